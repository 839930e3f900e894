!
! mdk_c.F90 -- iso_c_binding declarations of the C ABI in include/moldy_hip.h.
! Every interface binds one extern "C" entry point of libmoldy_hip.so; the
! enumerations repeat the header's values.
!
module mdk_c
  use iso_c_binding
  implicit none
  public

  integer(c_int), parameter :: MDK_OK = 0, MDK_ERR_HIP = 1, MDK_ERR_NBR_OVERFLOW = 2
  integer(c_int), parameter :: MDK_ERR_BAD_ARG = 3, MDK_ERR_SINGULAR_BOX = 4, MDK_ERR_STATE = 5
  integer(c_int), parameter :: MDK_PREC_FP64 = 0, MDK_PREC_FP32 = 1
  ! per-atom fields
  integer(c_int), parameter :: MDK_X0 = 0, MDK_Y0 = 1, MDK_Z0 = 2, MDK_X1 = 3, MDK_Y1 = 4, MDK_Z1 = 5
  integer(c_int), parameter :: MDK_X2 = 6, MDK_Y2 = 7, MDK_Z2 = 8, MDK_FX = 9, MDK_FY = 10, MDK_FZ = 11
  integer(c_int), parameter :: MDK_RHO = 12, MDK_AFRHO = 13, MDK_DAFRHO = 14, MDK_MASS = 15, MDK_DEL2 = 16
  integer(c_int), parameter :: MDK_X3 = 17, MDK_Y3 = 18, MDK_Z3 = 19
  ! tensors
  integer(c_int), parameter :: MDK_B0 = 0, MDK_B1 = 1, MDK_B2 = 2, MDK_TK = 3, MDK_TC = 4, MDK_TG = 5
  integer(c_int), parameter :: MDK_TP = 6, MDK_TGINV = 7, MDK_TGDOT = 8, MDK_TGID = 9, MDK_FB = 10
  integer(c_int), parameter :: MDK_B3 = 11
  ! scalars
  integer, parameter :: MDK_S_PE = 1, MDK_S_KE = 2, MDK_S_VOL = 3, MDK_S_S1 = 4, MDK_S_S2 = 5
  integer, parameter :: MDK_S_S3 = 6, MDK_NSCALARS = 11

  type, bind(C) :: mdk_params_t
     integer(c_int) :: nm, ivol, nnbrs, ntcm, nlcx, nlcy, nlcz, precision, rebuild, max_full
     real(c_double) :: deltat, rnear, rskin, press, bdel2
  end type mdk_params_t

  interface
     integer(c_int) function mdk_create(h, p) bind(C, name="mdk_create")
       import :: c_int, c_ptr, mdk_params_t
       type(c_ptr), intent(out) :: h
       type(mdk_params_t), intent(in) :: p
     end function mdk_create
     integer(c_int) function mdk_destroy(h) bind(C, name="mdk_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function mdk_destroy
     type(c_ptr) function mdk_last_error() bind(C, name="mdk_last_error")
       import :: c_ptr
     end function mdk_last_error
     integer(c_int) function mdk_set_potential(h, nv, av, rv, np, ap, rp) bind(C, name="mdk_set_potential")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int), value :: nv, np
       real(c_double), intent(in) :: av(*), rv(*), ap(*), rp(*)
     end function mdk_set_potential
     integer(c_int) function mdk_set_iron_carbon(h, iron_in) bind(C, name="mdk_set_iron_carbon")
       import :: c_int, c_ptr
       type(c_ptr), value :: h, iron_in
     end function mdk_set_iron_carbon
     integer(c_int) function mdk_scalev(h, nose, rqke) bind(C, name="mdk_scalev")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), value :: nose, rqke
     end function mdk_scalev
     integer(c_int) function mdk_get_snhv(h, v) bind(C, name="mdk_get_snhv")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double) :: v
     end function mdk_get_snhv
     integer(c_int) function mdk_set_snhv(h, v) bind(C, name="mdk_set_snhv")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), value :: v
     end function mdk_set_snhv
     integer(c_int) function mdk_predic(h) bind(C, name="mdk_predic")
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function mdk_predic
     integer(c_int) function mdk_correc(h) bind(C, name="mdk_correc")
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function mdk_correc
     integer(c_int) function mdk_energy_calc(h, en_atom) bind(C, name="mdk_energy_calc")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double) :: en_atom(*)
     end function mdk_energy_calc
     integer(c_int) function mdk_atomic_stress(h, sx, sy, sz) bind(C, name="mdk_atomic_stress")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double) :: sx(*), sy(*), sz(*)
     end function mdk_atomic_stress
     integer(c_int) function mdk_set_lookup(h, nlookup) bind(C, name="mdk_set_lookup")
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), value :: nlookup
     end function mdk_set_lookup
     integer(c_int) function mdk_eval_potential(h, which, x, y, count) bind(C, name="mdk_eval_potential")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int), value :: which, count
       real(c_double), intent(in) :: x(*)
       real(c_double), intent(out) :: y(*)
     end function mdk_eval_potential
     integer(c_int) function mdk_upload(h, field, data) bind(C, name="mdk_upload")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int), value :: field
       real(c_double), intent(in) :: data(*)
     end function mdk_upload
     integer(c_int) function mdk_download(h, field, data) bind(C, name="mdk_download")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int), value :: field
       real(c_double), intent(out) :: data(*)
     end function mdk_download
     integer(c_int) function mdk_upload_species(h, z) bind(C, name="mdk_upload_species")
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), intent(in) :: z(*)
     end function mdk_upload_species
     integer(c_int) function mdk_set_tensor(h, which, t) bind(C, name="mdk_set_tensor")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int), value :: which
       real(c_double), intent(in) :: t(3,3)
     end function mdk_set_tensor
     integer(c_int) function mdk_get_tensor(h, which, t) bind(C, name="mdk_get_tensor")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int), value :: which
       real(c_double), intent(out) :: t(3,3)
     end function mdk_get_tensor
     integer(c_int) function mdk_get_scalars(h, out) bind(C, name="mdk_get_scalars")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: out(*)
     end function mdk_get_scalars
     integer(c_int) function mdk_set_ntc(h, ntc) bind(C, name="mdk_set_ntc")
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), value :: ntc
     end function mdk_set_ntc
     integer(c_int) function mdk_get_ntc(h, ntc) bind(C, name="mdk_get_ntc")
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), intent(out) :: ntc
     end function mdk_get_ntc
     integer(c_int) function mdk_linkup(h) bind(C, name="mdk_linkup")
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function mdk_linkup
     integer(c_int) function mdk_relink(h) bind(C, name="mdk_relink")
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function mdk_relink
     integer(c_int) function mdk_get_linkcells(h, l0, link, ic) bind(C, name="mdk_get_linkcells")
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), intent(out) :: l0(*), link(*), ic(*)
     end function mdk_get_linkcells
     integer(c_int) function mdk_update_neighbourlist(h) bind(C, name="mdk_update_neighbourlist")
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function mdk_update_neighbourlist
     integer(c_int) function mdk_get_neighbourlist(h, numn, nlist) bind(C, name="mdk_get_neighbourlist")
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), intent(out) :: numn(*), nlist(*)
     end function mdk_get_neighbourlist
     integer(c_int) function mdk_force(h) bind(C, name="mdk_force")
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function mdk_force
     integer(c_int) function mdk_rhoset(h) bind(C, name="mdk_rhoset")
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function mdk_rhoset
     integer(c_int) function mdk_embed(h, pe) bind(C, name="mdk_embed")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: pe
     end function mdk_embed
     integer(c_int) function mdk_velocityverlet(h) bind(C, name="mdk_velocityverlet")
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function mdk_velocityverlet
     integer(c_int) function mdk_update_therm_avs(h, s1, s2, s3) bind(C, name="mdk_update_therm_avs")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: s1, s2, s3
     end function mdk_update_therm_avs
     integer(c_int) function mdk_md_steps(h, nsteps, trace) bind(C, name="mdk_md_steps")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int), value :: nsteps
       real(c_double), intent(out) :: trace(*)
     end function mdk_md_steps
  end interface

contains

  !> Message of the last failing C-ABI call.
  function mdk_error_string() result(s)
    character(len=512) :: s
    character(kind=c_char), pointer :: cs(:)
    type(c_ptr) :: p
    integer :: k
    s = ' '
    p = mdk_last_error()
    if (.not. c_associated(p)) return
    call c_f_pointer(p, cs, [512])
    do k = 1, 512
       if (cs(k) == c_null_char) exit
       s(k:k) = cs(k)
    end do
  end function mdk_error_string

  !> Stops like the reference does on fatal errors, with the C-side message.
  subroutine mdk_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc == MDK_OK) return
    write(0,'(a,a,a,i0,a,a)') 'MDK: ', where, ' failed (', rc, '): ', trim(mdk_error_string())
    if (rc == MDK_ERR_NBR_OVERFLOW) stop 'NEIGHBOURLIST: EXCESS NEIGHBOURS!'
    if (rc == MDK_ERR_SINGULAR_BOX) stop "Error in matrix_invert: determinant is or is close to zero. STOP."
    stop 'MDK: device call failed'
  end subroutine mdk_check
end module mdk_c
