!
! mdk_device_m.F90 -- the device handle behind the drop-in modules.  The host
! arrays of particles_m/parrinellorahman_m/system_m stay authoritative between
! calls: every drop-in entry point pushes the host state it reads, runs the
! kernels through the C ABI and pulls back what the reference routine writes.
! (The MD loop can instead stay on the device: mdk_md_steps.)
!
module mdk_device_m
  use iso_c_binding
  use mdk_c
  use constants_m
  use params_m
  use particles_m
  use parrinellorahman_m
  use system_m
  implicit none
  private
  public :: mdk_dev_create, mdk_dev_destroy, mdk_dev_handle, mdk_dev_ready
  public :: mdk_dev_set_potential, mdk_dev_set_lookup, mdk_dev_push, mdk_dev_pull
  public :: mdk_dev_set_iron_carbon
  public :: mdk_dev_push_positions
  ! device options with no counterpart in the reference's simparameters, so
  ! the drop-in also builds against the reference's own params_m: set them
  ! before the first init_linkcell_m.  precision 0 FP64 / 1 FP32 form (an
  ! unchanged caller selects the FP32 form with MDK_PRECISION=fp32 in the
  ! environment); rebuild 0 the reference's ntcm cadence plus the skin
  ! monitor, 1 skin monitor only; rskin the list skin (<= 0: the reference's
  ! rpad)
  integer, public, save :: mdk_precision = 0, mdk_rebuild = 0
  real(c_double), public, save :: mdk_rskin = 0.0d0

  type(c_ptr), save :: handle = c_null_ptr
  logical, save :: pot_pending = .false.
  integer, save :: nv_p = 0, np_p = 0, nlookup_p = 0
  real(c_double), save :: av_p(8), rv_p(8), ap_p(4), rp_p(4)
  ! iron_carbon material: coefficients from an iron.in, or the built-in ones
  logical, save :: fec_p = .false., fec_file_p = .false.
  real(c_double), target, save :: fec_c_p(38)

contains

  logical function mdk_dev_ready()
    mdk_dev_ready = c_associated(handle)
  end function mdk_dev_ready

  type(c_ptr) function mdk_dev_handle()
    if (.not. c_associated(handle)) stop 'MDK: device handle not created (call init_linkcell_m)'
    mdk_dev_handle = handle
  end function mdk_dev_handle

  !> Creates the device state from the current parameters (the reference's
  !> link-cell grid must already be set, src/moldin.F90:252-257).
  subroutine mdk_dev_create()
    type(simparameters) :: sp
    type(mdk_params_t) :: p
    character(len=16) :: envv
    integer :: envlen, envst
    if (c_associated(handle)) return
    sp = get_params()
    p%nm = sp%nm; p%ivol = sp%ivol; p%nnbrs = sp%nnbrs; p%ntcm = sp%ntcm
    p%nlcx = sp%nlcx; p%nlcy = sp%nlcy; p%nlcz = sp%nlcz
    p%precision = mdk_precision; p%rebuild = mdk_rebuild; p%max_full = 0
    call get_environment_variable('MDK_PRECISION', envv, envlen, envst)
    if (envst == 0 .and. (trim(envv) == 'fp32' .or. trim(envv) == 'FP32' .or. trim(envv) == '1')) &
       p%precision = 1
    p%deltat = sp%deltat; p%rnear = sp%rnear; p%press = sp%press; p%bdel2 = sp%bdel2
    p%rskin = mdk_rskin
    if (p%rskin <= 0.0d0) p%rskin = sp%rpad
    call mdk_check(mdk_create(handle, p), 'mdk_create')
    if (pot_pending) call apply_potential()
  end subroutine mdk_dev_create

  subroutine mdk_dev_destroy()
    integer(c_int) :: rc
    if (c_associated(handle)) rc = mdk_destroy(handle)
    handle = c_null_ptr
  end subroutine mdk_dev_destroy

  subroutine apply_potential()
    if (fec_p) then
       if (fec_file_p) then
          call mdk_check(mdk_set_iron_carbon(handle, c_loc(fec_c_p)), 'mdk_set_iron_carbon')
       else
          call mdk_check(mdk_set_iron_carbon(handle, c_null_ptr), 'mdk_set_iron_carbon')
       end if
    else
       call mdk_check(mdk_set_potential(handle, int(nv_p, c_int), av_p, rv_p, int(np_p, c_int), ap_p, rp_p), &
                      'mdk_set_potential')
    end if
    call mdk_check(mdk_set_lookup(handle, int(nlookup_p, c_int)), 'mdk_set_lookup')
  end subroutine apply_potential

  subroutine mdk_dev_set_potential(nv, av, rv, np, ap, rp)
    integer, intent(in) :: nv, np
    real(kind_wp), intent(in) :: av(nv), rv(nv), ap(np), rp(np)
    nv_p = nv; np_p = np
    av_p(1:nv) = av; rv_p(1:nv) = rv; ap_p(1:np) = ap; rp_p(1:np) = rp
    pot_pending = .true.
    if (c_associated(handle)) call apply_potential()
  end subroutine mdk_dev_set_potential

  !> iron_carbon material; c = the 38 values of an iron.in when from_file
  subroutine mdk_dev_set_iron_carbon(c, from_file)
    real(kind_wp), intent(in) :: c(38)
    logical, intent(in) :: from_file
    fec_p = .true.
    fec_file_p = from_file
    fec_c_p = c
    pot_pending = .true.
    if (c_associated(handle)) call apply_potential()
  end subroutine mdk_dev_set_iron_carbon

  subroutine mdk_dev_set_lookup(nlookup)
    integer, intent(in) :: nlookup
    nlookup_p = nlookup
    if (c_associated(handle) .and. pot_pending) &
         call mdk_check(mdk_set_lookup(handle, int(nlookup_p, c_int)), 'mdk_set_lookup')
  end subroutine mdk_dev_set_lookup

  subroutine mdk_dev_push_positions()
    call mdk_check(mdk_upload(handle, MDK_X0, x0), 'upload x0')
    call mdk_check(mdk_upload(handle, MDK_Y0, y0), 'upload y0')
    call mdk_check(mdk_upload(handle, MDK_Z0, z0), 'upload z0')
    call mdk_check(mdk_set_tensor(handle, MDK_B0, b0), 'set b0')
  end subroutine mdk_dev_push_positions

  !> host -> device: every array the hot path reads
  subroutine mdk_dev_push()
    type(simparameters) :: sp
    integer(c_int), allocatable :: z(:)
    call mdk_dev_push_positions()
    call mdk_check(mdk_upload(handle, MDK_X1, x1), 'upload x1')
    call mdk_check(mdk_upload(handle, MDK_Y1, y1), 'upload y1')
    call mdk_check(mdk_upload(handle, MDK_Z1, z1), 'upload z1')
    call mdk_check(mdk_upload(handle, MDK_X2, x2), 'upload x2')
    call mdk_check(mdk_upload(handle, MDK_Y2, y2), 'upload y2')
    call mdk_check(mdk_upload(handle, MDK_Z2, z2), 'upload z2')
    call mdk_check(mdk_upload(handle, MDK_X3, x3), 'upload x3')
    call mdk_check(mdk_upload(handle, MDK_Y3, y3), 'upload y3')
    call mdk_check(mdk_upload(handle, MDK_Z3, z3), 'upload z3')
    call mdk_check(mdk_upload(handle, MDK_MASS, atomic_mass), 'upload mass')
    call mdk_check(mdk_upload(handle, MDK_DEL2, del2), 'upload del2')
    call mdk_check(mdk_upload(handle, MDK_AFRHO, afrho), 'upload afrho')
    call mdk_check(mdk_upload(handle, MDK_DAFRHO, dafrho), 'upload dafrho')
    allocate(z(size(atomic_number)))
    z = atomic_number
    call mdk_check(mdk_upload_species(handle, z), 'upload species')
    call mdk_check(mdk_set_tensor(handle, MDK_B1, b1), 'set b1')
    call mdk_check(mdk_set_tensor(handle, MDK_B2, b2), 'set b2')
    call mdk_check(mdk_set_tensor(handle, MDK_B3, b3), 'set b3')
    call mdk_check(mdk_set_tensor(handle, MDK_TGID, tgid), 'set tgid')
    sp = get_params()
    call mdk_check(mdk_set_ntc(handle, int(sp%ntc, c_int)), 'set ntc')
  end subroutine mdk_dev_push

  !> device -> host: state written by force / velocityverlet
  subroutine mdk_dev_pull(with_positions)
    logical, intent(in) :: with_positions
    type(simparameters) :: sp
    real(c_double) :: sc(MDK_NSCALARS)
    integer(c_int) :: ntc
    if (with_positions) then
       call mdk_check(mdk_download(handle, MDK_X0, x0), 'download x0')
       call mdk_check(mdk_download(handle, MDK_Y0, y0), 'download y0')
       call mdk_check(mdk_download(handle, MDK_Z0, z0), 'download z0')
       call mdk_check(mdk_download(handle, MDK_X1, x1), 'download x1')
       call mdk_check(mdk_download(handle, MDK_Y1, y1), 'download y1')
       call mdk_check(mdk_download(handle, MDK_Z1, z1), 'download z1')
       call mdk_check(mdk_download(handle, MDK_X2, x2), 'download x2')
       call mdk_check(mdk_download(handle, MDK_Y2, y2), 'download y2')
       call mdk_check(mdk_download(handle, MDK_Z2, z2), 'download z2')
       call mdk_check(mdk_download(handle, MDK_X3, x3), 'download x3')
       call mdk_check(mdk_download(handle, MDK_Y3, y3), 'download y3')
       call mdk_check(mdk_download(handle, MDK_Z3, z3), 'download z3')
       call mdk_check(mdk_get_tensor(handle, MDK_B0, b0), 'get b0')
       call mdk_check(mdk_get_tensor(handle, MDK_B1, b1), 'get b1')
       call mdk_check(mdk_get_tensor(handle, MDK_B2, b2), 'get b2')
       call mdk_check(mdk_get_tensor(handle, MDK_B3, b3), 'get b3')
    end if
    call mdk_check(mdk_download(handle, MDK_RHO, rho), 'download rho')
    call mdk_check(mdk_download(handle, MDK_AFRHO, afrho), 'download afrho')
    call mdk_check(mdk_download(handle, MDK_DAFRHO, dafrho), 'download dafrho')
    call mdk_check(mdk_get_tensor(handle, MDK_TP, tp), 'get tp')
    call mdk_check(mdk_get_tensor(handle, MDK_TK, tk), 'get tk')
    call mdk_check(mdk_get_tensor(handle, MDK_TC, tc), 'get tc')
    call mdk_check(mdk_get_tensor(handle, MDK_FB, fb), 'get fb')
    call mdk_check(mdk_get_tensor(handle, MDK_TG, tg), 'get tg')
    call mdk_check(mdk_get_tensor(handle, MDK_TGDOT, tgdot), 'get tgdot')
    call mdk_check(mdk_get_scalars(handle, sc), 'get scalars')
    pe = sc(MDK_S_PE)
    vol = sc(MDK_S_VOL)
    call mdk_check(mdk_get_ntc(handle, ntc), 'get ntc')
    sp = get_params()
    sp%ntc = ntc
    call set_params(sp)
  end subroutine mdk_dev_pull
end module mdk_device_m
