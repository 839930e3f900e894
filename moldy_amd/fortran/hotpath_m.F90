!
! hotpath_m.F90 -- drop-in replacements of the reference's hot-path modules
! with the same names and argument lists, running on the GPU through the C ABI:
!   potential_m      src/potential_m.F90:100-118   (fetransmetal Fe-Fe, src/fetransmetal.F90;
!                    with -DIRON_CARBON the iron_carbon Fe-Fe functions, src/iron_carbon.F90)
!   linkcell_m       src/linkcell_m.F90:52-54
!   neighbourlist_m  src/neighbourlist_m.F90:52-58
!   thermostat_m     src/thermostat_m.F90:94-395   (init, kinten, scalev, snhv)
!   dynamics_m       src/dynamics_m.F90:69-83      (force, rhoset, energy_calc, update_therm_avs)
!   quench_m         src/quench_m.F90:59-63        (velocityverlet, predic, correc)
!
module potential_m
  use iso_c_binding
  use mdk_c
  use constants_m
  use params_m
  use particles_m
  use system_m
  use mdk_device_m
  implicit none
  private
  public :: vee, dvee, vee_src, phi, dphi, emb, demb, embed, set_up, potin
  public :: potential_set_exact, potential_set_lookup, cleanup_potential_lookups
  public :: get_available_potential_range, get_potential

  integer, save :: nv = 0, np = 0
  real(kind_wp), save :: av(8), rv(8), ap(4), rp(4)
  logical, save :: have_coeffs = .false.

contains

#ifdef IRON_CARBON
  !> check_supported_atomic_numbers of iron_carbon (src/iron_carbon.F90:745-803):
  !> range pot_rmax = 5.4 A, or vcut1 when an "iron.in" supplies the 38
  !> coefficients.  Only Fe (26) is supported here: the device kernels
  !> evaluate the Fe-Fe functions (ierror = 1 for other atomic numbers).
  subroutine get_potential(spnum, spna, range, ierror)
    integer, intent(in) :: spnum
    integer, intent(in) :: spna(spnum)
    real(kind_wp), intent(out) :: range
    integer, intent(out) :: ierror
    integer :: u, ios, i
    character(len=30) :: titlefe
    real(kind_wp) :: c(38)
    ierror = 0
    range = 5.4_kind_wp
    do i = 1, spnum
       if (spna(i) /= 26) ierror = 1
    end do
    c = 0.0_kind_wp
    open(newunit=u, file="iron.in", status='old', action='read', iostat=ios)
    if (ios == 0) then
       read(u, *, iostat=ios) titlefe
       if (ios == 0) read(u, *, iostat=ios) c(1:7)
       if (ios == 0) read(u, *, iostat=ios) c(8:13)
       if (ios == 0) read(u, *, iostat=ios) c(14:20)
       if (ios == 0) read(u, *, iostat=ios) c(21:26)
       if (ios == 0) read(u, *, iostat=ios) c(27:30)
       if (ios == 0) read(u, *, iostat=ios) c(31:33)
       if (ios == 0) read(u, *, iostat=ios) c(34:36)
       if (ios == 0) read(u, *, iostat=ios) c(37:38)
       close(u)
       if (ios /= 0) then
          ierror = 2
          return
       end if
       range = c(1)
       call mdk_dev_set_iron_carbon(c, .true.)
    else
       call mdk_dev_set_iron_carbon(c, .false.)
    end if
    have_coeffs = .true.
  end subroutine get_potential

  subroutine get_available_potential_range(rmin, rmax)
    real(kind_wp), intent(out) :: rmin, rmax
    rmin = 0.
    rmax = 5.4
  end subroutine get_available_potential_range

  !> iron_carbon has no set-up step (src/potential_m.F90:257-261)
  subroutine set_up
  end subroutine set_up
#else
  !> Reads pot_026_026.in (AVTF layout, src/fetransmetal.F90:603-766) for the
  !> Fe-Fe pair; range is the potential's declared 6 A (:589-593).
  subroutine get_potential(spnum, spna, range, ierror)
    integer, intent(in) :: spnum
    integer, intent(in) :: spna(spnum)
    real(kind_wp), intent(out) :: range
    integer, intent(out) :: ierror
    integer :: u, za, zb
    character(len=100) :: title
    real(kind_wp) :: rmin_, rmax_
    range = 6.0
    open(newunit=u, file="pot_026_026.in", status='old', action='read', iostat=ierror)
    if (ierror /= 0) return
    read(u, '(a100)', iostat=ierror) title
    read(u, *, iostat=ierror) za
    read(u, *, iostat=ierror) zb
    read(u, *, iostat=ierror) nv
    read(u, *, iostat=ierror) np
    read(u, *, iostat=ierror) rmin_
    read(u, *, iostat=ierror) rmax_
    if (ierror == 0 .and. (nv > 8 .or. np > 4 .or. nv < 1 .or. np < 1)) ierror = 3
    if (ierror == 0) read(u, *, iostat=ierror) av(1:nv)
    if (ierror == 0) read(u, *, iostat=ierror) rv(1:nv)
    if (ierror == 0) read(u, *, iostat=ierror) ap(1:np)
    if (ierror == 0) read(u, *, iostat=ierror) rp(1:np)
    close(u)
    have_coeffs = (ierror == 0)
  end subroutine get_potential

  subroutine get_available_potential_range(rmin, rmax)
    real(kind_wp), intent(out) :: rmin, rmax
    rmin = 0.
    rmax = 6.
  end subroutine get_available_potential_range

  !> set_up_m: join constants and flat-phi cap, derived on the device side's host code
  subroutine set_up
    if (.not. have_coeffs) stop 'potential_m: get_potential must run before set_up'
    call mdk_dev_set_potential(nv, av(1:nv), rv(1:nv), np, ap(1:np), rp(1:np))
  end subroutine set_up
#endif

  subroutine potential_set_lookup(nlookup)
    integer, optional :: nlookup
    if (present(nlookup)) then
       call mdk_dev_set_lookup(nlookup)
    else
       call mdk_dev_set_lookup(5000)
    end if
  end subroutine potential_set_lookup

  subroutine potential_set_exact()
    call mdk_dev_set_lookup(0)
  end subroutine potential_set_exact

  subroutine cleanup_potential_lookups()
    call mdk_dev_set_lookup(0)
  end subroutine cleanup_potential_lookups

  real(kind_wp) function eval1(which, x)
    integer(c_int), intent(in) :: which
    real(kind_wp), intent(in) :: x
    real(c_double) :: xin(1), yout(1)
    xin(1) = x
    call mdk_check(mdk_eval_potential(mdk_dev_handle(), which, xin, yout, 1_c_int), 'mdk_eval_potential')
    eval1 = yout(1)
  end function eval1

  ! Per-pair functions stay host-callable (set-up, checks); the force loops
  ! never cross the ABI per pair.
  real(kind_wp) function vee(r, na1, na2)
    real(kind_wp), intent(in) :: r
    integer, intent(in) :: na1, na2
    vee = eval1(0_c_int, r)
  end function vee
  real(kind_wp) function vee_src(r, na1, na2)
    real(kind_wp), intent(in) :: r
    integer, intent(in) :: na1, na2
    vee_src = eval1(0_c_int, r)
  end function vee_src
  real(kind_wp) function dvee(r, na1, na2)
    real(kind_wp), intent(in) :: r
    integer, intent(in) :: na1, na2
    dvee = eval1(1_c_int, r)
  end function dvee
  real(kind_wp) function phi(r, na1, na2)
    real(kind_wp), intent(in) :: r
    integer, intent(in) :: na1, na2
    phi = eval1(2_c_int, r)
  end function phi
  real(kind_wp) function dphi(r, na1, na2)
    real(kind_wp), intent(in) :: r
    integer, intent(in) :: na1, na2
    dphi = eval1(3_c_int, r)
  end function dphi
  real(kind_wp) function emb(rho_, na1)
    real(kind_wp), intent(in) :: rho_
    integer, intent(in) :: na1
    emb = eval1(4_c_int, rho_)
  end function emb
  real(kind_wp) function demb(rho_, na1)
    real(kind_wp), intent(in) :: rho_
    integer, intent(in) :: na1
    demb = eval1(5_c_int, rho_)
  end function demb

  !> embed: F(rho), F'(rho) and the embedding energy (src/potential_m.F90:148-169),
  !> with the densities recomputed on the device from the current positions
  subroutine embed(pexx)
    real(kind_wp) :: pexx
    real(c_double) :: pe_c
    call mdk_dev_push()
    call mdk_check(mdk_embed(mdk_dev_handle(), pe_c), 'mdk_embed')
    pexx = pe_c
    call mdk_check(mdk_download(mdk_dev_handle(), MDK_RHO, rho), 'download rho')
    call mdk_check(mdk_download(mdk_dev_handle(), MDK_AFRHO, afrho), 'download afrho')
    call mdk_check(mdk_download(mdk_dev_handle(), MDK_DAFRHO, dafrho), 'download dafrho')
  end subroutine embed

  !> potin writes potential.dat for plotting; not part of the device path
  subroutine potin
  end subroutine potin
end module potential_m

module linkcell_m
  use iso_c_binding
  use mdk_c
  use params_m
  use particles_m
  use mdk_device_m
  implicit none
  private
  public :: init_linkcell_m, cleanup_linkcell_m, get_l0, get_link, linkup, relink, ic
  integer, allocatable :: ic(:)
  type(simparameters), save :: simparam
contains
  !> also creates the device state: the link-cell grid is the last parameter
  !> the reference fixes before its first force call
  subroutine init_linkcell_m
    simparam = get_params()
    allocate(ic(simparam%nm))
    ic = 0
    call mdk_dev_create()
  end subroutine init_linkcell_m
  subroutine cleanup_linkcell_m
    deallocate(ic)
  end subroutine cleanup_linkcell_m

  subroutine linkup
    integer(c_int), allocatable :: l0(:), lk(:), icc(:)
    call mdk_dev_push_positions()
    call mdk_check(mdk_linkup(mdk_dev_handle()), 'mdk_linkup')
    allocate(l0(simparam%nlcx*simparam%nlcy*simparam%nlcz), lk(simparam%nm), icc(simparam%nm))
    call mdk_check(mdk_get_linkcells(mdk_dev_handle(), l0, lk, icc), 'mdk_get_linkcells')
    ic = icc
  end subroutine linkup

  subroutine relink
    integer(c_int), allocatable :: l0(:), lk(:), icc(:)
    call mdk_dev_push_positions()
    call mdk_check(mdk_relink(mdk_dev_handle()), 'mdk_relink')
    allocate(l0(simparam%nlcx*simparam%nlcy*simparam%nlcz), lk(simparam%nm), icc(simparam%nm))
    call mdk_check(mdk_get_linkcells(mdk_dev_handle(), l0, lk, icc), 'mdk_get_linkcells')
    ic = icc
  end subroutine relink

  function get_l0()
    integer :: get_l0(simparam%nlcx*simparam%nlcy*simparam%nlcz)
    integer(c_int), allocatable :: l0(:), lk(:), icc(:)
    allocate(l0(size(get_l0)), lk(simparam%nm), icc(simparam%nm))
    call mdk_check(mdk_get_linkcells(mdk_dev_handle(), l0, lk, icc), 'mdk_get_linkcells')
    get_l0 = l0
  end function get_l0

  function get_link()
    integer :: get_link(simparam%nm)
    integer(c_int), allocatable :: l0(:), lk(:), icc(:)
    allocate(l0(simparam%nlcx*simparam%nlcy*simparam%nlcz), lk(simparam%nm), icc(simparam%nm))
    call mdk_check(mdk_get_linkcells(mdk_dev_handle(), l0, lk, icc), 'mdk_get_linkcells')
    get_link = lk
  end function get_link
end module linkcell_m

module neighbourlist_m
  use iso_c_binding
  use mdk_c
  use params_m
  use particles_m
  use mdk_device_m
  implicit none
  private
  public :: init_neighbourlist_m, cleanup_neighbourlist_m, update_neighbourlist
  integer, allocatable, public :: nlist(:,:), numn(:)
contains
  subroutine init_neighbourlist_m()
    type(simparameters) :: sp
    sp = get_params()
    allocate(nlist(sp%nnbrs, sp%nm), numn(sp%nm))
    numn = 0
  end subroutine init_neighbourlist_m
  subroutine cleanup_neighbourlist_m()
    deallocate(nlist, numn)
  end subroutine cleanup_neighbourlist_m

  !> the reference half list at rnear, built on the device and downloaded
  subroutine update_neighbourlist
    integer(c_int) :: rc
    integer(c_int), allocatable :: nn(:), nl(:)
    call mdk_dev_push_positions()
    rc = mdk_update_neighbourlist(mdk_dev_handle())
    if (rc == MDK_ERR_NBR_OVERFLOW) then
       write(0, "(' ATTEMPT TO LIST MORE THAN ',i5,' NEIGHBOURS ')") size(nlist, 1)
       stop 'NEIGHBOURLIST: EXCESS NEIGHBOURS!'
    end if
    call mdk_check(rc, 'mdk_update_neighbourlist')
    allocate(nn(size(numn)), nl(size(nlist)))
    call mdk_check(mdk_get_neighbourlist(mdk_dev_handle(), nn, nl), 'mdk_get_neighbourlist')
    numn = nn
    nlist = reshape(nl, shape(nlist))
  end subroutine update_neighbourlist
end module neighbourlist_m

module thermostat_m
  use iso_c_binding
  use mdk_c
  use constants_m
  use params_m
  use particles_m
  use parrinellorahman_m
  use mdk_device_m
  implicit none
  private
  public :: init_thermostat_m, cleanup_thermostat_m, kinten, scalev, set_snhv, get_snhv
contains
  !> scalev (src/thermostat_m.F90:317-395): Nose-Hoover velocity scaling on
  !> the device; snhv lives with the device handle
  subroutine scalev
    type(simparameters) :: sp
    sp = get_params()
    if (.not. (sp%nose > 0.0d0)) return
    call mdk_dev_push()
    call mdk_check(mdk_scalev(mdk_dev_handle(), real(sp%nose, c_double), real(sp%rqke, c_double)), 'mdk_scalev')
    call mdk_dev_pull(.true.)
  end subroutine scalev
  subroutine set_snhv(v)
    real(kind_wp), intent(in) :: v
    call mdk_check(mdk_set_snhv(mdk_dev_handle(), real(v, c_double)), 'mdk_set_snhv')
  end subroutine set_snhv
  real(kind_wp) function get_snhv()
    real(c_double) :: v
    call mdk_check(mdk_get_snhv(mdk_dev_handle(), v), 'mdk_get_snhv')
    get_snhv = v
  end function get_snhv

  subroutine init_thermostat_m
    type(simparameters) :: sp
    sp = get_params()
    sp%rqke = 1.5d0*bk*sp%temprq*(sp%nm-1)
    call set_params(sp)
  end subroutine init_thermostat_m
  subroutine cleanup_thermostat_m
  end subroutine cleanup_thermostat_m

  !> kinetic box stress tk and tgdot (src/thermostat_m.F90:127-177) on the host;
  !> inside force the device computes the same sums
  subroutine kinten
    integer :: i, j, k
    real(kind_wp) :: mvx, mvy, mvz
    type(simparameters) :: sp
    sp = get_params()
    tk = 0.0d0
    do i = 1, sp%nm
       if (atomic_number(i) == 0) cycle
       mvx = (b0(1,1)*x1(i)+b0(1,2)*y1(i)+b0(1,3)*z1(i))*atomic_mass(i)
       mvy = (b0(2,1)*x1(i)+b0(2,2)*y1(i)+b0(2,3)*z1(i))*atomic_mass(i)
       mvz = (b0(3,1)*x1(i)+b0(3,2)*y1(i)+b0(3,3)*z1(i))*atomic_mass(i)
       tk(1,1) = tk(1,1)+x1(i)*mvx; tk(2,1) = tk(2,1)+x1(i)*mvy; tk(3,1) = tk(3,1)+x1(i)*mvz
       tk(1,2) = tk(1,2)+y1(i)*mvx; tk(2,2) = tk(2,2)+y1(i)*mvy; tk(3,2) = tk(3,2)+y1(i)*mvz
       tk(1,3) = tk(1,3)+z1(i)*mvx; tk(2,3) = tk(2,3)+z1(i)*mvy; tk(3,3) = tk(3,3)+z1(i)*mvz
    end do
    tk = tk/(sp%deltat*sp%deltat)
    tgdot = 0.0d0
    do i = 1, nmat
       do k = 1, nmat
          do j = 1, nmat
             tgdot(i,j) = tgdot(i,j)+b0(k,i)*b1(k,j)+b1(k,i)*b0(k,j)
          end do
       end do
    end do
  end subroutine kinten
end module thermostat_m

module dynamics_m
  use iso_c_binding
  use mdk_c
  use constants_m
  use params_m
  use particles_m
  use potential_m
  use system_m
  use thermostat_m
  use parrinellorahman_m
  use neighbourlist_m
  use mdk_device_m
  implicit none
  private
  public :: init_dynamics_m, cleanup_dynamics_m, force, rhoset, update_therm_avs, energy_calc
  public :: update_neighbourlist, write_energy_forces_stress, dynamics_pull_forces
  real(kind_wp), allocatable, public :: fx(:), fy(:), fz(:)
  real(kind_wp), allocatable, public :: x1_dt2(:), y1_dt2(:), z1_dt2(:)
contains
  subroutine init_dynamics_m
    type(simparameters) :: sp
    sp = get_params()
    allocate(fx(sp%nm), fy(sp%nm), fz(sp%nm), x1_dt2(sp%nm), y1_dt2(sp%nm), z1_dt2(sp%nm))
    fx = 0; fy = 0; fz = 0; x1_dt2 = 0; y1_dt2 = 0; z1_dt2 = 0
  end subroutine init_dynamics_m
  subroutine cleanup_dynamics_m
    deallocate(fx, fy, fz, x1_dt2, y1_dt2, z1_dt2)
  end subroutine cleanup_dynamics_m

  subroutine dynamics_pull_forces()
    type(c_ptr) :: h
    h = mdk_dev_handle()
    call mdk_check(mdk_download(h, MDK_FX, fx), 'download fx')
    call mdk_check(mdk_download(h, MDK_FY, fy), 'download fy')
    call mdk_check(mdk_download(h, MDK_FZ, fz), 'download fz')
  end subroutine dynamics_pull_forces

  !> force (src/dynamics_m.F90:135-401): list maintenance (ntc/ntcm), density,
  !> embedding, pair forces, virial tp, tc, kinten and fb on the device
  subroutine force
    type(simparameters) :: sp
    logical :: list_due
    sp = get_params()
    list_due = (sp%ntc + 1 == 1)   ! the reference rebuilds NLIST/NUMN inside force here
    call mdk_dev_push()
    call mdk_check(mdk_force(mdk_dev_handle()), 'mdk_force')
    call mdk_dev_pull(.false.)
    call dynamics_pull_forces()
    if (list_due) call update_neighbourlist
  end subroutine force

  !> rhoset (src/dynamics_m.F90:497-579).  The density function is the
  !> registered potential's phi; phi_local is accepted for interface parity.
  subroutine rhoset(phi_local, rho_local)
    real(kind_wp), external :: phi_local
    real(kind_wp) :: rho_local(:)
    call mdk_dev_push()
    call mdk_check(mdk_rhoset(mdk_dev_handle()), 'mdk_rhoset')
    call mdk_check(mdk_download(mdk_dev_handle(), MDK_RHO, rho_local), 'download rho')
  end subroutine rhoset

  !> energy_calc (src/dynamics_m.F90:410-470): en_atom from the list and
  !> densities of the last force, with the exact pair function
  subroutine energy_calc
    call mdk_dev_push()
    call mdk_check(mdk_energy_calc(mdk_dev_handle(), en_atom), 'mdk_energy_calc')
  end subroutine energy_calc

  !> update_therm_avs (src/dynamics_m.F90:649-734)
  subroutine update_therm_avs(s1, s2, s3)
    real(kind_wp) :: s1, s2, s3
    real(c_double) :: a, b, c
    call mdk_dev_push()
    call mdk_check(mdk_update_therm_avs(mdk_dev_handle(), a, b, c), 'mdk_update_therm_avs')
    s1 = a; s2 = b; s3 = c
    call mdk_check(mdk_get_tensor(mdk_dev_handle(), MDK_TG, tg), 'get tg')
    call mdk_check(mdk_get_tensor(mdk_dev_handle(), MDK_TGDOT, tgdot), 'get tgdot')
    call mdk_check(mdk_get_tensor(mdk_dev_handle(), MDK_TGINV, tginv), 'get tginv')
    call mdk_check(mdk_get_tensor(mdk_dev_handle(), MDK_TGID, tgid), 'get tgid')
    call pr_get_metric_tensor
  end subroutine update_therm_avs

  !> efs.out: energy, Cartesian forces B0.f and stress (src/dynamics_m.F90:588-619)
  subroutine write_energy_forces_stress()
    integer :: u, i, j, k
    real(kind_wp) :: sigma(3,3)
    type(simparameters) :: sp
    sp = get_params()
    open(newunit=u, file="efs.out", status='unknown')
    write(u,*) "energy", pe
    do i = 1, sp%nm
       write(u,*) b0(1,1)*fx(i)+b0(1,2)*fy(i)+b0(1,3)*fz(i), &
                  b0(2,1)*fx(i)+b0(2,2)*fy(i)+b0(2,3)*fz(i), &
                  b0(3,1)*fx(i)+b0(3,2)*fy(i)+b0(3,3)*fz(i)
    end do
    do i = 1, 3
       do j = 1, 3
          sigma(i,j) = -1.d0*sum((/(b0(j,k)*tp(i,k), k = 1, 3)/))/vol
       end do
    end do
    write(u,*) sigma
    close(u)
  end subroutine write_energy_forces_stress
end module dynamics_m

module quench_m
  use iso_c_binding
  use mdk_c
  use constants_m
  use params_m
  use particles_m
  use parrinellorahman_m
  use dynamics_m
  use mdk_device_m
  implicit none
  private
  public :: velocityverlet, predic, correc
contains
  !> predic (src/quench_m.F90:636-671); a no-op until correc or
  !> velocityverlet has run, as the reference's (see mdk_predic)
  subroutine predic
    call mdk_dev_push()
    call mdk_check(mdk_predic(mdk_dev_handle()), 'mdk_predic')
    call mdk_dev_pull(.true.)
  end subroutine predic

  !> correc (src/quench_m.F90:682-759) with the forces of the last force call
  subroutine correc
    call mdk_dev_push()
    call mdk_check(mdk_correc(mdk_dev_handle()), 'mdk_correc')
    call mdk_dev_pull(.true.)
  end subroutine correc

  !> velocityverlet(dt) (src/quench_m.F90:562-624), force included
  subroutine velocityverlet(dt)
    real(kind_wp) :: dt
    call mdk_dev_push()
    call mdk_check(mdk_velocityverlet(mdk_dev_handle()), 'mdk_velocityverlet')
    call mdk_dev_pull(.true.)
    call dynamics_pull_forces()
    call mdk_check(mdk_get_tensor(mdk_dev_handle(), MDK_TGID, tgid), 'get tgid')
  end subroutine velocityverlet
end module quench_m
