!
! hotpath_m.F90 -- drop-in replacements of the reference's hot-path modules
! with the same names and argument lists, running on the GPU through the C ABI:
!   potential_m      src/potential_m.F90:100-118   (fetransmetal Fe-Fe, src/fetransmetal.F90;
!                    with -DIRON_CARBON the iron_carbon Fe-Fe functions, src/iron_carbon.F90)
!   linkcell_m       src/linkcell_m.F90:52-54
!   neighbourlist_m  src/neighbourlist_m.F90:52-58
!   thermostat_m     src/thermostat_m.F90:52-395   (init, temp, setvel, kinten, scalev, snhv)
!   dynamics_m       src/dynamics_m.F90:69-83      (force with atomstress, rhoset, energy_calc,
!                    update_therm_avs, efs.out, write_stress, write_atomic_stress, energy_pair_emb)
!   quench_m         src/quench_m.F90:59-63        (velocityverlet, predic, correc, quench)
! They build against this package's state modules (state_m.F90) or, unchanged,
! against the reference's own constants_m, params_m, random_m, particles_m,
! system_m, matrix_m and parrinellorahman_m.
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

  !> embed (src/potential_m.F90:148-169): F(rho) and F'(rho) of the caller's
  !> RHO array, atoms with rho <= 1e-40 keeping their AFRHO and DAFRHO, and
  !> the embedding energy.  The embedding functions are evaluated by the
  !> device for all atoms in one call each (the exact functions of the
  !> density pass); the sum runs in atom order.
  subroutine embed(pexx)
    real(kind_wp) :: pexx
    real(c_double), allocatable :: f(:), df(:)
    integer :: i, n
    n = size(rho)
    allocate(f(n), df(n))
    call mdk_check(mdk_eval_potential(mdk_dev_handle(), 4_c_int, rho, f, int(n, c_int)), 'mdk_eval_potential')
    call mdk_check(mdk_eval_potential(mdk_dev_handle(), 5_c_int, rho, df, int(n, c_int)), 'mdk_eval_potential')
    pexx = 0.0_kind_wp
    do i = 1, n
       if (rho(i) <= 1.0e-40_kind_wp) cycle
       afrho(i) = f(i)
       dafrho(i) = df(i)
       pexx = pexx + afrho(i)
    end do
    deallocate(f, df)
  end subroutine embed

  !> The functions of one kind (0 V, 1 V', 2 phi, 3 phi', 4 F, 5 F') at many
  !> points in one device call.
  subroutine evalv(which, x, y)
    integer(c_int), intent(in) :: which
    real(kind_wp), intent(in) :: x(:)
    real(kind_wp), intent(out) :: y(:)
    call mdk_check(mdk_eval_potential(mdk_dev_handle(), which, x, y, int(size(x), c_int)), 'mdk_eval_potential')
  end subroutine evalv

  !> potin (src/potential_m.F90:429-642): potential.dat, for 1000 nearest-
  !> neighbour distances over the potential's range, the pair and density
  !> functions there and the cohesive energies of bcc, fcc and ideal hcp at
  !> the same volume per atom (shell radii and counts, the literals' kinds and
  !> the reference's choice of embedding term for hcp kept, so the file
  !> agrees with the reference's).  The device holds the Fe-Fe functions
  !> only: blocks are written for Fe-Fe species pairs.
  subroutine potin
    integer, parameter :: npts = 1000
    real(kind_wp), parameter :: rminplot = 1.5_kind_wp
    ! shell weights of the pair sums (half the shell counts) and of the
    ! densities: bcc, fcc and ideal hcp
    real(kind_wp), parameter :: nbv(6) = (/4.0, 3.0, 6.0, 12.0, 4.0, 3.0/)
    real(kind_wp), parameter :: nbp(6) = (/8.0, 6.0, 12.0, 24.0, 8.0, 6.0/)
    real(kind_wp), parameter :: nfv(5) = (/6.0, 3.0, 12.0, 6.0, 12.0/)
    real(kind_wp), parameter :: nfp(5) = (/12.0, 6.0, 24.0, 12.0, 24.0/)
    real(kind_wp), parameter :: nhv(7) = (/6.0, 3.0, 1.0, 9.0, 6.0, 3.0, 9.0/)
    real(kind_wp) :: rb(npts, 6), rf(npts, 5), rh(npts, 7)
    real(kind_wp) :: vb(npts, 6), pb(npts, 6), vf(npts, 5), pf(npts, 5), vh(npts, 7), ph(npts, 7)
    real(kind_wp) :: ccb(npts), ccf(npts), eb(npts), ef(npts), dph(npts), dvh(npts)
    real(kind_wp) :: rmin, rmax, ecr, ecbccb, ecfcc, echcp, veff
    type(simparameters) :: sp
    integer :: i, j, ix, k, u
    sp = get_params()
    call get_available_potential_range(rmin, rmax)
    rmin = max(rmin, rminplot)
    open(newunit=u, file="potential.dat", status='unknown')
    do i = 1, sp%nspec
       do j = i, sp%nspec
          if (ispec(i) /= 26 .or. ispec(j) /= 26) cycle
          do ix = 1, npts
             rb(ix, 1) = 2.0 + (ix*(rmax - rmin))/npts
             rb(ix, 2) = rb(ix, 1)*(2.0/sqrt(3.0d0))
             rb(ix, 3) = rb(ix, 2)*sqrt(2.0d0)
             rb(ix, 4) = rb(ix, 1)*sqrt(11.0/3.0)
             rb(ix, 5) = rb(ix, 1)*2.0
             rb(ix, 6) = rb(ix, 1)*sqrt(16.0/3.0)
             ! fcc at the same volume: second shell x2a, first x2a/sqrt(2)
             rf(ix, 2) = rb(ix, 2)*2d0**(0.3333333333)
             rf(ix, 1) = rf(ix, 2)/sqrt(2.0d0)
             rf(ix, 3) = rf(ix, 1)*sqrt(3.0d0)
             rf(ix, 4) = rf(ix, 1)*2d0
             rf(ix, 5) = rf(ix, 1)*sqrt(5d0)
             ! ideal hcp, the same x2a
             rh(ix, 1) = rf(ix, 1)
             rh(ix, 2) = rf(ix, 2)
             rh(ix, 3) = rh(ix, 1)*sqrt(8.0d0/3.0d0)
             rh(ix, 4) = rh(ix, 1)*sqrt(3.0d0)
             rh(ix, 5) = rh(ix, 1)*sqrt(11.0d0/3.0d0)
             rh(ix, 6) = rh(ix, 1)*2d0
             rh(ix, 7) = rh(ix, 1)*sqrt(5d0)
          end do
          do k = 1, 6
             call evalv(0_c_int, rb(:, k), vb(:, k))
             call evalv(2_c_int, rb(:, k), pb(:, k))
          end do
          do k = 1, 5
             call evalv(0_c_int, rf(:, k), vf(:, k))
             call evalv(2_c_int, rf(:, k), pf(:, k))
          end do
          do k = 1, 7
             call evalv(0_c_int, rh(:, k), vh(:, k))
             call evalv(2_c_int, rh(:, k), ph(:, k))
          end do
          do ix = 1, npts
             ccb(ix) = nbp(1)*pb(ix, 1) + nbp(2)*pb(ix, 2) + nbp(3)*pb(ix, 3) + nbp(4)*pb(ix, 4) &
                       + nbp(5)*pb(ix, 5) + nbp(6)*pb(ix, 6)
             ccf(ix) = nfp(1)*pf(ix, 1) + nfp(2)*pf(ix, 2) + nfp(3)*pf(ix, 3) + nfp(4)*pf(ix, 4) &
                       + nfp(5)*pf(ix, 5)
          end do
          call evalv(4_c_int, ccb, eb)
          call evalv(4_c_int, ccf, ef)
          call evalv(3_c_int, rh(:, 1), dph)
          call evalv(1_c_int, rh(:, 1), dvh)
          do ix = 1, npts
             ecr = nbv(1)*vb(ix, 1) + nbv(2)*vb(ix, 2) + nbv(3)*vb(ix, 3) + nbv(4)*vb(ix, 4) &
                   + nbv(5)*vb(ix, 5) + nbv(6)*vb(ix, 6)
             ecbccb = ecr + eb(ix)
             ecr = nfv(1)*vf(ix, 1) + nfv(2)*vf(ix, 2) + nfv(3)*vf(ix, 3) + nfv(4)*vf(ix, 4) &
                   + nfv(5)*vf(ix, 5)
             ecfcc = ecr + ef(ix)
             ecr = nhv(1)*vh(ix, 1) + nhv(2)*vh(ix, 2) + nhv(3)*vh(ix, 3) + nhv(4)*vh(ix, 4) &
                   + nhv(5)*vh(ix, 5) + nhv(6)*vh(ix, 6) + nhv(7)*vh(ix, 7)
             echcp = ecr + ef(ix)   ! the fcc embedding term, as the reference
             if (ph(ix, 1) > 1d-10) then
                veff = vh(ix, 1) + ph(ix, 1)/ef(ix)
                write(u, '(9e15.7)') rh(ix, 1), vh(ix, 1), ph(ix, 1), veff, dph(ix), dvh(ix), ecbccb, ecfcc, echcp
             end if
          end do
       end do
    end do
    close(u)
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
  use system_m, only : ke
  use random_m, only : rand1
  use mdk_device_m
  implicit none
  private
  public :: init_thermostat_m, cleanup_thermostat_m, kinten, scalev, set_snhv, get_snhv
  public :: set_temp, get_temp, setvel
  real(kind_wp), save :: temp = 0.0d0   ! the run's temperature (src/thermostat_m.F90:58,106-115)
contains
  !> setvel (src/thermostat_m.F90:191-306): initial velocities on the host
  !> (once per run).  Movable atoms draw x1, y1, z1 = rand1() - 1/2, the
  !> mean momentum per movable atom is removed, and all velocities are scaled
  !> so the kinetic energy is rqke nm/nmove; for ivol 1, 2, 3 and 5 the box
  !> velocities b1 are drawn after the atoms, rows centred by sum/nm and
  !> scaled to the box temperature temprq.  boxtem is set as the reference
  !> leaves it: from the sums of squares before and after the scaling.
  subroutine setvel
    type(simparameters) :: sp
    integer :: i, j, nmove
    real(kind_wp) :: px, py, pz, acc, vx, vy, vz, f, rsum
    sp = get_params()
    nmove = 0
    do i = 1, sp%nm
       if (atomic_number(i) == 0 .or. atomic_mass(i) < 0.1) cycle
       x1(i) = rand1() - 0.5
       y1(i) = rand1() - 0.5
       z1(i) = rand1() - 0.5
       nmove = nmove + 1
    end do
    write(*,*) sp%nm, "atoms, of which  ", nmove, "  are moveable"
    px = 0.0d0; py = 0.0d0; pz = 0.0d0
    do i = 1, sp%nm
       px = px + x1(i)*atomic_mass(i)
       py = py + y1(i)*atomic_mass(i)
       pz = pz + z1(i)*atomic_mass(i)
    end do
    px = px/nmove; py = py/nmove; pz = pz/nmove
    acc = 0.0d0
    do i = 1, sp%nm
       if (atomic_number(i) == 0 .or. atomic_mass(i) < 0.1) cycle
       x1(i) = x1(i) - px/atomic_mass(i)
       y1(i) = y1(i) - py/atomic_mass(i)
       z1(i) = z1(i) - pz/atomic_mass(i)
       vx = b0(1,1)*x1(i) + b0(1,2)*y1(i) + b0(1,3)*z1(i)
       vy = b0(2,1)*x1(i) + b0(2,2)*y1(i) + b0(2,3)*z1(i)
       vz = b0(3,1)*x1(i) + b0(3,2)*y1(i) + b0(3,3)*z1(i)
       acc = acc + atomic_mass(i)*(vx*vx + vy*vy + vz*vz)
    end do
    ke = acc/sp%deltat**2/2.d0
    ! the atom-count ratio and its root are single precision in the reference
    f = sqrt(sp%rqke/ke)*sqrt(real(nmove)/real(sp%nm))
    x1 = x1*f; y1 = y1*f; z1 = z1*f
    call set_temp(sp%rqke/(1.5d0*bk))
    ke = sp%rqke
    if (sp%ivol < 1 .or. sp%ivol == 4) return
    acc = 0.0d0
    do i = 1, nmat
       rsum = 0.0d0
       do j = 1, nmat
          b1(i,j) = rand1() - 0.5d0
          rsum = rsum + b1(i,j)
       end do
       rsum = rsum/sp%nm
       do j = 1, nmat
          b1(i,j) = b1(i,j) - rsum
          acc = acc + b1(i,j)**2
       end do
    end do
    sp%boxtem = (acc/sp%bdel2)/(4.5d0*bk)
    f = sqrt(sp%temprq/sp%boxtem)
    do i = 1, nmat
       do j = 1, nmat
          b1(i,j) = b1(i,j)*f
          acc = acc + b1(i,j)**2
       end do
    end do
    sp%boxtem = (acc/sp%bdel2)/(4.5d0*bk)
    call set_params(sp)
  end subroutine setvel
  subroutine set_temp(t)
    real(kind_wp), intent(in) :: t
    temp = t
  end subroutine set_temp
  pure real(kind_wp) function get_temp()
    get_temp = temp
  end function get_temp
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
  public :: write_stress, write_atomic_stress, energy_pair_emb, dynamics_after_force
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
    call dynamics_after_force(list_due)
  end subroutine force

  !> What the reference's force leaves besides forces, after a device force
  !> evaluation (also the one inside velocityverlet): NLIST/NUMN rebuilt when
  !> its ntc was 1 (list_due), and with atomstress the per-atom diagonal
  !> stress over that half list (src/dynamics_m.F90:186-198, 279-283).  The
  !> positions are those of the evaluation.
  subroutine dynamics_after_force(list_due)
    logical, intent(in) :: list_due
    type(simparameters) :: sp
    sp = get_params()
    if (list_due) call update_neighbourlist
    if (sp%atomstress) call mdk_check(mdk_atomic_stress(mdk_dev_handle(), stressx, stressy, stressz), &
                                      'mdk_atomic_stress')
  end subroutine dynamics_after_force

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

  !> write_stress (src/dynamics_m.F90:629-647): the stress tensor
  !> -B0 (tp + tk)^T / vol in GPa, one row per line
  subroutine write_stress(unit)
    integer :: unit
    integer :: i, j, k
    real(kind_wp) :: sigma(3,3)
    do i = 1, 3
       do j = 1, 3
          sigma(i,j) = -1.d0*sum((/(b0(j,k)*(tp(i,k)+tk(i,k)), k = 1, 3)/))/vol
       end do
    end do
    do i = 1, 3
       write(unit, '("STRESS I=",I2, 3e14.6," GPa")') i, (sigma(i,j)*press_to_gpa, j = 1, 3)
    end do
  end subroutine write_stress

  !> write_atomic_stress (src/dynamics_m.F90:765-792): atomicStressNN.out
  !> (NN the counter, zero-padded below 10), or lastAtomicStress.out for -1;
  !> one line per atom: stressx, stressy, stressz, x0, y0, z0
  subroutine write_atomic_stress(file_counter)
    integer :: file_counter
    integer :: u, i
    character(len=40) :: fname
    type(simparameters) :: sp
    sp = get_params()
    if (file_counter == -1) then
       fname = 'lastAtomicStress.out'
    else if (file_counter < 10) then
       write(fname, '("atomicStress0",i0,".out")') file_counter
    else
       write(fname, '("atomicStress",i0,".out")') file_counter
    end if
    open(newunit=u, file=trim(fname), status='unknown')
    do i = 1, sp%nm
       write(u,*) stressx(i), stressy(i), stressz(i), x0(i), y0(i), z0(i)
    end do
    close(u)
  end subroutine write_atomic_stress

  !> energy_pair_emb (src/dynamics_m.F90:797-806): total potential energy,
  !> the sum of energy_calc's per-atom energies
  function energy_pair_emb()
    real(kind_wp) :: energy_pair_emb
    call energy_calc
    energy_pair_emb = sum(en_atom)
  end function energy_pair_emb
end module dynamics_m

module quench_m
  use iso_c_binding
  use mdk_c
  use constants_m
  use params_m
  use particles_m
  use parrinellorahman_m
  use dynamics_m
  use system_m
  use mdk_device_m
  implicit none
  private
  public :: velocityverlet, predic, correc
  public :: quench, init_quench_m, cleanup_quench_m
  integer, save :: nvari = 0, nevals = 0
  real(kind_wp), allocatable, save :: qx(:), qg(:)   ! variables and gradient
contains
  !> init_quench_m / cleanup_quench_m (src/quench_m.F90:83-91): the work
  !> vectors of 3 nm + 9 variables
  subroutine init_quench_m
    type(simparameters) :: sp
    integer :: istat
    sp = get_params()
    nvari = 3*sp%nm + 9
    allocate(qx(nvari), qg(nvari), stat=istat)
    if (istat /= 0) stop 'ALLOCATION ERROR (init_quench_m) - Possibly out of memory.'
  end subroutine init_quench_m
  subroutine cleanup_quench_m
    if (allocated(qx)) deallocate(qx, qg)
  end subroutine cleanup_quench_m

  !> The objective of quench (derivs, src/quench_m.F90:214-322): H = pe +
  !> press vol at the point x (fractional positions atom by atom, then B0 row
  !> by row) and its gradient, dH/ds_i = -tg f_i with tg = B0^T B0, dH/dB0 =
  !> -tp + press tc for the box elements that move (all for ivol < 1, zz for
  !> ivol = 4), from one device force evaluation.
  subroutine q_eval(x, hval, g)
    real(kind_wp), intent(in) :: x(:)
    real(kind_wp), intent(out) :: hval, g(:)
    type(simparameters) :: sp
    integer :: i, j, m
    real(kind_wp) :: t(3,3)
    sp = get_params()
    m = 3*sp%nm
    do i = 1, 3
       do j = 1, 3
          m = m + 1
          b0(i,j) = x(m)
       end do
    end do
    do i = 1, sp%nm
       x0(i) = x(3*i-2)
       y0(i) = x(3*i-1)
       z0(i) = x(3*i)
    end do
    call force
    nevals = nevals + 1
    t = matmul(transpose(b0), b0)
    do i = 1, sp%nm
       g(3*i-2) = -(t(1,1)*fx(i) + t(1,2)*fy(i) + t(1,3)*fz(i))
       g(3*i-1) = -(t(2,1)*fx(i) + t(2,2)*fy(i) + t(2,3)*fz(i))
       g(3*i)   = -(t(3,1)*fx(i) + t(3,2)*fy(i) + t(3,3)*fz(i))
    end do
    m = 3*sp%nm
    do i = 1, 3
       do j = 1, 3
          m = m + 1
          g(m) = 0.0_kind_wp
          if (sp%ivol < 1 .or. (sp%ivol == 4 .and. i*j == 9)) g(m) = -tp(i,j) + sp%press*tc(i,j)
       end do
    end do
    hval = pe + sp%press*vol
  end subroutine q_eval

  !> Minimum of phi(a) = H(x + a d), a > 0: bracketing with doubling steps,
  !> then safeguarded secant steps on phi'(a) = g(x + a d).d; accepts a point
  !> with |phi'| <= 0.1 |phi'(0)| and H below H(x), else returns the lowest
  !> point seen (a = 0 when none is lower).
  subroutine q_line(x, hx, gx, d, step0, hb, xb, gb, ab)
    real(kind_wp), intent(in) :: x(:), hx, gx(:), d(:), step0
    real(kind_wp), intent(out) :: hb, xb(:), gb(:), ab
    real(kind_wp), allocatable :: xa(:), ga(:)
    real(kind_wp) :: d0, lo, flo, dlo, hi, fhi, dhi, a, fa, da, t, w
    logical :: have_hi
    integer :: it
    allocate(xa(size(x)), ga(size(x)))
    d0 = dot_product(gx, d)
    lo = 0.0_kind_wp; flo = hx; dlo = d0
    hi = 0.0_kind_wp; fhi = 0.0_kind_wp; dhi = 0.0_kind_wp; have_hi = .false.
    a = step0
    hb = hx; xb = x; gb = gx; ab = 0.0_kind_wp
    do it = 1, 40
       xa = x + a*d
       call q_eval(xa, fa, ga)
       da = dot_product(ga, d)
       if (fa < hb) then
          hb = fa; xb = xa; gb = ga; ab = a
       end if
       if (abs(da) <= 0.1_kind_wp*abs(d0) .and. fa <= hx) then
          hb = fa; xb = xa; gb = ga; ab = a
          exit
       end if
       if (da > 0.0_kind_wp .or. fa > flo) then
          hi = a; fhi = fa; dhi = da; have_hi = .true.
       else
          lo = a; flo = fa; dlo = da
       end if
       if (.not. have_hi) then
          if (da < 0.0_kind_wp) a = 2.0_kind_wp*a
          cycle
       end if
       if (dhi > 0.0_kind_wp .and. dlo < 0.0_kind_wp) then
          t = lo + (hi - lo)*(-dlo)/(dhi - dlo)
       else
          t = 0.5_kind_wp*(lo + hi)
       end if
       w = hi - lo
       a = min(max(t, lo + 0.1_kind_wp*w), hi - 0.1_kind_wp*w)
    end do
    deallocate(xa, ga)
  end subroutine q_line

  !> quench (src/quench_m.F90:100-205): relaxes positions and, for ivol < 1
  !> (zz only for ivol = 4), the box to the minimum of H = pe + press vol,
  !> each point evaluated by the device force.  The minimiser is this
  !> project's Polak-Ribiere conjugate gradient (PR+, restarted every
  !> max(3 nm + 9, 50) iterations or on a non-descent direction) with the line
  !> search above; as the reference's VA14CD call, the first step is sized for
  !> an expected decrease dfn = 0.1 and the run ends once two successive
  !> iterations lower H by less than acc = 5e-7 nm (src/quench_m.F90:134-136).
  !> Afterwards velocities, higher derivatives and box rates are zero, as
  !> there.
  subroutine quench
    type(simparameters) :: sp
    real(kind_wp), allocatable :: d(:), xn(:), gn(:)
    real(kind_wp) :: hx, hn, a, acc, dfn, gd, gdn, beta, step, dec
    integer :: i, j, m, it, small
    sp = get_params()
    if (.not. allocated(qx)) call init_quench_m
    write(*,'(A,I9,A)') ' QUENCH, FINDING MINIMUM ENERGY CONFIGURATION OF', sp%nm, ' ATOMS'
    do i = 1, sp%nm
       qx(3*i-2) = x0(i)
       qx(3*i-1) = y0(i)
       qx(3*i) = z0(i)
    end do
    m = 3*sp%nm
    do i = 1, 3
       do j = 1, 3
          m = m + 1
          qx(m) = b0(i,j)
       end do
    end do
    allocate(d(nvari), xn(nvari), gn(nvari))
    dfn = 0.1_kind_wp
    acc = 5.0e-7_kind_wp*sp%nm
    nevals = 0
    call q_eval(qx, hx, qg)
    d = -qg
    step = 2.0_kind_wp*dfn/max(dot_product(qg, qg), 1.0e-300_kind_wp)
    small = 0
    do it = 1, 20000
       gd = dot_product(qg, d)
       if (gd >= 0.0_kind_wp) then
          d = -qg
          gd = -dot_product(qg, qg)
       end if
       call q_line(qx, hx, qg, d, step, hn, xn, gn, a)
       if (a == 0.0_kind_wp) then
          if (all(d == -qg)) exit
          d = -qg
          cycle
       end if
       dec = hx - hn
       beta = max(0.0_kind_wp, dot_product(gn, gn - qg)/max(dot_product(qg, qg), 1.0e-300_kind_wp))
       if (mod(it, max(3*sp%nm + 9, 50)) == 0) beta = 0.0_kind_wp
       d = -gn + beta*d
       gdn = dot_product(gn, d)
       if (gdn < 0.0_kind_wp) then
          step = a*gd/gdn
       else
          step = a
       end if
       qx = xn; hx = hn; qg = gn
       if (dec < acc) then
          small = small + 1
       else
          small = 0
       end if
       if (small >= 2) exit
    end do
    call q_eval(qx, hx, qg)   ! the host and device state at the minimum
    write(*,'(A,ES18.10,A,F10.6,A,ES12.4,A,ES13.7)') 'FINAL U= ', pe/sp%nm, ' V=', vol/sp%nm, &
         ' P=', sp%press*press_to_gpa, ' H=', hx/sp%nm
    deallocate(d, xn, gn)
    b1 = 0.0_kind_wp; b2 = 0.0_kind_wp; b3 = 0.0_kind_wp
    call pr_get_metric_tensor
    x1 = 0.0_kind_wp; y1 = 0.0_kind_wp; z1 = 0.0_kind_wp
    x2 = 0.0_kind_wp; y2 = 0.0_kind_wp; z2 = 0.0_kind_wp
    x3 = 0.0_kind_wp; y3 = 0.0_kind_wp; z3 = 0.0_kind_wp
    ke = 0.0_kind_wp
    te = ke + pe
    h = te + sp%press*vol
    th = h
    sp = get_params()
    sp%boxtem = 0.0_kind_wp
    call set_params(sp)
  end subroutine quench

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
    type(simparameters) :: sp
    logical :: list_due
    sp = get_params()
    list_due = (sp%ntc + 1 == 1)   ! the force inside rebuilds NLIST/NUMN
    call mdk_dev_push()
    call mdk_check(mdk_velocityverlet(mdk_dev_handle()), 'mdk_velocityverlet')
    call mdk_dev_pull(.true.)
    call dynamics_pull_forces()
    call mdk_check(mdk_get_tensor(mdk_dev_handle(), MDK_TGID, tgid), 'get tgid')
    call dynamics_after_force(list_due)
  end subroutine velocityverlet
end module quench_m
