!
! state_m.F90 -- the reference's base and state modules, restated for the GPU
! drop-in: same module names and the public names the hot path's callers use
! (src/constants_m.F90, src/params_m.F90:44-184, src/matrix_m.F90:60-119,
! src/system_m.F90:45-49, src/random_m.F90:54-143, src/particles_m.F90:52-94,
! src/parrinellorahman_m.F90:91-261).  Host arrays keep the reference's layout;
! the device copies live behind the C ABI (mdk_device_m).
!
module constants_m
  implicit none
  public
  integer, parameter :: kind_dp = kind(1.d0), kind_sp = kind(1.0), kind_wp = kind_dp
  integer, parameter :: stdin = 5, stdout = 6, stderr = 0, unit_stdout = 1, unit_stderr = stderr
  integer, parameter :: nmat = 3
  real(kind_wp), parameter :: avogadro = 6.02214179d+23, electron = 1.602176487d-19
  real(kind_wp), parameter :: angstrom = 1.d-10, boltzmann = 1.3806504d-23
  real(kind_wp), parameter :: bk = boltzmann/electron
  real(kind_wp), parameter :: time_to_sec = 1.018050530893617d-014
  real(kind_wp), parameter :: time_to_fs = time_to_sec/1.d-15
  real(kind_wp), parameter :: fs_to_time = 1.d0/time_to_fs
  real(kind_wp), parameter :: press_to_gpa = (electron/angstrom**3)/1.0d9
  real(kind_wp), parameter :: gpa_to_press = 1.0d0/press_to_gpa
end module constants_m

module params_m
  use constants_m
  implicit none
  private
  public :: simparameters, get_params, set_params

  !> The subset of the reference's simulation parameters the hot path reads.
  type simparameters
     integer :: ivol = -1, iverlet = 0, nnbrs = 150
     integer :: nlcx = 3, nlcy = 3, nlcz = 3
     integer :: nm = 0, nspec = 0
     real(kind_wp) :: dsp = 0._kind_wp, rpad = 0._kind_wp, rcut = 0._kind_wp, rnear = 0._kind_wp
     real(kind_wp) :: deltat = 0._kind_wp, temprq = 0._kind_wp, rqke = 0._kind_wp
     real(kind_wp) :: press = 0._kind_wp
     real(kind_wp) :: nose = 0._kind_wp, therm_con = -1._kind_wp   ! Nose thermostat (src/params_m.F90:102-103)
     integer :: nsteps = 0, ntcm = 10, ntc = 0
     logical :: uselookup = .false., atomstress = .false.
     real(kind_wp) :: boxtem = 0._kind_wp, bdel2 = 0._kind_wp, bmass = 0._kind_wp
  end type simparameters

  type(simparameters), save :: simparam_global

contains
  function get_params()
    type(simparameters) :: get_params
    get_params = simparam_global
  end function get_params
  subroutine set_params(p)
    type(simparameters), intent(in) :: p
    simparam_global = p
  end subroutine set_params
end module params_m

module matrix_m
  use constants_m
  implicit none
  private
  public :: matrix_multiply, matrix_invert
contains
  !> c = a b for n x n matrices (src/matrix_m.F90:60-79)
  subroutine matrix_multiply(n, a, b, c)
    integer :: n, i, j, k
    real(kind_wp), intent(in) :: a(n,n), b(n,n)
    real(kind_wp), intent(out) :: c(n,n)
    do i = 1, n
       c(i,:) = 0.0d0
       do k = 1, n
          do j = 1, n
             c(i,j) = c(i,j) + a(i,k)*b(k,j)
          end do
       end do
    end do
  end subroutine matrix_multiply

  !> cofactor inverse of a 3x3 matrix (src/matrix_m.F90:89-119)
  subroutine matrix_invert(a, b)
    real(kind_wp) :: a(3,3), b(3,3), deta
    deta = a(1,1)*(a(2,2)*a(3,3)-a(2,3)*a(3,2))
    deta = deta + a(2,1)*(a(3,2)*a(1,3)-a(1,2)*a(3,3))
    deta = deta + a(3,1)*(a(1,2)*a(2,3)-a(1,3)*a(2,2))
    b(1,1) = a(2,2)*a(3,3)-a(2,3)*a(3,2)
    b(1,2) = a(3,1)*a(2,3)-a(2,1)*a(3,3)
    b(1,3) = a(2,1)*a(3,2)-a(3,1)*a(2,2)
    b(2,1) = a(3,2)*a(1,3)-a(1,2)*a(3,3)
    b(2,2) = a(1,1)*a(3,3)-a(1,3)*a(3,1)
    b(2,3) = a(3,1)*a(1,2)-a(1,1)*a(3,2)
    b(3,1) = a(1,2)*a(2,3)-a(1,3)*a(2,2)
    b(3,2) = a(2,1)*a(1,3)-a(2,3)*a(1,1)
    b(3,3) = a(1,1)*a(2,2)-a(1,2)*a(2,1)
    if (abs(deta) == 0.d0) stop "Error in matrix_invert: determinant is or is close to zero. STOP."
    b(:,:) = b(:,:)/deta
  end subroutine matrix_invert
end module matrix_m

module system_m
  use constants_m, only : kind_wp
  implicit none
  public
  real(kind_wp) :: pe = 0, ke = 0, te = 0, h = 0, th = 0
end module system_m

!> random_m (src/random_m.F90:54-143): the Marsaglia-Zaman-Tsang RANMAR
!> generator (Marsaglia, Zaman and Tsang, Stat. Prob. Lett. 9 (1990) 35-39):
!> a lagged-Fibonacci difference sequence over a 97-entry table seeded by
!> rstart, combined with an arithmetic sequence of step 7654321/2^24.
module random_m
  use constants_m, only : kind_wp
  implicit none
  private
  public :: rand1, rstart
  real(kind_wp), save :: tab(97) = 0.0_kind_wp, cc = 0.0_kind_wp, cstep = 0.0_kind_wp, cmod = 0.0_kind_wp
  integer, save :: ip = 97, jp = 33
contains
  !> next uniform deviate in [0, 1)
  function rand1()
    real(kind_wp) :: rand1, x
    x = tab(ip) - tab(jp)
    if (x < 0.0_kind_wp) x = x + 1.0_kind_wp
    tab(ip) = x
    ip = merge(97, ip - 1, ip == 1)
    jp = merge(97, jp - 1, jp == 1)
    cc = cc - cstep
    if (cc < 0.0_kind_wp) cc = cc + cmod
    x = x - cc
    if (x < 0.0_kind_wp) x = x + 1.0_kind_wp
    rand1 = x
  end function rand1

  !> seeds in 1..168, not all 1 (moldin uses 97, 33, 21, 82): each table
  !> entry takes 24 bits from a 3-lag multiplicative sequence mod 179 and a
  !> linear congruential sequence mod 169
  subroutine rstart(iw, ix, iy, iz)
    integer, intent(in) :: iw, ix, iy, iz
    integer :: a, b, c, d, m, e, bit
    real(kind_wp) :: acc, w
    a = iw; b = ix; c = iy; d = iz
    do e = 1, 97
       acc = 0.0_kind_wp
       w = 0.5_kind_wp
       do bit = 1, 24
          m = mod(mod(a*b, 179)*c, 179)
          a = b; b = c; c = m
          d = mod(53*d + 1, 169)
          if (mod(d*m, 64) >= 32) acc = acc + w
          w = 0.5_kind_wp*w
       end do
       tab(e) = acc
    end do
    cc = 362436.0_kind_wp/16777216.0_kind_wp
    cstep = 7654321.0_kind_wp/16777216.0_kind_wp
    cmod = 16777213.0_kind_wp/16777216.0_kind_wp
    ip = 97
    jp = 33
  end subroutine rstart
end module random_m

module particles_m
  use constants_m, only : kind_wp
  use params_m, only : simparameters, get_params
  implicit none
  private
  public :: init_particles_m, cleanup_particles_m

  integer, allocatable, public :: atomic_index(:), ispec(:), atomic_number(:)
  real(kind_wp), allocatable, public :: atomic_mass(:), del2(:)
  real(kind_wp), allocatable, public :: x0(:), y0(:), z0(:), x1(:), y1(:), z1(:)
  real(kind_wp), allocatable, public :: x2(:), y2(:), z2(:), x3(:), y3(:), z3(:)
  real(kind_wp), allocatable, public :: afrho(:), dafrho(:), rho(:), en_atom(:)
  real(kind_wp), allocatable, public :: stressx(:), stressy(:), stressz(:)   ! atomstress (:66)
contains
  subroutine init_particles_m()
    type(simparameters) :: sp
    integer :: n
    sp = get_params()
    n = sp%nm
    allocate(atomic_index(n), atomic_number(n), atomic_mass(n), del2(n), ispec(112))
    allocate(x0(n), y0(n), z0(n), x1(n), y1(n), z1(n), x2(n), y2(n), z2(n), x3(n), y3(n), z3(n))
    allocate(afrho(n), dafrho(n), rho(n), en_atom(n), stressx(n), stressy(n), stressz(n))
    stressx = 0; stressy = 0; stressz = 0
    ispec = 0
    x0 = 0; y0 = 0; z0 = 0; x1 = 0; y1 = 0; z1 = 0; x2 = 0; y2 = 0; z2 = 0; x3 = 0; y3 = 0; z3 = 0
    afrho = 0; dafrho = 0; rho = 0; en_atom = 0; del2 = 0
  end subroutine init_particles_m
  subroutine cleanup_particles_m()
    deallocate(atomic_index, atomic_number, atomic_mass, del2, ispec)
    deallocate(x0, y0, z0, x1, y1, z1, x2, y2, z2, x3, y3, z3, afrho, dafrho, rho, en_atom)
    deallocate(stressx, stressy, stressz)
  end subroutine cleanup_particles_m
end module particles_m

module parrinellorahman_m
  use constants_m
  use params_m
  use matrix_m
  implicit none
  private
  public :: init_parrinellorahman_m, cleanup_parrinellorahman_m, pr_get_metric_tensor
  public :: get_tgid, set_tgid, pr_get_tgid, pr_get_realsep_from_dx, pr_get_realsep2_from_dx

  real(kind_wp), public :: b0(nmat,nmat) = 0.d0, b1(nmat,nmat) = 0.d0, b2(nmat,nmat) = 0.d0
  real(kind_wp), public :: b3(nmat,nmat) = 0.d0
  real(kind_wp), public :: tk(nmat,nmat) = 0.d0, tc(nmat,nmat) = 0.d0, tg(nmat,nmat) = 0.d0
  real(kind_wp), public :: tp(nmat,nmat) = 0.d0, tginv(nmat,nmat) = 0.d0, tgdot(nmat,nmat) = 0.d0
  real(kind_wp), public :: tgid(nmat,nmat) = 0.d0, fb(nmat,nmat) = 0.d0
  real(kind_wp), public :: vol = 0.d0
  type(simparameters), save :: simparam
contains
  subroutine init_parrinellorahman_m
    simparam = get_params()
  end subroutine init_parrinellorahman_m
  subroutine cleanup_parrinellorahman_m
  end subroutine cleanup_parrinellorahman_m

  !> minimum image and real separation (host version, for set-up and checks;
  !> the device applies the same arithmetic per pair)
  subroutine pr_get_realsep_from_dx(r, dx, dy, dz)
    real(kind_wp), intent(out) :: r
    real(kind_wp), intent(inout) :: dx, dy, dz
    real(kind_wp) :: rsq
    call pr_get_realsep2_from_dx(rsq, dx, dy, dz)
    r = sqrt(rsq)
  end subroutine pr_get_realsep_from_dx
  subroutine pr_get_realsep2_from_dx(rsq, dx, dy, dz)
    real(kind_wp), intent(out) :: rsq
    real(kind_wp), intent(inout) :: dx, dy, dz
    real(kind_wp) :: rx, ry, rz
    select case (simparam%ivol)
    case (0, 1, 4)
       dx = dx - nint(dx); dy = dy - nint(dy); dz = dz - nint(dz)
    case (2)
       dx = dx - nint(dx); dy = dy - nint(dy)
    case (5)
       dz = dz - nint(dz)
    end select
    rx = b0(1,1)*dx + b0(1,2)*dy + b0(1,3)*dz
    ry = b0(2,1)*dx + b0(2,2)*dy + b0(2,3)*dz
    rz = b0(3,1)*dx + b0(3,2)*dy + b0(3,3)*dz
    rsq = rx*rx + ry*ry + rz*rz
  end subroutine pr_get_realsep2_from_dx

  !> box volume det(B0)
  subroutine pr_get_metric_tensor
    vol = b0(1,1)*(b0(2,2)*b0(3,3)-b0(2,3)*b0(3,2)) + &
          b0(2,1)*(b0(3,2)*b0(1,3)-b0(1,2)*b0(3,3)) + &
          b0(3,1)*(b0(1,2)*b0(2,3)-b0(2,2)*b0(1,3))
  end subroutine pr_get_metric_tensor

  function get_tgid()
    real(kind_wp) :: get_tgid(nmat,nmat)
    get_tgid = tgid
  end function get_tgid
  subroutine set_tgid(t)
    real(kind_wp), intent(in) :: t(:,:)
    tgid(:,:) = t(:,:)
  end subroutine set_tgid

  subroutine pr_get_tgid
    call matrix_invert(tg, tginv)
    call matrix_multiply(nmat, tginv, tgdot, tgid)
  end subroutine pr_get_tgid
end module parrinellorahman_m
