!
! ref_driver.F90 -- drives the REFERENCE Fortran modules of gjackland/moldy
! (compiled from /root/reference/src by oracle/build_ref.sh into oracle/_ref/)
! on explicit inputs and dumps their outputs.  Test infrastructure only: it
! produces the golden vectors under tests/golden/ and times the reference's
! own hot path for the CPU baseline.  It contains no reference code; it only
! calls the reference modules' public interfaces in the order of
! src/moldin.F90:172-300 (set-up), :407 (linkup), :446-477 (first force and
! accelerations) and :752-791 (MD step with iverlet=1).  set_up is called
! BEFORE potential_set_lookup so that lookup mode is usable with the
! fetransmetal potential (moldin calls them the other way round).
! Built with -DFETRANSMETAL (fetransmetal module) or -DIRON_CARBON
! (iron_carbon module: one species, Fe, as read_system's generic branch sets it).
!
! usage: ref_driver <input.bin> <output.bin>     (stream-unformatted, little endian)
! input : int32 n, ivol, ntcm, nsteps, uselookup, nnbrs, flags
!         real64 deltat_fs, rpad, press_gpa, boxmass, temprq, nose, b0(3,3)
!         real64 x0(n),y0(n),z0(n),x1(n),y1(n),z1(n),mass(n); int32 z(n)
!         int32 ngrid; real64 rgrid(ngrid); int32 nrho; real64 rhogrid(nrho)
! flags : bit 0 = timing only (no dumps, prints wall-clock seconds of the MD loop)
!         bit 1 = velocities as moldin makes them for a new run: rstart(97,33,21,82)
!                 (moldin.F90:175), the 3 draws per atom of read_system
!                 (io_m.F90:233-235), then setvel and scalev (moldin.F90:307-308)
!         bit 7 = extra record after the final one: embed(pexx) on RHO scaled by
!                 1.01 with RHO(1) = 0, then pexx, AFRHO, DAFRHO
!         bit 6 = potin (potential.dat) for the Fe-Fe pair, one species, after
!                 init_linkcell_m (moldin.F90:266-282)
!         bit 5 = quench after the first force (moldin.F90:566-572: init_quench_m,
!                 quench, cleanup_quench_m) instead of MD; pe and vol follow
!                 the final record
!         bit 4 = Nose thermostat as moldin drives it (moldin.F90:828-835): scalev
!                 whenever the accumulated time reaches fs = 0.098227 (REAL(4)),
!                 with RQKE of the strain loop (moldin.F90:437)
!         bit 3 = Gear predictor-corrector (iverlet=0): PREDIC, FORCE, CORREC
!                 (moldin.F90:754-758) instead of velocityverlet; x3, y3, z3
!                 and b3 follow the final record
!         bit 2 = extra record: per step pe, ke, vol, box kinetic energy
!                 (moldin.F90:768-774), b0(3,3); after the run en_atom (energy_calc)
!         bit 8 = atomstress on (the force loop's per-atom diagonal stress,
!                 dynamics_m.F90:279-283); extra record after all others: the
!                 stress after the first force as (3, n), then stressx,
!                 stressy, stressz after the last
!
program ref_driver
  use constants_m
  use params_m
  use particles_m
  use system_m
  use parrinellorahman_m
  use matrix_m
  use potential_m
  use linkcell_m
  use neighbourlist_m
  use thermostat_m
  use dynamics_m
  use quench_m
  use random_m
  implicit none

  character(len=1024) :: inpath, outpath
  integer :: n, ivol, ntcm, nsteps, uselk, nnbrs_in, flags
  integer :: ngrid, nrho, i, k, istep, ierror, c0, c1, crate, nspec_save, ispec_save
  real(kind_wp) :: dt_fs, rpad_in, press_gpa, boxmass, b0in(3,3)
  real(kind_wp) :: rmaxpot, s1, s2, s3, a1, a2, a3, temprq_in, rdum, bke, nose_in, thermotimer
  real(kind_wp) :: fs_ref = 0.098227, pexx
  real(kind_wp), allocatable :: rgrid(:), rhogrid(:), trace(:,:), xtrace(:,:), stress0(:,:)
  type(simparameters) :: sp
  logical :: timing

  call get_command_argument(1, inpath)
  call get_command_argument(2, outpath)

  open(10, file=trim(inpath), access='stream', form='unformatted', status='old')
  read(10) n, ivol, ntcm, nsteps, uselk, nnbrs_in, flags
  read(10) dt_fs, rpad_in, press_gpa, boxmass, temprq_in, nose_in
  read(10) b0in
  timing = iand(flags, 1) /= 0

  ! parameters as read_params would leave them (params_m.F90:187-210, :412, :454)
  sp = get_params()
  sp%nm = n
  sp%ivol = ivol
  sp%iverlet = 1
  if (iand(flags, 8) /= 0) sp%iverlet = 0
  sp%ntcm = ntcm
  sp%nnbrs = nnbrs_in
  sp%deltat = dt_fs*fs_to_time
  sp%rpad = rpad_in
  sp%press = press_gpa*gpa_to_press
  sp%rcut = 0.0d0
  sp%temprq = temprq_in
  sp%nose = nose_in
  sp%ntc = 0
#ifdef IRON_CARBON
  sp%nspec = 1
#else
  sp%nspec = 24
#endif
  sp%uselookup = (uselk /= 0)
  sp%bmass = boxmass
  sp%bmass = sp%nm*sp%bmass
  sp%bdel2 = sp%deltat*sp%deltat/(2.d0*sp%bmass)
  sp%atomstress = iand(flags, 256) /= 0
  call set_params(sp)

  call init_particles_m
  call init_neighbourlist_m
  call init_thermostat_m
  call init_parrinellorahman_m

  b0 = b0in
  read(10) x0, y0, z0, x1, y1, z1, atomic_mass
  read(10) atomic_number
  read(10) ngrid
  allocate(rgrid(max(ngrid,1)))
  if (ngrid > 0) read(10) rgrid
  read(10) nrho
  allocate(rhogrid(max(nrho,1)))
  if (nrho > 0) read(10) rhogrid
  close(10)
  x2 = 0.0d0; y2 = 0.0d0; z2 = 0.0d0

#ifdef IRON_CARBON
  ! species bookkeeping of read_system's generic branch (io_m.F90:119-125): Fe only
  ispec(1) = 26
  do i = 1, n
     if (atomic_number(i) == 26) then
        atomic_index(i) = 1
     else
        atomic_index(i) = 0
     end if
  end do
#else
  ! species bookkeeping of read_system's FETRANSMETAL branch (io_m.F90:133-163)
  ispec(1:24) = (/ 22,23,24,25,26,27,28,29, 40,41,42,43,44,45,46,47, 72,73,74,75,76,77,78,79 /)
  do i = 1, n
     if (atomic_number(i) == 26) then
        atomic_index(i) = 5
     else
        atomic_index(i) = 0
     end if
  end do
#endif

  call get_potential(sp%nspec, ispec(1:sp%nspec), rmaxpot, ierror)
  sp = get_params()
  if (rmaxpot > sp%rcut) sp%rcut = rmaxpot
  call set_params(sp)
#ifndef IRON_CARBON
  call set_up
#endif
  if (sp%uselookup) then
     call potential_set_lookup
  else
     call potential_set_exact
  end if

  sp = get_params()
  sp%rnear = sp%rcut + sp%rpad
  sp%nlcx = int(b0(1,1)/sp%rnear)
  sp%nlcy = int(b0(2,2)/sp%rnear)
  sp%nlcz = int(b0(3,3)/sp%rnear)
  if (sp%nlcx < 3) sp%nlcx = 3
  if (sp%nlcy < 3) sp%nlcy = 3
  if (sp%nlcz < 3) sp%nlcz = 3
  call set_params(sp)
  call init_linkcell_m

  if (iand(flags, 64) /= 0) then
     ! potin's potential.dat for the Fe-Fe pair alone (one species)
     sp = get_params()
     nspec_save = sp%nspec
     ispec_save = ispec(1)
     sp%nspec = 1
     call set_params(sp)
     ispec(1) = 26
     call potin
     ispec(1) = ispec_save
     sp%nspec = nspec_save
     call set_params(sp)
  end if

  call init_dynamics_m
  call pr_get_metric_tensor
  do i = 1, n
     if (atomic_mass(i) > 0.01) then
        del2(i) = sp%deltat*sp%deltat/(2.0*atomic_mass(i))
     else
        del2(i) = 0.0
     end if
  end do

  if (iand(flags, 2) /= 0) then
     call rstart(97, 33, 21, 82)
     do i = 1, 3*n
        rdum = rand1()
     end do
     call setvel
     call scalev
  end if

  call linkup
  call force
  if (iand(flags, 256) /= 0) then
     allocate(stress0(3, n))
     stress0(1,:) = stressx; stress0(2,:) = stressy; stress0(3,:) = stressz
  end if
  call matrix_invert(tg, tginv)
  call matrix_multiply(nmat, tginv, tgdot, tgid)
  do i = 1, n
     if (atomic_index(i) == 0) cycle
     a1 = tgid(1,1)*x1(i)+tgid(1,2)*y1(i)+tgid(1,3)*z1(i)
     a2 = tgid(2,1)*x1(i)+tgid(2,2)*y1(i)+tgid(2,3)*z1(i)
     a3 = tgid(3,1)*x1(i)+tgid(3,2)*y1(i)+tgid(3,3)*z1(i)
     x2(i) = del2(i)*fx(i)-0.5*a1
     y2(i) = del2(i)*fy(i)-0.5*a2
     z2(i) = del2(i)*fz(i)-0.5*a3
  end do
  b2 = sp%bdel2*fb

  if (.not. timing) then
     open(20, file=trim(outpath), access='stream', form='unformatted', status='replace')
     ! --- first force (state after moldin.F90:446-477)
     write(20) sp%nlcx, sp%nlcy, sp%nlcz
     write(20) sp%rnear, sp%deltat, sp%bdel2, sp%press, vol, pe
     write(20) rho, afrho, dafrho, fx, fy, fz, x2, y2, z2
     write(20) tp, tk, tc, fb, tg, tgdot, tgid, b2
     write(20) numn
     do i = 1, n
        write(20) (nlist(k,i), k = 1, numn(i))
     end do
     ! --- potential functions through potential_m's dispatch
     write(20) (vee(rgrid(k),26,26), k = 1, ngrid)
     write(20) (dvee(rgrid(k),26,26), k = 1, ngrid)
     write(20) (phi(rgrid(k),26,26), k = 1, ngrid)
     write(20) (dphi(rgrid(k),26,26), k = 1, ngrid)
     write(20) (emb(rhogrid(k),26), k = 1, nrho)
     write(20) (demb(rhogrid(k),26), k = 1, nrho)
  end if

  if (iand(flags, 32) /= 0) then
     call init_quench_m
     call quench
     call cleanup_quench_m
  end if
  allocate(trace(6, max(nsteps,1)), xtrace(13, max(nsteps,1)))
  sp = get_params()
  sp%rqke = 1.5d0*bk*sp%temprq*(sp%nm-3)
  call set_params(sp)
  thermotimer = 0.0d0
  call system_clock(c0, crate)
  do istep = 1, nsteps
     if (iand(flags, 8) /= 0) then
        call predic
        call force
        call correc
     else
        call velocityverlet(sp%deltat)
     end if
     call update_therm_avs(s1, s2, s3)
     ke = s3/sp%deltat**2/2.d0
     trace(1, istep) = pe
     trace(2, istep) = ke
     trace(3, istep) = s1
     trace(4, istep) = s3
     trace(5, istep) = vol
     sp = get_params()
     trace(6, istep) = real(sp%ntc, kind_wp)
     bke = 0.0d0
     do i = 1, 3
        do k = 1, 3
           bke = bke + (b1(i,k)**2)/(sp%bdel2)
        end do
     end do
     xtrace(1, istep) = pe
     xtrace(2, istep) = ke
     xtrace(3, istep) = vol
     xtrace(4, istep) = bke
     xtrace(5:13, istep) = reshape(b0, (/ 9 /))
     if (sp%ntc == 0) call relink
     if (iand(flags, 16) /= 0) then
        thermotimer = thermotimer + sp%deltat
        if (thermotimer >= fs_ref) then
           thermotimer = thermotimer - fs_ref
           if (sp%nose /= 0) call scalev
        end if
     end if
  end do
  call system_clock(c1)

  if (timing) then
     write(*,'(a,i0,a,es16.8)') 'REF_TIMING steps=', nsteps, ' seconds=', real(c1-c0,kind_wp)/real(crate,kind_wp)
  else
     ! --- after nsteps MD steps
     write(20) trace
     write(20) x0, y0, z0, x1, y1, z1, x2, y2, z2, fx, fy, fz, rho, dafrho
     write(20) b0, b1, b2, tp, fb, tgid, tk
     if (iand(flags, 8) /= 0) write(20) x3, y3, z3, b3
     if (iand(flags, 32) /= 0) write(20) pe, vol
     if (iand(flags, 4) /= 0) then
        call energy_calc
        write(20) xtrace
        write(20) en_atom
        write(20) x1, y1, z1
     end if
     if (iand(flags, 128) /= 0) then
        ! embed on a caller-changed RHO: scaled by 1.01, atom 1 set to 0 (below the
        ! 1e-40 cut: its AFRHO and DAFRHO keep their values)
        rho = 1.01d0*rho
        rho(1) = 0.0d0
        call embed(pexx)
        write(20) pexx, afrho, dafrho
     end if
     if (iand(flags, 256) /= 0) write(20) stress0, stressx, stressy, stressz
     close(20)
  end if
end program ref_driver
