! fdtd_gpu_driver.F90 -- Fortran driver for the gfx950 time-stepping library.
!
! Keeps the call structure of semba-fdtd: the set-up of launch_simulation
! (timestepping.F90:305-1014: spacings, calc_G1G2Gm1Gm2, InitCPMLBorders,
! InitMURBorders, InitEDispersives, InitMDispersives, crea_timevector, nodal
! sources, plane waves, probes) done in Fortran
! through the library's table builders, then the loop of ciclo_temporal
! (timestepping.F90:1248-1981) with one library call per reference call site.
!
! usage: fdtd_gpu_driver CASE_FILE OUT [fast|exact] [fused|split|graph|tables]
!   CASE_FILE  binary case written by semba_fdtd_amd.case_file.write_case
!   OUT        int32 nsamples, nprobe; real(4) samples(nprobe, nsamples)
!   fused      fdtd_advance_e / fdtd_advance_h per step (default)
!   split      one call per reference routine (Advance_Ex/Ey/Ez,
!              AdvanceelectricCPML, AdvanceEDispersiveE, AdvanceNodalE, ...)
!   graph      fdtd_run over the whole run (captured HIP graph)
!   tables     stop after the host set-up and write the coefficient tables to
!              OUT instead (g1,g2,gm1,gm2; per axis ide,idh,pbe,pce,pbm,pcm;
!              kappa,beta,g3 of the electric and then the magnetic dispersive
!              media; the Mur coefficients); needs no GPU
program fdtd_gpu_driver
   use iso_c_binding
   use sembafdtd_gpu
   implicit none

   integer, parameter :: PMAX = 8
   type :: i16box
      integer(c_int16_t), allocatable :: a(:,:,:)
   end type
   type :: r4vec
      real(c_float), allocatable :: v(:)
   end type

   character(len=1024) :: casefile, outfile, arg
   integer(c_int32_t) :: mode, loopkind
   integer :: u, ios, c, ax, s, m, n, nd, d, i
   character(len=4) :: magic
   integer(c_int32_t) :: version, nx, ny, nz, nmedia, nsteps, nn(3), hi(3)
   real(c_float) :: eps0, mu0
   real(c_double) :: dt
   real(c_float), allocatable :: epr(:), sigma(:), mur(:), sigmam(:), eps11(:), sig11(:)
   integer(c_int32_t), allocatable :: kind(:), npoles(:)
   real(c_double), allocatable :: pole_c(:,:,:), pole_a(:,:,:)
   real(c_float), allocatable :: mu11(:), sigm11(:)
   integer(c_int32_t), allocatable :: mnpoles(:)
   real(c_double), allocatable :: mpole_c(:,:,:), mpole_a(:,:,:)
   integer(c_int32_t) :: murf(6)
   real(c_float) :: idhf(6)
   real(c_float), allocatable :: cab1(:)
   real(c_float), allocatable :: g1(:), g2(:), gm1(:), gm2(:)
   type(r4vec) :: dd(3), de(3), dh(3), ide(3), idh(3), pbe(3), pce(3), pbm(3), pcm(3)
   integer(c_int32_t) :: layers(6)
   real(c_float) :: order(6), refl(6), alphamax, alphaorden, kappamax
   type(i16box) :: mi(0:5)
   integer(c_int32_t) :: nnodal, ncomp, nhard, nbox(6), numus
   real(c_float) :: amp, deltaevol
   real(c_float), allocatable :: evol(:)
   integer(c_int32_t) :: nprobe
   integer(c_int32_t), allocatable :: pc(:), pi(:), pj(:), pk(:)
   integer(c_int32_t), allocatable :: dmed(:), dnp(:), mdmed(:), mdnp(:)
   real(c_double), allocatable :: kappa(:,:,:), beta(:,:,:), g3(:,:,:)
   real(c_double), allocatable :: mkappa(:,:,:), mbeta(:,:,:), gm3(:,:,:)
   integer :: nmd
   real(c_double) :: kk(2, PMAX), bb(2, PMAX), gg(2, PMAX)
   real(c_float) :: a1(1), a2(1)
   real(c_double), allocatable :: tiempo(:)
   real(c_float), allocatable :: samples(:,:), chunk(:,:)
   integer(c_int32_t) :: nsamp, got, bufsteps, left, k, n0
   integer(c_int32_t) :: npw, pwbox(6), nmodes
   real(c_float), allocatable :: pwp(:), pwe(:), pwinc(:)
   type(r4vec) :: lines(3)
   logical :: has_cpml
   type(fdtd_grid_desc) :: desc
   type(c_ptr) :: h

   call get_command_argument(1, casefile)
   call get_command_argument(2, outfile)
   if (len_trim(casefile) == 0 .or. len_trim(outfile) == 0) then
      write (0, '(a)') 'usage: fdtd_gpu_driver CASE_FILE OUT [fast|exact] [fused|split|graph|tables]'
      stop 2
   end if
   mode = FDTD_MODE_FAST
   call get_command_argument(3, arg)
   if (trim(arg) == 'exact') mode = FDTD_MODE_EXACT
   loopkind = 0
   call get_command_argument(4, arg)
   if (trim(arg) == 'split') loopkind = 1
   if (trim(arg) == 'graph') loopkind = 2
   if (trim(arg) == 'tables') loopkind = 3

   ! ---- case file -----------------------------------------------------------
   open (newunit=u, file=trim(casefile), access='stream', form='unformatted', &
         status='old', action='read', iostat=ios)
   if (ios /= 0) call die('cannot open '//trim(casefile))
   read (u) magic, version
   if (magic /= 'SFDC' .or. version /= 3) call die('not a version-3 case file')
   read (u) nx, ny, nz, nmedia, nsteps
   nn = [nx, ny, nz]
   read (u) eps0, mu0, dt
   allocate (epr(0:nmedia), sigma(0:nmedia), mur(0:nmedia), sigmam(0:nmedia), &
             kind(0:nmedia), npoles(0:nmedia), eps11(0:nmedia), sig11(0:nmedia))
   allocate (pole_c(2, PMAX, 0:nmedia), pole_a(2, PMAX, 0:nmedia))
   read (u) epr, sigma, mur, sigmam, kind, npoles, pole_c, pole_a, eps11, sig11
   allocate (mnpoles(0:nmedia), mpole_c(2, PMAX, 0:nmedia), mpole_a(2, PMAX, 0:nmedia), &
             mu11(0:nmedia), sigm11(0:nmedia))
   read (u) mnpoles, mpole_c, mpole_a, mu11, sigm11, murf
   do ax = 1, 3
      allocate (dd(ax)%v(-2:nn(ax)+1))
      read (u) dd(ax)%v
   end do
   read (u) layers, order, refl, alphamax, alphaorden, kappamax
   do c = 0, 5
      do ax = 1, 3
         hi(ax) = sweep_hi(c, ax - 1)
      end do
      ! the Alloc box of launch_simulation (timestepping.F90:380-386)
      allocate (mi(c)%a(-1:hi(1)+1, -1:hi(2)+1, -1:hi(3)+1))
      read (u) mi(c)%a
   end do

   ! ---- launch_simulation set-up, reference order -----------------------------
   allocate (g1(0:nmedia), g2(0:nmedia), gm1(0:nmedia), gm2(0:nmedia))
   call check_host(fdtd_build_g_tables(nmedia, epr, sigma, mur, sigmam, kind, dt, eps0, mu0, &
                                       g1, g2, gm1, gm2), 'fdtd_build_g_tables')
   has_cpml = any(layers /= 0)
   do ax = 1, 3
      n = nn(ax)
      allocate (de(ax)%v(-1:n+1), dh(ax)%v(-1:n+1), ide(ax)%v(-1:n+1), idh(ax)%v(-1:n+1))
      allocate (pbe(ax)%v(-1:n+1), pce(ax)%v(-1:n+1), pbm(ax)%v(-1:n+1), pcm(ax)%v(-1:n+1))
      call check_host(fdtd_build_spacing(n, dd(ax)%v, de(ax)%v, dh(ax)%v, ide(ax)%v, idh(ax)%v), &
                      'fdtd_build_spacing')
      pbe(ax)%v = 1.0; pce(ax)%v = 0.0; pbm(ax)%v = 1.0; pcm(ax)%v = 0.0
      ! InitCPMLBorders builds all three axes once any face has layers
      if (has_cpml) then
         call check_host(fdtd_build_cpml_axis(n, layers(2*ax-1), layers(2*ax), order(2*ax-1), &
                 order(2*ax), refl(2*ax-1), refl(2*ax), alphamax, alphaorden, kappamax, dt, eps0, &
                 mu0, de(ax)%v, dh(ax)%v, pbe(ax)%v, pce(ax)%v, pbm(ax)%v, pcm(ax)%v, &
                 ide(ax)%v, idh(ax)%v), 'fdtd_build_cpml_axis')
      end if
   end do
   nd = count(iand(kind(1:), FDTD_KIND_EDISP) /= 0)
   allocate (dmed(max(nd, 1)), dnp(max(nd, 1)))
   allocate (kappa(2, PMAX, max(nd, 1)), beta(2, PMAX, max(nd, 1)), g3(2, PMAX, max(nd, 1)))
   kappa = 0; beta = 0; g3 = 0
   d = 0
   do m = 1, nmedia
      if (iand(kind(m), FDTD_KIND_EDISP) == 0) cycle
      d = d + 1
      kk = 0; bb = 0; gg = 0
      call check_host(fdtd_build_ade(npoles(m), pole_c(:, :, m), pole_a(:, :, m), eps11(m), &
                      sig11(m), dt, kk, bb, gg, a1, a2), 'fdtd_build_ade')
      dmed(d) = m; dnp(d) = npoles(m)
      kappa(:, :, d) = kk; beta(:, :, d) = bb; g3(:, :, d) = gg
      ! InitEDispersives overrides G1/G2 of the medium (electricdispersive.F90:140-146)
      g1(m) = a1(1); g2(m) = a2(1)
   end do
   ! InitMDispersives (timestepping.F90:943): the same tables from mu11 and
   ! Sigmam11, overriding GM1/GM2 (magneticdispersive.F90:122-135)
   nmd = count(iand(kind(1:), FDTD_KIND_MDISP) /= 0)
   allocate (mdmed(max(nmd, 1)), mdnp(max(nmd, 1)))
   allocate (mkappa(2, PMAX, max(nmd, 1)), mbeta(2, PMAX, max(nmd, 1)), gm3(2, PMAX, max(nmd, 1)))
   mkappa = 0; mbeta = 0; gm3 = 0
   d = 0
   do m = 1, nmedia
      if (iand(kind(m), FDTD_KIND_MDISP) == 0) cycle
      d = d + 1
      kk = 0; bb = 0; gg = 0
      call check_host(fdtd_build_ade(mnpoles(m), mpole_c(:, :, m), mpole_a(:, :, m), mu11(m), &
                      sigm11(m), dt, kk, bb, gg, a1, a2), 'fdtd_build_ade')
      mdmed(d) = m; mdnp(d) = mnpoles(m)
      mkappa(:, :, d) = kk; mbeta(:, :, d) = bb; gm3(:, :, d) = gg
      gm1(m) = a1(1); gm2(m) = a2(1)
   end do
   ! InitMURBorders (timestepping.F90:669): dual inverse spacings at the Alloc
   ! ends after the CPML stretching
   allocate (cab1(6 * (nmedia + 1)))
   cab1 = 0
   if (any(murf /= 0)) then
      idhf = [idh(1)%v(-1), idh(1)%v(nx), idh(2)%v(-1), idh(2)%v(ny), idh(3)%v(-1), idh(3)%v(nz)]
      call check_host(fdtd_build_mur(nmedia, epr, mur, dt, eps0, mu0, idhf, cab1), 'fdtd_build_mur')
   end if
   allocate (tiempo(0:nsteps+2))
   tiempo(0) = 0.0d0
   do i = 1, nsteps + 2
      tiempo(i) = tiempo(i-1) + dt
   end do

   if (loopkind == 3) then
      open (newunit=u, file=trim(outfile), access='stream', form='unformatted', &
            status='replace', action='write')
      write (u) g1, g2, gm1, gm2
      do ax = 1, 3
         write (u) ide(ax)%v, idh(ax)%v, pbe(ax)%v, pce(ax)%v, pbm(ax)%v, pcm(ax)%v
      end do
      write (u) nd, kappa, beta, g3
      write (u) nmd, mkappa, mbeta, gm3
      write (u) cab1
      close (u)
      stop
   end if

   ! ---- device state ---------------------------------------------------------
   desc%nx = nx; desc%ny = ny; desc%nz = nz
   desc%mode = mode; desc%storage = FDTD_STORE_FP32; desc%device = 0
   desc%stream = c_null_ptr
   h = c_null_ptr
   if (fdtd_create(desc, h) /= FDTD_OK) call die('fdtd_create: '//trim(fdtd_error_message(c_null_ptr)))
   call check(fdtd_set_media(h, nmedia, g1, g2, gm1, gm2, kind), 'fdtd_set_media')
   do c = 0, 5
      call check(fdtd_set_media_index(h, c, mi(c)%a), 'fdtd_set_media_index')
   end do
   ! Idxh/Idyh/Idzh span Alloc(iEx..iEz) = [-1, n]
   call check(fdtd_set_spacing(h, ide(1)%v, ide(2)%v, ide(3)%v, idh(1)%v(-1:nx), &
                               idh(2)%v(-1:ny), idh(3)%v(-1:nz)), 'fdtd_set_spacing')
   if (has_cpml) then
      call check(fdtd_set_cpml(h, layers, pbe(1)%v, pce(1)%v, pbm(1)%v, pcm(1)%v, &
                               pbe(2)%v, pce(2)%v, pbm(2)%v, pcm(2)%v, &
                               pbe(3)%v, pce(3)%v, pbm(3)%v, pcm(3)%v), 'fdtd_set_cpml')
   end if
   if (nd > 0) then
      call check(fdtd_set_edispersive(h, nd, dmed, dnp, kappa, beta, g3), 'fdtd_set_edispersive')
   end if
   if (nmd > 0) then
      call check(fdtd_set_mdispersive(h, nmd, mdmed, mdnp, mkappa, mbeta, gm3), &
                 'fdtd_set_mdispersive')
   end if
   if (any(murf /= 0)) call check(fdtd_set_mur(h, murf, cab1), 'fdtd_set_mur')
   call check(fdtd_set_time(h, dt, nsteps + 2, tiempo), 'fdtd_set_time')
   read (u) nnodal
   do s = 1, nnodal
      read (u) ncomp, nhard, nbox, amp, deltaevol, numus
      allocate (evol(0:numus))
      read (u) evol
      call check(fdtd_add_nodal(h, ncomp, nhard, nbox, amp, deltaevol, numus, evol), &
                 'fdtd_add_nodal')
      deallocate (evol)
   end do
   read (u) nprobe
   allocate (pc(max(nprobe, 1)), pi(max(nprobe, 1)), pj(max(nprobe, 1)), pk(max(nprobe, 1)))
   if (nprobe > 0) read (u) pc(1:nprobe), pi(1:nprobe), pj(1:nprobe), pk(1:nprobe)
   ! TF/SF plane waves (InitPlaneWave, planewaves.F90:72) placed on the
   ! discretisation lines of the preprocessing (preprocess_geom.F90:5672-5712)
   read (u) npw
   if (npw > 0) then
      do ax = 1, 3
         allocate (lines(ax)%v(-1:nn(ax)+1))
         call check_host(fdtd_build_lines(nn(ax), layers(2*ax-1), layers(2*ax), dd(ax)%v, &
                                          lines(ax)%v), 'fdtd_build_lines')
      end do
      call check(fdtd_set_lines(h, lines(1)%v, lines(2)%v, lines(3)%v), 'fdtd_set_lines')
   end if
   do s = 1, npw
      read (u) pwbox, nmodes
      allocate (pwp(3*nmodes), pwe(3*nmodes), pwinc(nmodes))
      read (u) pwp, pwe, pwinc, deltaevol, numus
      allocate (evol(0:numus))
      read (u) evol
      call check(fdtd_add_planewave(h, pwbox, nmodes, pwp, pwe, pwinc, deltaevol, numus, evol, &
                                    eps0, mu0), 'fdtd_add_planewave')
      deallocate (pwp, pwe, pwinc, evol)
   end do
   close (u)
   bufsteps = max(1, min(nsteps, 4096))
   if (nprobe > 0) then
      call check(fdtd_set_probes(h, nprobe, pc, pi, pj, pk, bufsteps), 'fdtd_set_probes')
   end if
   allocate (samples(max(nprobe, 1), max(nsteps, 1)), chunk(max(nprobe, 1), bufsteps))
   nsamp = 0

   ! ---- ciclo_temporal --------------------------------------------------------
   if (loopkind == 2) then
      left = nsteps; n0 = 0
      do while (left > 0)
         k = min(left, bufsteps)
         call check(fdtd_run(h, n0, k), 'fdtd_run')
         call drain()
         n0 = n0 + k; left = left - k
      end do
   else
      do n = 0, nsteps - 1
         if (loopkind == 1) then
            ! Advance_Ex/Ey/Ez (timestepping.F90:1286-1297)
            call check(fdtd_advance_e_yee(h), 'fdtd_advance_e_yee')
            ! AdvanceelectricCPML (timestepping.F90:1381)
            if (has_cpml) call check(fdtd_advance_e_cpml(h), 'fdtd_advance_e_cpml')
            ! AdvanceEDispersiveE (timestepping.F90:1416)
            if (nd > 0) call check(fdtd_advance_e_dispersive(h), 'fdtd_advance_e_dispersive')
            ! AdvancePlaneWaveE (timestepping.F90:1425)
            if (npw > 0) call check(fdtd_advance_planewave_e(h, n), 'fdtd_advance_planewave_e')
            ! AdvanceNodalE (timestepping.F90:1438-1446)
            call check(fdtd_advance_nodal_e(h, n), 'fdtd_advance_nodal_e')
            ! Advance_Hx/Hy/Hz (timestepping.F90:1497-1507)
            call check(fdtd_advance_h_yee(h), 'fdtd_advance_h_yee')
            ! AdvanceMagneticCPML (timestepping.F90:1535)
            if (has_cpml) call check(fdtd_advance_h_cpml(h), 'fdtd_advance_h_cpml')
            ! AdvanceMDispersiveH (timestepping.F90:1586)
            if (nmd > 0) call check(fdtd_advance_h_dispersive(h), 'fdtd_advance_h_dispersive')
            ! AdvancePlaneWaveH (timestepping.F90:1598)
            if (npw > 0) call check(fdtd_advance_planewave_h(h, n), 'fdtd_advance_planewave_h')
            ! AdvanceNodalH (timestepping.F90:1596-1635)
            call check(fdtd_advance_nodal_h(h, n), 'fdtd_advance_nodal_h')
            ! AdvanceMagneticMUR (timestepping.F90:1711-1712)
            if (any(murf /= 0)) call check(fdtd_advance_mur(h), 'fdtd_advance_mur')
         else
            call check(fdtd_advance_e(h, n), 'fdtd_advance_e')
            call check(fdtd_advance_h(h, n), 'fdtd_advance_h')
         end if
         ! UpdateObservation (timestepping.F90:1725-1742)
         if (nprobe > 0) then
            call check(fdtd_sample_probes(h), 'fdtd_sample_probes')
            if (mod(n + 1, bufsteps) == 0) call drain()
         end if
      end do
      call drain()
   end if
   call check(fdtd_synchronize(h), 'fdtd_synchronize')

   open (newunit=u, file=trim(outfile), access='stream', form='unformatted', &
         status='replace', action='write')
   write (u) nsamp, nprobe
   if (nprobe > 0 .and. nsamp > 0) write (u) samples(1:nprobe, 1:nsamp)
   close (u)
   if (fdtd_destroy(h) /= FDTD_OK) call die('fdtd_destroy failed')
   write (*, '(a,i0,a,i0,a,i0,a,i0,a)') 'fdtd_gpu_driver: ', nx, 'x', ny, 'x', nz, ' cells, ', &
      nsamp, ' steps'

contains

   integer(c_int32_t) function sweep_hi(comp, a)
      ! last swept index of component comp along axis a (total coordinates)
      integer, intent(in) :: comp, a
      logical :: node
      if (comp < 3) then
         node = comp /= a
      else
         node = comp - 3 == a
      end if
      sweep_hi = nn(a + 1)
      if (.not. node) sweep_hi = nn(a + 1) - 1
   end function

   subroutine drain()
      if (nprobe == 0) return
      call check(fdtd_read_probes(h, chunk, bufsteps, got), 'fdtd_read_probes')
      if (nsamp + got > size(samples, 2)) call die('more probe samples than steps')
      samples(1:nprobe, nsamp+1:nsamp+got) = chunk(1:nprobe, 1:got)
      nsamp = nsamp + got
   end subroutine

   subroutine check(rc, what)
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: what
      if (rc /= FDTD_OK) call die(what//': '//trim(fdtd_error_message(h)))
   end subroutine

   subroutine check_host(rc, what)
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: what
      if (rc /= FDTD_OK) call die(what//' failed')
   end subroutine

   subroutine die(msg)
      character(len=*), intent(in) :: msg
      write (0, '(a)') 'fdtd_gpu_driver: '//trim(msg)
      error stop 1
   end subroutine

end program fdtd_gpu_driver
