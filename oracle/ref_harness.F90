! ref_harness.F90 -- TEST INFRASTRUCTURE ONLY.
!
! A driver of the reference's own Fortran modules, built from their sources
! where they lie (fdetypes, errorreport, calc_constants, borderscpml,
! electricdispersive, nodalsources, planewaves, magneticdispersive, bordersmur;
! see oracle/Makefile), so that
! the tests can compare the C restatement in fdtd_oracle.c against the
! reference itself:
!   - calc_G1G2Gm1Gm2          calc_constants.F90:36-148
!   - InitCPMLBorders, AdvanceelectricCPML, AdvanceMagneticCPML
!                              borderscpml.F90:90-473, 584-1176, 1465-1733
!   - InitEDispersives, AdvanceEDispersiveE
!                              electricdispersive.F90:77-362
!   - InitnodalSources, AdvanceNodalE/H
!                              nodalsources.F90:76-305, 396-777
!   - InitPlaneWave, AdvancePlaneWaveE/H
!                              planewaves.F90:72-770, 878-1475
!   - InitMdispersives, AdvanceMdispersiveH
!                              magneticdispersive.F90:77-362
!   - InitMURBorders, AdvanceMagneticMUR
!                              bordersmur.F90:79-363, 502-1600
! The six Yee loops live in timestepping.F90, which cannot be compiled here
! without stand-ins for the MTLN modules it uses, so they are restated below
! from timestepping.F90:2165-2409 (one statement each).  Everything else that
! touches the fields runs the reference's code.
!
! Coordinates: total cells (0 = outer face of the low CPML slab), arrays
! allocated with the reference's Alloc bounds (fullsize +-1,
! semba_fdtd.F90:1285-1300).  Exchange with C uses the uniform layout
! ((k+1)*(ny+3)+(j+1))*(nx+3)+(i+1).
module ref_harness
   use iso_c_binding
   use fdetypes
   use report
   use calc_constants
   use borders_cpml
   use edispersives
   use nodalsources
   use ilumina
   use Mdispersives
   use Borders_MUR
   implicit none

   type(SGGFDTDINFO), save, target :: sgg
   type(bounds_t), save :: b
   integer(kind=4), save :: gnx, gny, gnz
   real(kind=RKIND), pointer, dimension(:,:,:), save :: Ex, Ey, Ez, Hx, Hy, Hz
   integer(kind=INTEGERSIZEOFMEDIAMATRICES), pointer, dimension(:,:,:), save :: &
      miEx, miEy, miEz, miHx, miHy, miHz
   real(kind=RKIND), pointer, dimension(:), save :: G1, G2, GM1, GM2
   real(kind=RKIND), allocatable, dimension(:), save :: dxe, dye, dze, dxh, dyh, dzh, &
      Idxe, Idye, Idze, Idxh, Idyh, Idzh
   logical, save :: thereCPML, thereEdisp, thereNodE, thereNodH
   logical, save :: therePW = .false., stillPW = .true.
   logical, save :: thereMdisp = .false., thereMur = .false.
   real(kind=RKIND), save :: eps0h, mu0h
   type(limit_t), dimension(1:6), save :: sinpml

contains

   subroutine set_limits(field, xi, xe, yi, ye, zi, ze, lim)
      integer, intent(in) :: field, xi, xe, yi, ye, zi, ze
      type(XYZlimit_t), dimension(1:6), intent(inout) :: lim
      lim(field)%XI = xi; lim(field)%XE = xe
      lim(field)%YI = yi; lim(field)%YE = ye
      lim(field)%ZI = zi; lim(field)%ZE = ze
   end subroutine

   ! findbounds, restated from timestepping.F90:2883-3169 (bounds of every
   ! array = Alloc, sweeps = Sweep, spacing bounds = Alloc of the matching field)
   subroutine fill_bounds()
      call one(b%Ex, sgg%Alloc(iEx)); call one(b%Ey, sgg%Alloc(iEy)); call one(b%Ez, sgg%Alloc(iEz))
      call one(b%Hx, sgg%Alloc(iHx)); call one(b%Hy, sgg%Alloc(iHy)); call one(b%Hz, sgg%Alloc(iHz))
      call one(b%sggMiEx, sgg%Alloc(iEx)); call one(b%sggMiEy, sgg%Alloc(iEy)); call one(b%sggMiEz, sgg%Alloc(iEz))
      call one(b%sggMiHx, sgg%Alloc(iHx)); call one(b%sggMiHy, sgg%Alloc(iHy)); call one(b%sggMiHz, sgg%Alloc(iHz))
      call one(b%sweepEx, sgg%Sweep(iEx)); call one(b%sweepEy, sgg%Sweep(iEy)); call one(b%sweepEz, sgg%Sweep(iEz))
      call one(b%sweepHx, sgg%Sweep(iHx)); call one(b%sweepHy, sgg%Sweep(iHy)); call one(b%sweepHz, sgg%Sweep(iHz))
      b%dxe%XI = sgg%Alloc(iHx)%XI; b%dxe%XE = sgg%Alloc(iHx)%XE; b%dxe%NX = b%dxe%XE - b%dxe%XI + 1
      b%dye%YI = sgg%Alloc(iHy)%YI; b%dye%YE = sgg%Alloc(iHy)%YE; b%dye%NY = b%dye%YE - b%dye%YI + 1
      b%dze%ZI = sgg%Alloc(iHz)%ZI; b%dze%ZE = sgg%Alloc(iHz)%ZE; b%dze%NZ = b%dze%ZE - b%dze%ZI + 1
      b%dxh%XI = sgg%Alloc(iEx)%XI; b%dxh%XE = sgg%Alloc(iEx)%XE; b%dxh%NX = b%dxh%XE - b%dxh%XI + 1
      b%dyh%YI = sgg%Alloc(iEy)%YI; b%dyh%YE = sgg%Alloc(iEy)%YE; b%dyh%NY = b%dyh%YE - b%dyh%YI + 1
      b%dzh%ZI = sgg%Alloc(iEz)%ZI; b%dzh%ZE = sgg%Alloc(iEz)%ZE; b%dzh%NZ = b%dzh%ZE - b%dzh%ZI + 1
   contains
      subroutine one(l, a)
         type(limit_t), intent(out) :: l
         type(XYZlimit_t), intent(in) :: a
         l%XI = a%XI; l%XE = a%XE; l%YI = a%YI; l%YE = a%YE; l%ZI = a%ZI; l%ZE = a%ZE
         l%NX = a%XE - a%XI + 1; l%NY = a%YE - a%YI + 1; l%NZ = a%ZE - a%ZI + 1
      end subroutine
   end subroutine

   integer(c_size_t) function uidx(i, j, k)
      integer, intent(in) :: i, j, k
      uidx = (int(k + 1, c_size_t) * int(gny + 3, c_size_t) + int(j + 1, c_size_t)) &
             * int(gnx + 3, c_size_t) + int(i + 1, c_size_t) + 1_c_size_t
   end function

   subroutine ref_init(nx, ny, nz, nmedia, epr, sigma, mur, sigmam, kind, &
                       npoles, pc, pa, eps11, sig11, lay, orden, refl, &
                       alphamax, alphaorden, kappamax, dt, dxa, dya, dza, &
                       mi, g_eps0, g_mu0) bind(C, name="ref_init")
      integer(c_int), value :: nx, ny, nz, nmedia
      real(c_float), intent(in) :: epr(0:nmedia), sigma(0:nmedia), mur(0:nmedia), sigmam(0:nmedia)
      integer(c_int), intent(in) :: kind(0:nmedia), npoles(0:nmedia)
      real(c_double), intent(in) :: pc(2, 8, 0:nmedia), pa(2, 8, 0:nmedia)
      real(c_float), intent(in) :: eps11(0:nmedia), sig11(0:nmedia)
      integer(c_int), intent(in) :: lay(2, 3)
      real(c_float), intent(in) :: orden(2, 3), refl(2, 3)
      real(c_float), value :: alphamax, alphaorden, kappamax, g_eps0, g_mu0
      real(c_double), value :: dt
      real(c_float), intent(in) :: dxa(-2:nx+1), dya(-2:ny+1), dza(-2:nz+1)
      integer(c_int16_t), intent(in) :: mi(*)
      integer :: f, r, i, j, k, p, o
      integer(c_size_t) :: nall
      logical :: resume

      gnx = nx; gny = ny; gnz = nz
      eps0h = g_eps0; mu0h = g_mu0
      nullify(sgg%tiempo)
      sgg%dt = dt
      sgg%NumMedia = nmedia
      ! fullsize (preprocess_geom.F90:5563-5591) in total coordinates
      call set_limits(iEx, 0, nx-1, 0, ny, 0, nz, sgg%Sweep)
      call set_limits(iEy, 0, nx, 0, ny-1, 0, nz, sgg%Sweep)
      call set_limits(iEz, 0, nx, 0, ny, 0, nz-1, sgg%Sweep)
      call set_limits(iHx, 0, nx, 0, ny-1, 0, nz-1, sgg%Sweep)
      call set_limits(iHy, 0, nx-1, 0, ny, 0, nz-1, sgg%Sweep)
      call set_limits(iHz, 0, nx-1, 0, ny-1, 0, nz, sgg%Sweep)
      do f = 1, 6
         sgg%Alloc(f)%XI = sgg%Sweep(f)%XI - 1; sgg%Alloc(f)%XE = sgg%Sweep(f)%XE + 1
         sgg%Alloc(f)%YI = sgg%Sweep(f)%YI - 1; sgg%Alloc(f)%YE = sgg%Sweep(f)%YE + 1
         sgg%Alloc(f)%ZI = sgg%Sweep(f)%ZI - 1; sgg%Alloc(f)%ZE = sgg%Sweep(f)%ZE + 1
         ! SINPML_fullsize: the sweep without the CPML slabs
         sinpml(f)%XI = sgg%Sweep(f)%XI + lay(1, 1); sinpml(f)%XE = sgg%Sweep(f)%XE - lay(2, 1)
         sinpml(f)%YI = sgg%Sweep(f)%YI + lay(1, 2); sinpml(f)%YE = sgg%Sweep(f)%YE - lay(2, 2)
         sinpml(f)%ZI = sgg%Sweep(f)%ZI + lay(1, 3); sinpml(f)%ZE = sgg%Sweep(f)%ZE - lay(2, 3)
         sgg%SINPMLSweep(f)%XI = sinpml(f)%XI; sgg%SINPMLSweep(f)%XE = sinpml(f)%XE
         sgg%SINPMLSweep(f)%YI = sinpml(f)%YI; sgg%SINPMLSweep(f)%YE = sinpml(f)%YE
         sgg%SINPMLSweep(f)%ZI = sinpml(f)%ZI; sgg%SINPMLSweep(f)%ZE = sinpml(f)%ZE
      end do
      call fill_bounds()

      allocate(sgg%DX(-2:nx+1), sgg%DY(-2:ny+1), sgg%DZ(-2:nz+1))
      sgg%DX = dxa; sgg%DY = dya; sgg%DZ = dza

      ! spacings, timestepping.F90:314-370
      allocate(dxe(-1:nx+1), dye(-1:ny+1), dze(-1:nz+1), Idxe(-1:nx+1), Idye(-1:ny+1), Idze(-1:nz+1))
      allocate(dxh(-1:nx), dyh(-1:ny), dzh(-1:nz), Idxh(-1:nx), Idyh(-1:ny), Idzh(-1:nz))
      do i = -1, nx+1; dxe(i) = sgg%DX(i); end do
      do j = -1, ny+1; dye(j) = sgg%DY(j); end do
      do k = -1, nz+1; dze(k) = sgg%DZ(k); end do
      do i = -1, nx; dxh(i) = (sgg%DX(i) + sgg%DX(i-1)) / 2.0_RKIND; end do
      do j = -1, ny; dyh(j) = (sgg%DY(j) + sgg%DY(j-1)) / 2.0_RKIND; end do
      do k = -1, nz; dzh(k) = (sgg%DZ(k) + sgg%DZ(k-1)) / 2.0_RKIND; end do
      Idxe = 1.0_RKIND/dxe; Idye = 1.0_RKIND/dye; Idze = 1.0_RKIND/dze
      Idxh = 1.0_RKIND/dxh; Idyh = 1.0_RKIND/dyh; Idzh = 1.0_RKIND/dzh

      ! media
      allocate(sgg%Med(0:nmedia))
      do r = 0, nmedia
         sgg%Med(r)%Epr = epr(r); sgg%Med(r)%Sigma = sigma(r)
         sgg%Med(r)%Mur = mur(r); sgg%Med(r)%SigmaM = sigmam(r)
         sgg%Med(r)%Priority = 0.0_RKIND
         sgg%Med(r)%Is = Exists_t(.false., .false., .false., .false., .false., .false., .false., .false., &
            .false., .false., .false., .false., .false., .false., .false., .false., .false., .false., &
            .false., .false., .false., .false., .false., .false., .false., .false., .false.)
         sgg%Med(r)%Is%PEC = (iand(kind(r), 1) /= 0) .or. (r == 0)
         sgg%Med(r)%Is%PMC = (iand(kind(r), 2) /= 0)
         sgg%Med(r)%Is%EDispersive = (iand(kind(r), 4) /= 0)
         sgg%Med(r)%Is%MDispersive = (iand(kind(r), 8) /= 0)
         if (sgg%Med(r)%Is%EDispersive) then
            allocate(sgg%Med(r)%EDispersive(1))
            sgg%Med(r)%EDispersive(1)%NumPolRes11 = npoles(r)
            allocate(sgg%Med(r)%EDispersive(1)%C11(npoles(r)), sgg%Med(r)%EDispersive(1)%A11(npoles(r)))
            do p = 1, npoles(r)
               sgg%Med(r)%EDispersive(1)%C11(p) = cmplx(pc(1, p, r), pc(2, p, r), kind=CKIND)
               sgg%Med(r)%EDispersive(1)%A11(p) = cmplx(pa(1, p, r), pa(2, p, r), kind=CKIND)
            end do
            sgg%Med(r)%EDispersive(1)%eps11 = eps11(r)
            sgg%Med(r)%EDispersive(1)%Sigma11 = sig11(r)
         end if
      end do

      ! fields and media matrices with the reference's Alloc bounds
      call alloc3(Ex, iEx); call alloc3(Ey, iEy); call alloc3(Ez, iEz)
      call alloc3(Hx, iHx); call alloc3(Hy, iHy); call alloc3(Hz, iHz)
      call allocm(miEx, iEx, 0); call allocm(miEy, iEy, 1); call allocm(miEz, iEz, 2)
      call allocm(miHx, iHx, 3); call allocm(miHy, iHy, 4); call allocm(miHz, iHz, 5)

      allocate(G1(0:nmedia), G2(0:nmedia), GM1(0:nmedia), GM2(0:nmedia))
      call calc_G1G2Gm1Gm2(sgg, G1, G2, GM1, GM2, eps0h, mu0h)

      ! CPML: borderscpml.F90:90
      sgg%PML%NumLayers = 0
      do o = 1, 3
         do p = 1, 2
            sgg%PML%NumLayers(o, p) = lay(p, o)
            sgg%PML%orden(o, p) = orden(p, o)
            sgg%PML%CoeffReflPML(o, p) = refl(p, o)
         end do
      end do
      sgg%Border%IsBackMUR = .false.; sgg%Border%IsFrontMUR = .false.
      sgg%Border%IsLeftMUR = .false.; sgg%Border%IsRightMUR = .false.
      sgg%Border%IsDownMUR = .false.; sgg%Border%IsUpMUR = .false.
      sgg%Border%IsBackPML = lay(1, 1) > 0; sgg%Border%IsFrontPML = lay(2, 1) > 0
      sgg%Border%IsLeftPML = lay(1, 2) > 0; sgg%Border%IsRightPML = lay(2, 2) > 0
      sgg%Border%IsDownPML = lay(1, 3) > 0; sgg%Border%IsUpPML = lay(2, 3) > 0
      resume = .false.
      call InitCPMLBorders(sgg, 0, sinpml, thereCPML, resume, dxe, dye, dze, dxh, dyh, dzh, &
                           Idxe, Idye, Idze, Idxh, Idyh, Idzh, alphamax, alphaorden, kappamax, &
                           eps0h, mu0h, .false.)
      call InitEDispersives(sgg, miEx, miEy, miEz, thereEdisp, resume, G1, G2, Ex, Ey, Ez)
      thereNodE = .false.; thereNodH = .false.
      nall = 0
   contains
      subroutine alloc3(a, f)
         real(kind=RKIND), pointer, dimension(:,:,:) :: a
         integer, intent(in) :: f
         allocate(a(sgg%Alloc(f)%XI:sgg%Alloc(f)%XE, sgg%Alloc(f)%YI:sgg%Alloc(f)%YE, &
                    sgg%Alloc(f)%ZI:sgg%Alloc(f)%ZE))
         a = 0.0_RKIND
      end subroutine
      subroutine allocm(a, f, slot)
         integer(kind=INTEGERSIZEOFMEDIAMATRICES), pointer, dimension(:,:,:) :: a
         integer, intent(in) :: f, slot
         integer :: ii, jj, kk
         integer(c_size_t) :: base
         allocate(a(sgg%Alloc(f)%XI:sgg%Alloc(f)%XE, sgg%Alloc(f)%YI:sgg%Alloc(f)%YE, &
                    sgg%Alloc(f)%ZI:sgg%Alloc(f)%ZE))
         base = int(slot, c_size_t) * int(nx + 3, c_size_t) * int(ny + 3, c_size_t) * int(nz + 3, c_size_t)
         do kk = sgg%Alloc(f)%ZI, sgg%Alloc(f)%ZE
            do jj = sgg%Alloc(f)%YI, sgg%Alloc(f)%YE
               do ii = sgg%Alloc(f)%XI, sgg%Alloc(f)%XE
                  a(ii, jj, kk) = mi(base + uidx(ii, jj, kk))
               end do
            end do
         end do
      end subroutine
   end subroutine ref_init

   ! Nodal sources through the reference's InitnodalSources (nodalsources.F90:76)
   subroutine ref_nodal(nsrc, comp, hard, box, amp, delta, numus, samples, nsteps) bind(C, name="ref_nodal")
      integer(c_int), value :: nsrc, nsteps
      integer(c_int), intent(in) :: comp(nsrc), hard(nsrc), box(6, nsrc), numus(nsrc)
      real(c_float), intent(in) :: amp(nsrc), delta(nsrc)
      real(c_float), intent(in) :: samples(*)
      integer :: s, off, i
      sgg%NumNodalSources = nsrc
      allocate(sgg%NodalSource(nsrc))
      off = 0
      do s = 1, nsrc
         sgg%NodalSource(s)%numpuntos = 1
         allocate(sgg%NodalSource(s)%punto(1))
         sgg%NodalSource(s)%punto(1)%XI = box(1, s); sgg%NodalSource(s)%punto(1)%XE = box(2, s)
         sgg%NodalSource(s)%punto(1)%YI = box(3, s); sgg%NodalSource(s)%punto(1)%YE = box(4, s)
         sgg%NodalSource(s)%punto(1)%ZI = box(5, s); sgg%NodalSource(s)%punto(1)%ZE = box(6, s)
         sgg%NodalSource(s)%punto(1)%xc = 0.0_RKIND
         sgg%NodalSource(s)%punto(1)%yc = 0.0_RKIND
         sgg%NodalSource(s)%punto(1)%zc = 0.0_RKIND
         select case (mod(comp(s), 3))
         case (0); sgg%NodalSource(s)%punto(1)%xc = amp(s)
         case (1); sgg%NodalSource(s)%punto(1)%yc = amp(s)
         case (2); sgg%NodalSource(s)%punto(1)%zc = amp(s)
         end select
         sgg%NodalSource(s)%IsElec = comp(s) < 3
         sgg%NodalSource(s)%IsHard = hard(s) /= 0
         sgg%NodalSource(s)%IsInitialValue = .false.
         sgg%NodalSource(s)%Fichero%NumSamples = numus(s)
         sgg%NodalSource(s)%Fichero%DeltaSamples = delta(s)
         allocate(sgg%NodalSource(s)%Fichero%Samples(0:numus(s)))
         do i = 0, numus(s)
            sgg%NodalSource(s)%Fichero%Samples(i) = samples(off + i + 1)
         end do
         off = off + numus(s) + 1
      end do
      call InitnodalSources(sgg, 0, nsrc, sgg%NodalSource, sgg%Sweep, thereNodE, thereNodH)
      call ref_time(nsteps)
   end subroutine ref_nodal

   ! Magnetic pole-residue media through the reference's InitMdispersives
   ! (magneticdispersive.F90:77), after InitEDispersives as in
   ! timestepping.F90:927-943: per medium its pole count, residues/poles
   ! (2, 8, 0:nmedia) and mu11 / Sigmam11.
   subroutine ref_mdisp(nmedia, npoles, pc, pa, mu11, sigm11) bind(C, name="ref_mdisp")
      integer(c_int), value :: nmedia
      integer(c_int), intent(in) :: npoles(0:nmedia)
      real(c_double), intent(in) :: pc(2, 8, 0:nmedia), pa(2, 8, 0:nmedia)
      real(c_float), intent(in) :: mu11(0:nmedia), sigm11(0:nmedia)
      integer :: r, p
      logical :: resume
      do r = 0, nmedia
         if (.not. sgg%Med(r)%Is%MDispersive) cycle
         allocate(sgg%Med(r)%MDispersive(1))
         sgg%Med(r)%MDispersive(1)%NumPolRes11 = npoles(r)
         allocate(sgg%Med(r)%MDispersive(1)%C11(npoles(r)), sgg%Med(r)%MDispersive(1)%A11(npoles(r)))
         do p = 1, npoles(r)
            sgg%Med(r)%MDispersive(1)%C11(p) = cmplx(pc(1, p, r), pc(2, p, r), kind=CKIND)
            sgg%Med(r)%MDispersive(1)%A11(p) = cmplx(pa(1, p, r), pa(2, p, r), kind=CKIND)
         end do
         sgg%Med(r)%MDispersive(1)%mu11 = mu11(r)
         sgg%Med(r)%MDispersive(1)%Sigmam11 = sigm11(r)
      end do
      resume = .false.
      call InitMdispersives(sgg, miHx, miHy, miHz, thereMdisp, resume, GM1, GM2, Hx, Hy, Hz)
   end subroutine ref_mdisp

   ! First-order Mur faces through the reference's InitMURBorders
   ! (bordersmur.F90:79), faces = back, front, left, right, down, up.
   subroutine ref_mur(faces) bind(C, name="ref_mur")
      integer(c_int), intent(in) :: faces(6)
      logical :: resume
      sgg%Border%IsBackMUR = faces(1) /= 0; sgg%Border%IsFrontMUR = faces(2) /= 0
      sgg%Border%IsLeftMUR = faces(3) /= 0; sgg%Border%IsRightMUR = faces(4) /= 0
      sgg%Border%IsDownMUR = faces(5) /= 0; sgg%Border%IsUpMUR = faces(6) /= 0
      resume = .false.
      call InitMURBorders(sgg, thereMur, resume, Idxh, Idyh, Idzh, eps0h, mu0h)
   end subroutine ref_mur

   ! crea_timevector, timestepping.F90:3230-3240 (once)
   subroutine ref_time(nsteps) bind(C, name="ref_time")
      integer(c_int), value :: nsteps
      integer :: i
      if (associated(sgg%tiempo)) return
      allocate(sgg%tiempo(0:nsteps+2))
      sgg%tiempo(0) = 0.0_RKIND_tiempo
      do i = 1, nsteps + 2
         sgg%tiempo(i) = sgg%tiempo(i-1) + sgg%dt
      end do
   end subroutine ref_time

   ! One TF/SF plane wave through the reference's InitPlaneWave
   ! (planewaves.F90:72): box esq = x1,x2,y1,y2,z1,z2, modes p(3,m), e(3,m),
   ! incert(m), excitation samples and the node coordinates of each axis over
   ! [-1, n+1] (sgg%LineX/Y/Z).
   subroutine ref_planewave(esq, nmodes, p, e, incert, delta, numus, samples, lx, ly, lz) &
         bind(C, name="ref_planewave")
      integer(c_int), intent(in) :: esq(6)
      integer(c_int), value :: nmodes, numus
      real(c_float), intent(in) :: p(3, nmodes), e(3, nmodes), incert(nmodes)
      real(c_float), value :: delta
      real(c_float), intent(in) :: samples(0:numus)
      real(c_float), intent(in) :: lx(-1:gnx+1), ly(-1:gny+1), lz(-1:gnz+1)
      logical :: resume
      sgg%NumPlaneWaves = 1
      allocate(sgg%PlaneWave(1))
      sgg%PlaneWave(1)%esqx1 = esq(1); sgg%PlaneWave(1)%esqx2 = esq(2)
      sgg%PlaneWave(1)%esqy1 = esq(3); sgg%PlaneWave(1)%esqy2 = esq(4)
      sgg%PlaneWave(1)%esqz1 = esq(5); sgg%PlaneWave(1)%esqz2 = esq(6)
      sgg%PlaneWave(1)%nummodes = nmodes
      sgg%PlaneWave(1)%isRC = .false.
      sgg%PlaneWave(1)%INCERTMAX = 0.0_RKIND
      allocate(sgg%PlaneWave(1)%px(nmodes), sgg%PlaneWave(1)%py(nmodes), sgg%PlaneWave(1)%pz(nmodes), &
               sgg%PlaneWave(1)%ex(nmodes), sgg%PlaneWave(1)%ey(nmodes), sgg%PlaneWave(1)%ez(nmodes), &
               sgg%PlaneWave(1)%incert(nmodes))
      sgg%PlaneWave(1)%px = p(1, :); sgg%PlaneWave(1)%py = p(2, :); sgg%PlaneWave(1)%pz = p(3, :)
      sgg%PlaneWave(1)%ex = e(1, :); sgg%PlaneWave(1)%ey = e(2, :); sgg%PlaneWave(1)%ez = e(3, :)
      sgg%PlaneWave(1)%incert = incert
      sgg%PlaneWave(1)%Fichero%Name = 'harness'
      sgg%PlaneWave(1)%Fichero%NumSamples = numus
      sgg%PlaneWave(1)%Fichero%DeltaSamples = delta
      allocate(sgg%PlaneWave(1)%Fichero%Samples(0:numus))
      sgg%PlaneWave(1)%Fichero%Samples = samples
      allocate(sgg%LineX(-1:gnx+1), sgg%LineY(-1:gny+1), sgg%LineZ(-1:gnz+1))
      sgg%LineX = lx; sgg%LineY = ly; sgg%LineZ = lz
      resume = .false.
      call InitPlaneWave(sgg, miEx, miEy, miEz, miHx, miHy, miHz, 0, 1, sinpml, therePW, resume, &
                         eps0h, mu0h)
      stillPW = .true.
   end subroutine ref_planewave

   ! Yee loops restated from timestepping.F90:2187-2197, 2228-2236, 2272-2280,
   ! 2316-2325, 2356-2364, 2396-2404 (global indices instead of offsets)
   subroutine yee_e()
      integer :: i, j, k, m
!$OMP PARALLEL DO DEFAULT(SHARED) COLLAPSE(2) PRIVATE(i, j, k, m)
      do k = sgg%Sweep(iEx)%ZI, sgg%Sweep(iEx)%ZE
         do j = sgg%Sweep(iEx)%YI, sgg%Sweep(iEx)%YE
            do i = sgg%Sweep(iEx)%XI, sgg%Sweep(iEx)%XE
               m = miEx(i, j, k)
               Ex(i,j,k) = G1(m)*Ex(i,j,k) + G2(m)*((Hz(i,j,k)-Hz(i,j-1,k))*Idyh(j) - (Hy(i,j,k)-Hy(i,j,k-1))*Idzh(k))
            end do
         end do
      end do
!$OMP END PARALLEL DO
!$OMP PARALLEL DO DEFAULT(SHARED) COLLAPSE(2) PRIVATE(i, j, k, m)
      do k = sgg%Sweep(iEy)%ZI, sgg%Sweep(iEy)%ZE
         do j = sgg%Sweep(iEy)%YI, sgg%Sweep(iEy)%YE
            do i = sgg%Sweep(iEy)%XI, sgg%Sweep(iEy)%XE
               m = miEy(i, j, k)
               Ey(i,j,k) = G1(m)*Ey(i,j,k) + G2(m)*((Hx(i,j,k)-Hx(i,j,k-1))*Idzh(k) - (Hz(i,j,k)-Hz(i-1,j,k))*Idxh(i))
            end do
         end do
      end do
!$OMP END PARALLEL DO
!$OMP PARALLEL DO DEFAULT(SHARED) COLLAPSE(2) PRIVATE(i, j, k, m)
      do k = sgg%Sweep(iEz)%ZI, sgg%Sweep(iEz)%ZE
         do j = sgg%Sweep(iEz)%YI, sgg%Sweep(iEz)%YE
            do i = sgg%Sweep(iEz)%XI, sgg%Sweep(iEz)%XE
               m = miEz(i, j, k)
               Ez(i,j,k) = G1(m)*Ez(i,j,k) + G2(m)*((Hy(i,j,k)-Hy(i-1,j,k))*Idxh(i) - (Hx(i,j,k)-Hx(i,j-1,k))*Idyh(j))
            end do
         end do
      end do
!$OMP END PARALLEL DO
   end subroutine

   subroutine yee_h()
      integer :: i, j, k, m
!$OMP PARALLEL DO DEFAULT(SHARED) COLLAPSE(2) PRIVATE(i, j, k, m)
      do k = sgg%Sweep(iHx)%ZI, sgg%Sweep(iHx)%ZE
         do j = sgg%Sweep(iHx)%YI, sgg%Sweep(iHx)%YE
            do i = sgg%Sweep(iHx)%XI, sgg%Sweep(iHx)%XE
               m = miHx(i, j, k)
               Hx(i,j,k) = GM1(m)*Hx(i,j,k) + GM2(m)*((Ey(i,j,k+1)-Ey(i,j,k))*Idze(k) - (Ez(i,j+1,k)-Ez(i,j,k))*Idye(j))
            end do
         end do
      end do
!$OMP END PARALLEL DO
!$OMP PARALLEL DO DEFAULT(SHARED) COLLAPSE(2) PRIVATE(i, j, k, m)
      do k = sgg%Sweep(iHy)%ZI, sgg%Sweep(iHy)%ZE
         do j = sgg%Sweep(iHy)%YI, sgg%Sweep(iHy)%YE
            do i = sgg%Sweep(iHy)%XI, sgg%Sweep(iHy)%XE
               m = miHy(i, j, k)
               Hy(i,j,k) = GM1(m)*Hy(i,j,k) + GM2(m)*((Ez(i+1,j,k)-Ez(i,j,k))*Idxe(i) - (Ex(i,j,k+1)-Ex(i,j,k))*Idze(k))
            end do
         end do
      end do
!$OMP END PARALLEL DO
!$OMP PARALLEL DO DEFAULT(SHARED) COLLAPSE(2) PRIVATE(i, j, k, m)
      do k = sgg%Sweep(iHz)%ZI, sgg%Sweep(iHz)%ZE
         do j = sgg%Sweep(iHz)%YI, sgg%Sweep(iHz)%YE
            do i = sgg%Sweep(iHz)%XI, sgg%Sweep(iHz)%XE
               m = miHz(i, j, k)
               Hz(i,j,k) = GM1(m)*Hz(i,j,k) + GM2(m)*((Ex(i,j+1,k)-Ex(i,j,k))*Idye(j) - (Ey(i+1,j,k)-Ey(i,j,k))*Idxe(i))
            end do
         end do
      end do
!$OMP END PARALLEL DO
   end subroutine

   ! One step in the order of ciclo_temporal (timestepping.F90:1248-1981)
   subroutine ref_step(n) bind(C, name="ref_step")
      integer(c_int), value :: n
      call yee_e()
      if (thereCPML) call AdvanceelectricCPML(sgg%NumMedia, b, miEx, miEy, miEz, G2, Ex, Ey, Ez, Hx, Hy, Hz)
      if (thereEdisp) call AdvanceEDispersiveE(sgg)
      if (therePW .and. stillPW) call AdvancePlaneWaveE(sgg, n, b, G2, Idxh, Idyh, Idzh, Ex, Ey, Ez, stillPW)
      if (thereNodE) call AdvanceNodalE(sgg, miEx, miEy, miEz, sgg%NumMedia, n, b, G2, Idxh, Idyh, Idzh, &
                                        Ex, Ey, Ez, .false.)
      call yee_h()
      if (thereCPML) call AdvanceMagneticCPML(sgg%NumMedia, b, miHx, miHy, miHz, GM2, Hx, Hy, Hz, Ex, Ey, Ez)
      if (thereMdisp) call AdvanceMdispersiveH(sgg)
      if (therePW .and. stillPW) call AdvancePlaneWaveH(sgg, n, b, GM2, Idxe, Idye, Idze, Hx, Hy, Hz, stillPW)
      if (thereNodH) call AdvanceNodalH(sgg, miHx, miHy, miHz, sgg%NumMedia, n, b, GM2, Idxe, Idye, Idze, &
                                        Hx, Hy, Hz, .false.)
      if (thereMur) call AdvanceMagneticMUR(b, sgg, miHx, miHy, miHz, Hx, Hy, Hz, .false.)
   end subroutine

   subroutine ref_get_field(c, out) bind(C, name="ref_get_field")
      integer(c_int), value :: c
      real(c_float), intent(out) :: out(*)
      select case (c)
      case (0); call copyout(Ex)
      case (1); call copyout(Ey)
      case (2); call copyout(Ez)
      case (3); call copyout(Hx)
      case (4); call copyout(Hy)
      case (5); call copyout(Hz)
      end select
   contains
      subroutine copyout(a)
         real(kind=RKIND), pointer, dimension(:,:,:) :: a
         integer :: ii, jj, kk
         do kk = lbound(a, 3), ubound(a, 3)
            do jj = lbound(a, 2), ubound(a, 2)
               do ii = lbound(a, 1), ubound(a, 1)
                  out(uidx(ii, jj, kk)) = a(ii, jj, kk)
               end do
            end do
         end do
      end subroutine
   end subroutine

   ! G tables after InitEDispersives and the CPML-stretched inverse spacings
   subroutine ref_get_tables(g1o, g2o, gm1o, gm2o, ixe, iye, ize, ixh, iyh, izh) bind(C, name="ref_get_tables")
      real(c_float), intent(out) :: g1o(0:sgg%NumMedia), g2o(0:sgg%NumMedia), &
         gm1o(0:sgg%NumMedia), gm2o(0:sgg%NumMedia)
      real(c_float), intent(out) :: ixe(-1:gnx+1), iye(-1:gny+1), ize(-1:gnz+1)
      real(c_float), intent(out) :: ixh(-1:gnx+1), iyh(-1:gny+1), izh(-1:gnz+1)
      g1o = G1; g2o = G2; gm1o = GM1; gm2o = GM2
      ixe = Idxe; iye = Idye; ize = Idze
      ixh = 0.0; iyh = 0.0; izh = 0.0
      ixh(-1:gnx) = Idxh; iyh(-1:gny) = Idyh; izh(-1:gnz) = Idzh
   end subroutine

end module ref_harness
