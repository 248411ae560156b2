! sembafdtd_gpu.F90 -- iso_c_binding interface to libsembafdtd.so (the C ABI
! of include/sembafdtd.h) for the semba-fdtd Fortran code.
!
! Every entry point takes the reference's own arrays unchanged: the field and
! media-index matrices as allocated in launch_simulation
! (timestepping.F90:380-386; sggMiEx(sgg%Alloc(iEx)%XI:...) etc.), the G
! tables (0:NumMedia), the inverse spacings and CPML profiles over their Alloc
! ranges.  Status codes: 0 = OK, negative = error (fdtd_error_message).
module sembafdtd_gpu
   use iso_c_binding
   implicit none
   private

   integer(c_int), parameter, public :: FDTD_OK = 0
   integer(c_int), parameter, public :: FDTD_MODE_FAST = 0, FDTD_MODE_EXACT = 1
   integer(c_int), parameter, public :: FDTD_STORE_FP32 = 0, FDTD_STORE_FP16 = 1
   integer(c_int), parameter, public :: FDTD_EX = 0, FDTD_EY = 1, FDTD_EZ = 2, &
                                        FDTD_HX = 3, FDTD_HY = 4, FDTD_HZ = 5
   integer(c_int), parameter, public :: FDTD_KIND_PEC = 1, FDTD_KIND_PMC = 2, &
                                        FDTD_KIND_EDISP = 4, FDTD_KIND_MDISP = 8

   type, bind(C), public :: fdtd_grid_desc
      integer(c_int32_t) :: nx, ny, nz      ! swept cells per axis, CPML included
      integer(c_int32_t) :: mode, storage, device
      type(c_ptr) :: stream = c_null_ptr
   end type

   public :: fdtd_create, fdtd_destroy, fdtd_set_media, fdtd_set_media_index, &
             fdtd_set_spacing, fdtd_set_cpml, fdtd_set_edispersive, fdtd_set_time, &
             fdtd_set_lines, fdtd_add_planewave, fdtd_advance_planewave_e, &
             fdtd_advance_planewave_h, fdtd_build_lines, fdtd_add_nodal, fdtd_advance_e, &
             fdtd_advance_h, fdtd_advance_e_yee, &
             fdtd_advance_e_cpml, fdtd_advance_e_dispersive, fdtd_advance_nodal_e, &
             fdtd_advance_h_yee, fdtd_advance_h_cpml, fdtd_advance_nodal_h, fdtd_run, &
             fdtd_set_probes, fdtd_read_probes, fdtd_sample_probes, fdtd_get_field, fdtd_get_field_noalloc, &
             fdtd_set_field, fdtd_synchronize, fdtd_courant_dt, fdtd_build_g_tables, &
             fdtd_build_spacing, fdtd_build_cpml_axis, fdtd_build_ade, fdtd_evolucion, &
             fdtd_abi_version, fdtd_device_bytes, fdtd_device_memory, fdtd_dispersive_edges, &
             fdtd_media_fill, &
             fdtd_media_pec_walls, fdtd_media_paint_sphere, fdtd_media_paint_box, &
             fdtd_set_mdispersive, fdtd_advance_h_dispersive, fdtd_set_mur, fdtd_build_mur, &
             fdtd_advance_mur, fdtd_set_tiling, fdtd_get_tiling, fdtd_time_halves, &
             fdtd_add_dft, fdtd_dft_size, fdtd_read_dft, fdtd_update_observation, &
             fdtd_set_field_scale, fdtd_state_size, fdtd_save_state, fdtd_load_state, &
             fdtd_batch_create, fdtd_batch_destroy, fdtd_batch_admit, fdtd_batch_retire, &
             fdtd_batch_step_of, fdtd_batch_set_tiling, fdtd_batch_set_step_mode, &
             fdtd_batch_step, &
             fdtd_batch_synchronize, fdtd_set_step_mode, fdtd_get_step_mode, fdtd_time_step, &
             fdtd_step_classes, fdtd_set_step_class_rows, fdtd_get_step_class_rows, &
             fdtd_set_step_vector, fdtd_set_graph_steps, &
             fdtd_error_message

   interface
      integer(c_int) function fdtd_create(desc, h) bind(C, name="fdtd_create")
         import :: c_int, c_ptr, fdtd_grid_desc
         type(fdtd_grid_desc), intent(in) :: desc
         type(c_ptr), intent(out) :: h
      end function
      integer(c_int) function fdtd_destroy(h) bind(C, name="fdtd_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function
      type(c_ptr) function fdtd_last_error_c(h) bind(C, name="fdtd_last_error")
         import :: c_ptr
         type(c_ptr), value :: h
      end function
      ! calc_G1G2Gm1Gm2 output (calc_constants.F90:36-148)
      integer(c_int) function fdtd_set_media(h, nmedia, g1, g2, gm1, gm2, kind) &
            bind(C, name="fdtd_set_media")
         import :: c_int, c_ptr, c_int32_t, c_float
         type(c_ptr), value :: h
         integer(c_int32_t), value :: nmedia
         real(c_float), intent(in) :: g1(*), g2(*), gm1(*), gm2(*)
         integer(c_int32_t), intent(in) :: kind(*)
      end function
      ! sggMiEx..sggMiHz (semba_fdtd.F90:93)
      integer(c_int) function fdtd_set_media_index(h, comp, mi) &
            bind(C, name="fdtd_set_media_index")
         import :: c_int, c_ptr, c_int32_t, c_int16_t
         type(c_ptr), value :: h
         integer(c_int32_t), value :: comp
         integer(c_int16_t), intent(in) :: mi(*)
      end function
      integer(c_int) function fdtd_set_spacing(h, idxe, idye, idze, idxh, idyh, idzh) &
            bind(C, name="fdtd_set_spacing")
         import :: c_int, c_ptr, c_float
         type(c_ptr), value :: h
         real(c_float), intent(in) :: idxe(*), idye(*), idze(*), idxh(*), idyh(*), idzh(*)
      end function
      ! InitCPMLBorders result (borderscpml.F90:90-473)
      integer(c_int) function fdtd_set_cpml(h, layers, pbe_x, pce_x, pbm_x, pcm_x, &
            pbe_y, pce_y, pbm_y, pcm_y, pbe_z, pce_z, pbm_z, pcm_z) bind(C, name="fdtd_set_cpml")
         import :: c_int, c_ptr, c_int32_t, c_float
         type(c_ptr), value :: h
         integer(c_int32_t), intent(in) :: layers(6)
         real(c_float), intent(in) :: pbe_x(*), pce_x(*), pbm_x(*), pcm_x(*), &
            pbe_y(*), pce_y(*), pbm_y(*), pcm_y(*), pbe_z(*), pce_z(*), pbm_z(*), pcm_z(*)
      end function
      ! InitEDispersives result (electricdispersive.F90:77-298)
      integer(c_int) function fdtd_set_edispersive(h, ndisp, medium, npoles, kappa, beta, g3) &
            bind(C, name="fdtd_set_edispersive")
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: h
         integer(c_int32_t), value :: ndisp
         integer(c_int32_t), intent(in) :: medium(*), npoles(*)
         real(c_double), intent(in) :: kappa(*), beta(*), g3(*)
      end function
      integer(c_int) function fdtd_set_time(h, dt, nlast, tiempo) bind(C, name="fdtd_set_time")
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: h
         real(c_double), value :: dt
         integer(c_int32_t), value :: nlast
         real(c_double), intent(in) :: tiempo(*)
      end function
      ! sgg%LineX/Y/Z over [-1, n+1] (preprocess_geom.F90:5672-5712)
      integer(c_int) function fdtd_set_lines(h, line_x, line_y, line_z) bind(C, name="fdtd_set_lines")
         import :: c_int, c_ptr, c_float
         type(c_ptr), value :: h
         real(c_float), intent(in) :: line_x(*), line_y(*), line_z(*)
      end function
      ! sgg%PlaneWave(j) (InitPlaneWave, planewaves.F90:72-770)
      integer(c_int) function fdtd_add_planewave(h, box, nmodes, p, e, incert, deltaevol, numus, &
            evol, eps0, mu0) bind(C, name="fdtd_add_planewave")
         import :: c_int, c_ptr, c_int32_t, c_float
         type(c_ptr), value :: h
         integer(c_int32_t), intent(in) :: box(6)
         integer(c_int32_t), value :: nmodes, numus
         real(c_float), intent(in) :: p(*), e(*), incert(*), evol(*)
         real(c_float), value :: deltaevol, eps0, mu0
      end function
      integer(c_int) function fdtd_add_nodal(h, comp, hard, box, amp, deltaevol, numus, evol) &
            bind(C, name="fdtd_add_nodal")
         import :: c_int, c_ptr, c_int32_t, c_float
         type(c_ptr), value :: h
         integer(c_int32_t), value :: comp, hard, numus
         integer(c_int32_t), intent(in) :: box(6)
         real(c_float), value :: amp, deltaevol
         real(c_float), intent(in) :: evol(*)
      end function
      ! ---- the call sites of ciclo_temporal (timestepping.F90:1248-1981)
      integer(c_int) function fdtd_advance_e(h, n) bind(C, name="fdtd_advance_e")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: h
         integer(c_int32_t), value :: n
      end function
      integer(c_int) function fdtd_advance_h(h, n) bind(C, name="fdtd_advance_h")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: h
         integer(c_int32_t), value :: n
      end function
      integer(c_int) function fdtd_advance_e_yee(h) bind(C, name="fdtd_advance_e_yee")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function
      integer(c_int) function fdtd_advance_e_cpml(h) bind(C, name="fdtd_advance_e_cpml")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function
      integer(c_int) function fdtd_advance_e_dispersive(h) bind(C, name="fdtd_advance_e_dispersive")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function
      integer(c_int) function fdtd_advance_planewave_e(h, n) bind(C, name="fdtd_advance_planewave_e")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: h
         integer(c_int32_t), value :: n
      end function
      integer(c_int) function fdtd_advance_planewave_h(h, n) bind(C, name="fdtd_advance_planewave_h")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: h
         integer(c_int32_t), value :: n
      end function
      integer(c_int) function fdtd_advance_nodal_e(h, n) bind(C, name="fdtd_advance_nodal_e")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: h
         integer(c_int32_t), value :: n
      end function
      integer(c_int) function fdtd_advance_h_yee(h) bind(C, name="fdtd_advance_h_yee")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function
      integer(c_int) function fdtd_advance_h_cpml(h) bind(C, name="fdtd_advance_h_cpml")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function
      integer(c_int) function fdtd_advance_nodal_h(h, n) bind(C, name="fdtd_advance_nodal_h")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: h
         integer(c_int32_t), value :: n
      end function
      integer(c_int) function fdtd_run(h, n0, nsteps) bind(C, name="fdtd_run")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: h
         integer(c_int32_t), value :: n0, nsteps
      end function
      ! ---- observation (observation.F90:2620-2672)
      integer(c_int) function fdtd_set_probes(h, nprobe, comp, i, j, k, bufsteps) &
            bind(C, name="fdtd_set_probes")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: h
         integer(c_int32_t), value :: nprobe, bufsteps
         integer(c_int32_t), intent(in) :: comp(*), i(*), j(*), k(*)
      end function
      integer(c_int) function fdtd_read_probes(h, out, max_samples, nsamples) &
            bind(C, name="fdtd_read_probes")
         import :: c_int, c_ptr, c_int32_t, c_float
         type(c_ptr), value :: h
         real(c_float), intent(out) :: out(*)
         integer(c_int32_t), value :: max_samples
         integer(c_int32_t), intent(out) :: nsamples
      end function
      integer(c_int) function fdtd_sample_probes(h) bind(C, name="fdtd_sample_probes")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function
      integer(c_int) function fdtd_get_field(h, comp, f) bind(C, name="fdtd_get_field")
         import :: c_int, c_ptr, c_int32_t, c_float
         type(c_ptr), value :: h
         integer(c_int32_t), value :: comp
         real(c_float), intent(out) :: f(*)
      end function
      integer(c_int) function fdtd_get_field_noalloc(h, comp, f) &
            bind(C, name="fdtd_get_field_noalloc")
         import :: c_int, c_ptr, c_int32_t, c_float
         type(c_ptr), value :: h
         integer(c_int32_t), value :: comp
         real(c_float), intent(out) :: f(*)
      end function
      integer(c_int) function fdtd_set_field(h, comp, f) bind(C, name="fdtd_set_field")
         import :: c_int, c_ptr, c_int32_t, c_float
         type(c_ptr), value :: h
         integer(c_int32_t), value :: comp
         real(c_float), intent(in) :: f(*)
      end function
      integer(c_int) function fdtd_synchronize(h) bind(C, name="fdtd_synchronize")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function
      ! ---- table builders (the reference's expressions, real(4) kinds)
      real(c_double) function fdtd_courant_dt(dxmin, dymin, dzmin, eps0, mu0, cfl, forcecfl, &
            dt_in) bind(C, name="fdtd_courant_dt")
         import :: c_double, c_float, c_int32_t
         real(c_float), value :: dxmin, dymin, dzmin, eps0, mu0, cfl
         integer(c_int32_t), value :: forcecfl
         real(c_double), value :: dt_in
      end function
      integer(c_int) function fdtd_build_g_tables(nmedia, epr, sigma, mur, sigmam, kind, dt, &
            eps0, mu0, g1, g2, gm1, gm2) bind(C, name="fdtd_build_g_tables")
         import :: c_int, c_int32_t, c_float, c_double
         integer(c_int32_t), value :: nmedia
         real(c_float), intent(in) :: epr(*), sigma(*), mur(*), sigmam(*)
         integer(c_int32_t), intent(in) :: kind(*)
         real(c_double), value :: dt
         real(c_float), value :: eps0, mu0
         real(c_float), intent(out) :: g1(*), g2(*), gm1(*), gm2(*)
      end function
      integer(c_int) function fdtd_build_spacing(n, d, de, dh, ide, idh) &
            bind(C, name="fdtd_build_spacing")
         import :: c_int, c_int32_t, c_float
         integer(c_int32_t), value :: n
         real(c_float), intent(in) :: d(*)
         real(c_float), intent(out) :: de(*), dh(*), ide(*), idh(*)
      end function
      integer(c_int) function fdtd_build_cpml_axis(n, llo, lhi, orden_lo, orden_hi, r_lo, r_hi, &
            alphamaxpar, alphaorden, kappamaxpar, dt, eps0, mu0, de, dh, pbe, pce, pbm, pcm, &
            ide, idh) bind(C, name="fdtd_build_cpml_axis")
         import :: c_int, c_int32_t, c_float, c_double
         integer(c_int32_t), value :: n, llo, lhi
         real(c_float), value :: orden_lo, orden_hi, r_lo, r_hi, alphamaxpar, alphaorden, &
            kappamaxpar, eps0, mu0
         real(c_double), value :: dt
         real(c_float), intent(in) :: de(*), dh(*)
         real(c_float), intent(out) :: pbe(*), pce(*), pbm(*), pcm(*)
         real(c_float), intent(inout) :: ide(*), idh(*)
      end function
      integer(c_int) function fdtd_build_lines(n, llo, lhi, d, line) bind(C, name="fdtd_build_lines")
         import :: c_int, c_int32_t, c_float
         integer(c_int32_t), value :: n, llo, lhi
         real(c_float), intent(in) :: d(*)
         real(c_float), intent(out) :: line(*)
      end function
      integer(c_int) function fdtd_build_ade(npoles, c, a, eps11, sigma11, dt, kappa, beta, g3, &
            g1, g2) bind(C, name="fdtd_build_ade")
         import :: c_int, c_int32_t, c_float, c_double
         integer(c_int32_t), value :: npoles
         real(c_double), intent(in) :: c(*), a(*)
         real(c_float), value :: eps11, sigma11
         real(c_double), value :: dt
         real(c_double), intent(out) :: kappa(*), beta(*), g3(*)
         real(c_float), intent(out) :: g1(*), g2(*)
      end function
      real(c_float) function fdtd_evolucion(t, deltaevol, numus, evol) bind(C, name="fdtd_evolucion")
         import :: c_float, c_int32_t
         real(c_float), value :: t, deltaevol
         integer(c_int32_t), value :: numus
         real(c_float), intent(in) :: evol(*)
      end function
      integer(c_int) function fdtd_abi_version() bind(C, name="fdtd_abi_version")
         import :: c_int
      end function
      integer(c_int64_t) function fdtd_device_bytes(h) bind(C, name="fdtd_device_bytes")
         import :: c_int64_t, c_ptr
         type(c_ptr), value :: h
      end function
      integer(c_int) function fdtd_device_memory(device, free_bytes, total_bytes) &
            bind(C, name="fdtd_device_memory")
         import :: c_int, c_int32_t, c_int64_t
         integer(c_int32_t), value :: device
         integer(c_int64_t), intent(out) :: free_bytes, total_bytes
      end function
      integer(c_int64_t) function fdtd_dispersive_edges(h, comp) bind(C, name="fdtd_dispersive_edges")
         import :: c_int64_t, c_ptr, c_int32_t
         type(c_ptr), value :: h
         integer(c_int32_t), value :: comp
      end function
      ! ---- device-side scene building (healer.F90 rules)
      integer(c_int) function fdtd_media_fill(h, medium) bind(C, name="fdtd_media_fill")
         import :: c_int, c_ptr, c_int16_t
         type(c_ptr), value :: h
         integer(c_int16_t), value :: medium
      end function
      integer(c_int) function fdtd_media_pec_walls(h, faces) bind(C, name="fdtd_media_pec_walls")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: h
         integer(c_int32_t), value :: faces
      end function
      integer(c_int) function fdtd_media_paint_sphere(h, cx, cy, cz, radius, medium, nprio, prio) &
            bind(C, name="fdtd_media_paint_sphere")
         import :: c_int, c_ptr, c_float, c_int16_t, c_int32_t
         type(c_ptr), value :: h
         real(c_float), value :: cx, cy, cz, radius
         integer(c_int16_t), value :: medium
         integer(c_int32_t), value :: nprio
         integer(c_int32_t), intent(in) :: prio(*)
      end function
      integer(c_int) function fdtd_media_paint_box(h, x0, x1, y0, y1, z0, z1, medium, nprio, prio) &
            bind(C, name="fdtd_media_paint_box")
         import :: c_int, c_ptr, c_int16_t, c_int32_t
         type(c_ptr), value :: h
         integer(c_int32_t), value :: x0, x1, y0, y1, z0, z1
         integer(c_int16_t), value :: medium
         integer(c_int32_t), value :: nprio
         integer(c_int32_t), intent(in) :: prio(*)
      end function
      ! ---- magnetic pole-residue media (magneticdispersive.F90:77-362)
      integer(c_int) function fdtd_set_mdispersive(h, ndisp, medium, npoles, kappa, beta, gm3) &
            bind(C, name="fdtd_set_mdispersive")
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: h
         integer(c_int32_t), value :: ndisp
         integer(c_int32_t), intent(in) :: medium(*), npoles(*)
         real(c_double), intent(in) :: kappa(*), beta(*), gm3(*)
      end function
      integer(c_int) function fdtd_advance_h_dispersive(h) bind(C, name="fdtd_advance_h_dispersive")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function
      ! ---- first-order Mur faces (bordersmur.F90:79-363, 502-1600)
      integer(c_int) function fdtd_set_mur(h, faces, cab1) bind(C, name="fdtd_set_mur")
         import :: c_int, c_ptr, c_int32_t, c_float
         type(c_ptr), value :: h
         integer(c_int32_t), intent(in) :: faces(6)
         real(c_float), intent(in) :: cab1(*)
      end function
      integer(c_int) function fdtd_build_mur(nmedia, epr, mur, dt, eps0, mu0, idh_face, cab1) &
            bind(C, name="fdtd_build_mur")
         import :: c_int, c_int32_t, c_float, c_double
         integer(c_int32_t), value :: nmedia
         real(c_float), intent(in) :: epr(*), mur(*)
         real(c_double), value :: dt
         real(c_float), value :: eps0, mu0
         real(c_float), intent(in) :: idh_face(6)
         real(c_float), intent(out) :: cab1(*)
      end function
      integer(c_int) function fdtd_advance_mur(h) bind(C, name="fdtd_advance_mur")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function
      ! ---- work decomposition of the face passes and its timing
      integer(c_int) function fdtd_set_tiling(h, rows, zchunk, xcells) bind(C, name="fdtd_set_tiling")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: h
         integer(c_int32_t), value :: rows, zchunk, xcells
      end function
      integer(c_int) function fdtd_get_tiling(h, rows, zchunk, xcells) bind(C, name="fdtd_get_tiling")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: h
         integer(c_int32_t), intent(out) :: rows, zchunk, xcells
      end function
      integer(c_int) function fdtd_time_halves(h, reps, ms) bind(C, name="fdtd_time_halves")
         import :: c_int, c_ptr, c_int32_t, c_float
         type(c_ptr), value :: h
         integer(c_int32_t), value :: reps
         real(c_float), intent(out) :: ms(3)
      end function
      ! ---- steps per graph replay of fdtd_run
      integer(c_int) function fdtd_set_graph_steps(h, steps) bind(C, name="fdtd_set_graph_steps")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: h
         integer(c_int32_t), value :: steps
      end function
      ! ---- one launch per step in fdtd_run
      integer(c_int) function fdtd_set_step_mode(h, rows, zchunk) bind(C, name="fdtd_set_step_mode")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: h
         integer(c_int32_t), value :: rows, zchunk
      end function
      integer(c_int) function fdtd_get_step_mode(h, rows, zchunk) bind(C, name="fdtd_get_step_mode")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: h
         integer(c_int32_t), intent(out) :: rows, zchunk
      end function
      integer(c_int) function fdtd_set_step_class_rows(h, classes, rows) &
            bind(C, name="fdtd_set_step_class_rows")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: h
         integer(c_int32_t), value :: classes, rows
      end function
      integer(c_int) function fdtd_get_step_class_rows(h, rows) bind(C, name="fdtd_get_step_class_rows")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: h
         integer(c_int32_t), intent(out) :: rows(9)
      end function
      integer(c_int) function fdtd_set_step_vector(h, on) bind(C, name="fdtd_set_step_vector")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: h
         integer(c_int32_t), value :: on
      end function
      integer(c_int) function fdtd_time_step(h, reps, ms) bind(C, name="fdtd_time_step")
         import :: c_int, c_ptr, c_int32_t, c_float
         type(c_ptr), value :: h
         integer(c_int32_t), value :: reps
         real(c_float), intent(out) :: ms
      end function
      integer(c_int) function fdtd_step_classes(h, reps, cls, cells, bytes, ms) &
            bind(C, name="fdtd_step_classes")
         import :: c_int, c_ptr, c_int32_t, c_float, c_double
         type(c_ptr), value :: h
         integer(c_int32_t), value :: reps
         integer(c_int32_t), intent(out) :: cls(9)
         real(c_double), intent(out) :: cells(9)
         real(c_double), intent(out) :: bytes(9)
         real(c_float), intent(out) :: ms(9)
      end function
      ! ---- frequency-domain observation (observation.F90:3527-3600)
      integer(c_int) function fdtd_add_dft(h, kind, box, stride, nfreq, freq, id) &
            bind(C, name="fdtd_add_dft")
         import :: c_int, c_ptr, c_int32_t, c_float
         type(c_ptr), value :: h
         integer(c_int32_t), value :: kind, nfreq
         integer(c_int32_t), intent(in) :: box(6), stride(3)
         real(c_float), intent(in) :: freq(*)
         integer(c_int32_t), intent(out) :: id
      end function
      integer(c_int64_t) function fdtd_dft_size(h, id) bind(C, name="fdtd_dft_size")
         import :: c_int64_t, c_ptr, c_int32_t
         type(c_ptr), value :: h
         integer(c_int32_t), value :: id
      end function
      integer(c_int) function fdtd_read_dft(h, id, out, capacity) bind(C, name="fdtd_read_dft")
         import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
         type(c_ptr), value :: h
         integer(c_int32_t), value :: id
         real(c_double), intent(out) :: out(*)
         integer(c_int64_t), value :: capacity
      end function
      integer(c_int) function fdtd_update_observation(h, n) bind(C, name="fdtd_update_observation")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: h
         integer(c_int32_t), value :: n
      end function
      ! ---- binary16 storage scale, checkpoint / resume (resuming.F90:394-496)
      integer(c_int) function fdtd_set_field_scale(h, comp, scale) bind(C, name="fdtd_set_field_scale")
         import :: c_int, c_ptr, c_int32_t, c_float
         type(c_ptr), value :: h
         integer(c_int32_t), value :: comp
         real(c_float), value :: scale
      end function
      integer(c_int64_t) function fdtd_state_size(h) bind(C, name="fdtd_state_size")
         import :: c_int64_t, c_ptr
         type(c_ptr), value :: h
      end function
      integer(c_int) function fdtd_save_state(h, n, buf, size) bind(C, name="fdtd_save_state")
         import :: c_int, c_ptr, c_int32_t, c_int64_t
         type(c_ptr), value :: h, buf
         integer(c_int32_t), value :: n
         integer(c_int64_t), value :: size
      end function
      integer(c_int) function fdtd_load_state(h, buf, size, n) bind(C, name="fdtd_load_state")
         import :: c_int, c_ptr, c_int32_t, c_int64_t
         type(c_ptr), value :: h, buf
         integer(c_int64_t), value :: size
         integer(c_int32_t), intent(out) :: n
      end function
      ! ---- batches of independent grids
      integer(c_int) function fdtd_batch_create(device, stream, b) bind(C, name="fdtd_batch_create")
         import :: c_int, c_ptr, c_int32_t
         integer(c_int32_t), value :: device
         type(c_ptr), value :: stream
         type(c_ptr), intent(out) :: b
      end function
      integer(c_int) function fdtd_batch_destroy(b) bind(C, name="fdtd_batch_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: b
      end function
      integer(c_int) function fdtd_batch_admit(b, h, n0, slot) bind(C, name="fdtd_batch_admit")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: b, h
         integer(c_int32_t), value :: n0
         integer(c_int32_t), intent(out) :: slot
      end function
      integer(c_int) function fdtd_batch_retire(b, slot) bind(C, name="fdtd_batch_retire")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: b
         integer(c_int32_t), value :: slot
      end function
      integer(c_int64_t) function fdtd_batch_step_of(b, slot) bind(C, name="fdtd_batch_step_of")
         import :: c_int64_t, c_ptr, c_int32_t
         type(c_ptr), value :: b
         integer(c_int32_t), value :: slot
      end function
      integer(c_int) function fdtd_batch_set_tiling(b, rows, zchunk, xcells, blocks) &
            bind(C, name="fdtd_batch_set_tiling")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: b
         integer(c_int32_t), value :: rows, zchunk, xcells, blocks
      end function
      integer(c_int) function fdtd_batch_set_step_mode(b, rows, zchunk) &
            bind(C, name="fdtd_batch_set_step_mode")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: b
         integer(c_int32_t), value :: rows, zchunk
      end function
      integer(c_int) function fdtd_batch_step(b, nsteps) bind(C, name="fdtd_batch_step")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: b
         integer(c_int32_t), value :: nsteps
      end function
      integer(c_int) function fdtd_batch_synchronize(b) bind(C, name="fdtd_batch_synchronize")
         import :: c_int, c_ptr
         type(c_ptr), value :: b
      end function
   end interface

contains

   ! fdtd_last_error as a Fortran string
   function fdtd_error_message(h) result(msg)
      type(c_ptr), intent(in) :: h
      character(len=512) :: msg
      character(kind=c_char), pointer :: p(:)
      integer :: i
      msg = ' '
      call c_f_pointer(fdtd_last_error_c(h), p, [512])
      do i = 1, 512
         if (p(i) == c_null_char) exit
         msg(i:i) = p(i)
      end do
   end function

end module sembafdtd_gpu
