! Fortran host of the MI355X shallow-water path: the counterpart of the
! reference's tmcore_swm_mod (reference src/core/tmcore_swm_mod.F90:29-90)
! with the time loop on the device through libtmgpu.
!
!   tm_swm_init   params_parse_namelist + mesh + poly_fit + initial state +
!                 device context (tmcore_swm_init, :29-47, and
!                 test_cases_set_initial_condition)
!   tm_swm_run    the time loop with the per-step invariant line
!                 (tmcore_swm_run, :60-90)
!   tm_swm_final  teardown (tmcore_swm_final, :49-58)
!
! The operators of operators_mod keep their names and dummy arguments
! (operators_mod.F90:40-260) and run on the device.  A non-zero return code
! from the library prints its message and stops, as the reference's
! log_error does.  With use_gpu = .false. (namelist gpu_control) the same
! run takes its steps on the host CPU through the Fortran FP64 CPU path
! (tm_swm_cpu_mod) instead, for the configuration that path implements.
module tm_swm_gpu_mod

  use iso_c_binding
  use tm_c_binding_mod
  use tm_params_mod
  use tm_swm_cpu_mod, only: cpu_setup, cpu_initial_iap, cpu_velocity, cpu_invariants, cpu_steps, cpu_final
  use omp_lib, only: omp_get_wtime

  implicit none

  private

  public tm_swm_init, tm_swm_run, tm_swm_final
  public div_operator, curl_operator, scalar_c2e_interp_operator, scalar_c2v_interp_operator
  public scalar_v2c_interp_operator, scalar_v2e_interp_operator, calc_kinetic_energy
  public calc_tangent_wind, calc_tangent_vor_flux, add_upwind_correction_on_cell
  public calc_pv_on_vertex, calc_pv_on_edge, calc_u_tend_on_edge
  public nCells, nEdges, nVertices, gd, u, iap_u, ghs, tracers, run_inv, run_beta, run_ms

  ! Mesh (mesh_mod.F90:10-71 names), unit-sphere values as a grid file holds them.
  integer(c_int32_t) :: nCells, nEdges, nVertices
  integer(c_int32_t) :: maxEdges = 6, maxEdges2 = 12, vertexDegree = 3
  real(c_double), allocatable, target :: latCell(:), lonCell(:), xCell(:), yCell(:), zCell(:)
  real(c_double), allocatable, target :: latEdge(:), lonEdge(:), xEdge(:), yEdge(:), zEdge(:)
  real(c_double), allocatable, target :: latVertex(:), lonVertex(:), xVertex(:), yVertex(:), zVertex(:)
  integer(c_int32_t), allocatable, target :: nEdgesOnCell(:), nEdgesOnEdge(:)
  integer(c_int32_t), allocatable, target :: cellsOnCell(:,:), edgesOnCell(:,:), verticesOnCell(:,:)
  integer(c_int32_t), allocatable, target :: cellsOnEdge(:,:), verticesOnEdge(:,:), edgesOnEdge(:,:)
  real(c_double), allocatable, target :: weightsOnEdge(:,:)
  integer(c_int32_t), allocatable, target :: cellsOnVertex(:,:), edgesOnVertex(:,:)
  real(c_double), allocatable, target :: kiteAreasOnVertex(:,:)
  real(c_double), allocatable, target :: dvEdge(:), dv1Edge(:), dv2Edge(:), dcEdge(:), angleEdge(:)
  real(c_double), allocatable, target :: areaCell(:), areaTriangle(:)
  ! poly_fit_mod.F90:35-57: the fit families the options read (slots 0:max_fit-1;
  ! 13 holds the second-order stencils, the fourth-order ones need up to 22)
  integer :: max_fit = 13
  integer(c_int32_t), allocatable, target :: nFit2Vertices(:), fit2Vertices(:,:), nFit4Vertices(:), fit4Vertices(:,:)
  integer(c_int32_t), allocatable, target :: nFit2Cells(:), fit2Cells(:,:), nFit4Cells(:), fit4Cells(:,:)
  real(c_double), allocatable, target :: deriv2OnVertex(:,:,:), deriv4OnVertex(:,:,:)
  real(c_double), allocatable, target :: deriv2OnCell(:,:,:), deriv4OnCell(:,:,:)
  ! Prognostics of state(old) and static%cell%ghs
  real(c_double), allocatable, target :: gd(:), u(:), iap_u(:), ghs(:), tracers(:,:)
  real(c_double), allocatable :: run_inv(:,:), run_beta(:)
  real(c_double) :: run_ms = 0

  type(tm_mesh_desc), target :: mesh
  type(tm_fit_desc), target :: fit
  type(tm_opts) :: opts
  type(c_ptr) :: ctx = c_null_ptr
  ! history output (history_mod.F90:27-174)
  type(c_ptr) :: hist = c_null_ptr
  integer :: hist_every = 0          ! steps between history records (0: none)
  real(c_double), allocatable :: h_u(:), h_zonal(:), h_merid(:), h_h(:), h_pv(:), h_vor(:), h_div(:)
  type(c_ptr) :: mlp = c_null_ptr

contains

  subroutine check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    if (rc /= TM_OK) then
      write(*, '(A,I0,A)') '[Error]: ' // where // ' returned ', rc, ': ' // tm_error_text(ctx)
      stop 1
    end if
  end subroutine check

  subroutine check_file(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    if (rc /= TM_OK) then
      write(*, '(A,I0,A)') '[Error]: ' // where // ' returned ', rc, ': ' // c_text(tm_mesh_mpas_error())
      stop 1
    end if
  end subroutine check_file

  subroutine tm_swm_init(namelist_file_path)

    character(*), intent(in) :: namelist_file_path
    integer(c_int64_t), allocatable :: tseeds(:)
    integer :: k
    integer(c_int32_t) :: dims(6)
    real(c_double) :: sphere_r

    call params_parse_namelist(namelist_file_path)
    if (trim(mesh_file_path) /= 'N/A') then
      ! mesh_init (mesh_mod.F90:79-164): an MPAS grid file on the unit sphere
      call check_file(tm_mesh_mpas_dims(trim(mesh_file_path) // c_null_char, dims, sphere_r), 'tm_mesh_mpas_dims')
      if (sphere_r /= 1.0d0) then
        write(*, '(A)') '[Error]: ' // trim(mesh_file_path) // ' is not on the unit sphere (sphere_radius /= 1)'
        stop 1
      end if
      nCells = dims(1); nEdges = dims(2); nVertices = dims(3)
      maxEdges = dims(4); maxEdges2 = dims(5); vertexDegree = dims(6)
      call allocate_mesh()
      call check_file(tm_mesh_mpas_read(trim(mesh_file_path) // c_null_char, mesh), 'tm_mesh_mpas_read')
    else
      call check(tm_mesh_icos_dims(mesh_level, nCells, nEdges, nVertices), 'tm_mesh_icos_dims')
      call allocate_mesh()
      call check(tm_mesh_icos_generate(mesh_level, mesh), 'tm_mesh_icos_generate')
    end if

    ! fit families: vertex order 2 (pv_scheme 3, order >= 3), vertex order 4
    ! (order >= 5), cell orders 2 and 4 (adv_order >= 3, >= 5)
    if (adv_order >= 5 .or. (pv_scheme == 3 .and. interp_pv_order >= 5)) max_fit = 31
    fit%maxFit = max_fit
    fit%nFit2Cells = c_null_ptr;    fit%fit2Cells = c_null_ptr;    fit%deriv2OnCell = c_null_ptr
    fit%nFit4Cells = c_null_ptr;    fit%fit4Cells = c_null_ptr;    fit%deriv4OnCell = c_null_ptr
    fit%nFit4Vertices = c_null_ptr; fit%fit4Vertices = c_null_ptr; fit%deriv4OnVertex = c_null_ptr
    allocate(nFit2Vertices(nVertices), fit2Vertices(0:max_fit-1, nVertices), deriv2OnVertex(0:max_fit-1, 2, nEdges))
    fit%nFit2Vertices = c_loc(nFit2Vertices)
    fit%fit2Vertices = c_loc(fit2Vertices)
    fit%deriv2OnVertex = c_loc(deriv2OnVertex)
    if (pv_scheme == 3 .and. interp_pv_order >= 5) then
      allocate(nFit4Vertices(nVertices), fit4Vertices(0:max_fit-1, nVertices), deriv4OnVertex(0:max_fit-1, 2, nEdges))
      fit%nFit4Vertices = c_loc(nFit4Vertices)
      fit%fit4Vertices = c_loc(fit4Vertices)
      fit%deriv4OnVertex = c_loc(deriv4OnVertex)
    end if
    if (adv_order >= 3) then
      allocate(nFit2Cells(nCells), fit2Cells(0:max_fit-1, nCells), deriv2OnCell(0:max_fit-1, 2, nEdges))
      fit%nFit2Cells = c_loc(nFit2Cells)
      fit%fit2Cells = c_loc(fit2Cells)
      fit%deriv2OnCell = c_loc(deriv2OnCell)
    end if
    if (adv_order >= 5) then
      allocate(nFit4Cells(nCells), fit4Cells(0:max_fit-1, nCells), deriv4OnCell(0:max_fit-1, 2, nEdges))
      fit%nFit4Cells = c_loc(nFit4Cells)
      fit%fit4Cells = c_loc(fit4Cells)
      fit%deriv4OnCell = c_loc(deriv4OnCell)
    end if
    call check(tm_poly_fit(mesh, radius, fit), 'tm_poly_fit')

    allocate(gd(nCells), u(nEdges), iap_u(nEdges), ghs(nCells), tracers(nCells, max(ntracers, 1)))
    select case (trim(case_name))
    case ('seeded')
      allocate(tseeds(max(ntracers, 1)))
      do k = 1, max(ntracers, 1)
        tseeds(k) = tracer_seed0 + k - 1
      end do
      call check(tm_case_seeded(mesh, radius, seed, ntracers, tseeds, gd, u, ghs, c_loc(tracers)), 'tm_case_seeded')
    case ('steady_geostrophic_flow')
      call check(tm_case_steady_geostrophic(mesh, radius, gd, u, ghs), 'tm_case_steady_geostrophic')
      tracers = 0
    case default
      write(*, '(A)') '[Error]: unknown case_name ' // trim(case_name) // ' (seeded, steady_geostrophic_flow)'
      stop 1
    end select

    call tm_opts_default(opts)
    opts%time_scheme = time_scheme
    opts%pv_scheme = pv_scheme
    opts%interp_pv_order = interp_pv_order
    opts%adv_order = adv_order
    opts%adv_mono = merge(1, 0, adv_mono)
    opts%conserve_energy = merge(1, 0, conserve_energy)
    opts%apvm_weight = apvm_weight
    opts%dt = dt
    opts%radius = radius
    opts%ntracers = ntracers
    opts%precision = precision
    opts%device = device
    opts%renumber = renumber
    if (flux_4th_order_correct) then
      write(*, '(A)') '[Error]: flux_4th_order_correct = .true. is not reproducible in the reference (deriv3OnCell)'
      stop 1
    end if
    if (use_gpu) then
      call check(tm_ctx_create(ctx, mesh, fit, opts), 'tm_ctx_create')
      call check(tm_state_upload(ctx, c_loc(gd), c_loc(u), c_null_ptr, c_loc(ghs), c_loc(tracers)), 'tm_state_upload')
      if (trim(history_interval(1)) /= 'N/A') call history_init()
      if (trim(closure_file) /= 'N/A') call attach_closure(trim(closure_file))
    else
      if (time_scheme /= 2 .or. .not. conserve_energy .or. pv_scheme /= 3 .or. interp_pv_order < 2 .or. &
          interp_pv_order > 4 .or. adv_order /= 2 .or. precision /= 64 .or. trim(closure_file) /= 'N/A' .or. &
          trim(history_interval(1)) /= 'N/A') then
        write(*, '(A)') '[Error]: use_gpu = .false. runs CRK4 with conserve_energy, pv_scheme 3 at interp_pv_order' // &
          ' 2-4, adv_order 2 and precision 64, without closure_file or history'
        stop 1
      end if
      if (cpu_setup(mesh, fit, radius, interp_pv_order, ntracers) /= 0) then
        write(*, '(A)') '[Error]: the CPU path could not set up the mesh'
        stop 1
      end if
      call cpu_initial_iap(gd, u, iap_u)
    end if

    if (trim(mesh_file_path) /= 'N/A') then
      write(*, '(A,I0,A,I0,A,I0,A,F8.2,A,I0)') '[Notice]: mesh ' // trim(mesh_file_path) // ': ', nCells, ' cells, ', &
        nEdges, ' edges, ', nVertices, ' vertices; dt = ', dt, ' s; precision ', precision
    else
      write(*, '(A,I0,A,I0,A,I0,A,I0,A,F8.2,A,I0)') '[Notice]: mesh level ', mesh_level, ': ', nCells, ' cells, ', &
        nEdges, ' edges, ', nVertices, ' vertices; dt = ', dt, ' s; precision ', precision
    end if

  end subroutine tm_swm_init

  subroutine allocate_mesh()

    allocate(latCell(nCells), lonCell(nCells), xCell(nCells), yCell(nCells), zCell(nCells))
    allocate(latEdge(nEdges), lonEdge(nEdges), xEdge(nEdges), yEdge(nEdges), zEdge(nEdges))
    allocate(latVertex(nVertices), lonVertex(nVertices), xVertex(nVertices), yVertex(nVertices), zVertex(nVertices))
    allocate(nEdgesOnCell(nCells), nEdgesOnEdge(nEdges))
    allocate(cellsOnCell(maxEdges, nCells), edgesOnCell(maxEdges, nCells), verticesOnCell(maxEdges, nCells))
    allocate(cellsOnEdge(2, nEdges), verticesOnEdge(2, nEdges))
    allocate(edgesOnEdge(maxEdges2, nEdges), weightsOnEdge(maxEdges2, nEdges))
    allocate(cellsOnVertex(vertexDegree, nVertices), edgesOnVertex(vertexDegree, nVertices))
    allocate(kiteAreasOnVertex(vertexDegree, nVertices))
    allocate(dvEdge(nEdges), dv1Edge(nEdges), dv2Edge(nEdges), dcEdge(nEdges), angleEdge(nEdges))
    allocate(areaCell(nCells), areaTriangle(nVertices))
    mesh%nCells = nCells; mesh%nEdges = nEdges; mesh%nVertices = nVertices
    mesh%maxEdges = maxEdges; mesh%maxEdges2 = maxEdges2; mesh%vertexDegree = vertexDegree
    mesh%latCell = c_loc(latCell); mesh%lonCell = c_loc(lonCell)
    mesh%xCell = c_loc(xCell); mesh%yCell = c_loc(yCell); mesh%zCell = c_loc(zCell)
    mesh%latEdge = c_loc(latEdge); mesh%lonEdge = c_loc(lonEdge)
    mesh%xEdge = c_loc(xEdge); mesh%yEdge = c_loc(yEdge); mesh%zEdge = c_loc(zEdge)
    mesh%latVertex = c_loc(latVertex); mesh%lonVertex = c_loc(lonVertex)
    mesh%xVertex = c_loc(xVertex); mesh%yVertex = c_loc(yVertex); mesh%zVertex = c_loc(zVertex)
    mesh%nEdgesOnCell = c_loc(nEdgesOnCell); mesh%nEdgesOnEdge = c_loc(nEdgesOnEdge)
    mesh%cellsOnCell = c_loc(cellsOnCell); mesh%edgesOnCell = c_loc(edgesOnCell)
    mesh%verticesOnCell = c_loc(verticesOnCell)
    mesh%cellsOnEdge = c_loc(cellsOnEdge); mesh%verticesOnEdge = c_loc(verticesOnEdge)
    mesh%edgesOnEdge = c_loc(edgesOnEdge); mesh%weightsOnEdge = c_loc(weightsOnEdge)
    mesh%cellsOnVertex = c_loc(cellsOnVertex); mesh%edgesOnVertex = c_loc(edgesOnVertex)
    mesh%kiteAreasOnVertex = c_loc(kiteAreasOnVertex)
    mesh%dvEdge = c_loc(dvEdge); mesh%dv1Edge = c_loc(dv1Edge); mesh%dv2Edge = c_loc(dv2Edge)
    mesh%dcEdge = c_loc(dcEdge); mesh%angleEdge = c_loc(angleEdge)
    mesh%areaCell = c_loc(areaCell); mesh%areaTriangle = c_loc(areaTriangle)

  end subroutine allocate_mesh

  ! The time loop of tmcore_swm_run (tmcore_swm_mod.F90:75-88): one device
  ! call per print interval; every step's invariants and beta come back and
  ! are logged as the reference's log_print_diag does.
  subroutine tm_swm_run()

    integer :: nsteps, done, n, i
    real(c_double) :: inv0(4), ms
    real(c_double), allocatable :: inv(:,:), beta(:)

    nsteps = run_steps()
    allocate(run_inv(4, nsteps), run_beta(nsteps))
    if (.not. use_gpu) then
      call cpu_run(nsteps)
      return
    end if
    call check(tm_invariants(ctx, inv0), 'tm_invariants')
    write(*, '(A,I8,4(A,ES24.16))') 'step=', 0, ' total_mass=', inv0(1), ' total_energy=', inv0(2), &
      ' total_potential_enstropy=', inv0(3), ' total_absolute_vorticity=', inv0(4)
    done = 0
    run_ms = 0
    if (c_associated(hist)) call history_write(0)
    do while (done < nsteps)
      n = min(max(print_every, 1), nsteps - done)
      if (hist_every > 0) n = min(n, hist_every - mod(done, hist_every))
      allocate(inv(4, n), beta(n))
      call check(tm_step(ctx, n, inv, beta), 'tm_step')
      call check(tm_last_elapsed_ms(ctx, ms), 'tm_last_elapsed_ms')
      run_ms = run_ms + ms
      run_inv(:, done+1:done+n) = inv
      run_beta(done+1:done+n) = beta
      do i = 1, n
        if (mod(done + i, max(print_every, 1)) == 0 .or. done + i == nsteps) then
          write(*, '(A,I8,5(A,ES24.16))') 'step=', done + i, ' total_mass=', inv(1,i), ' total_energy=', inv(2,i), &
            ' total_potential_enstropy=', inv(3,i), ' total_absolute_vorticity=', inv(4,i), ' beta=', beta(i)
        end if
      end do
      deallocate(inv, beta)
      done = done + n
      if (hist_every > 0) then
        if (mod(done, hist_every) == 0) call history_write(done)
      end if
    end do
    call check(tm_state_download(ctx, c_loc(gd), c_loc(u), c_loc(iap_u), c_loc(tracers)), 'tm_state_download')
    write(*, '(A,I0,A,F12.3,A,F14.3)') '[Notice]: ', nsteps, ' steps, device time ', run_ms, ' ms, sim-days/hour ', &
      (nsteps * dt / 86400.0_r8) / (max(run_ms, 1.0e-9_r8) / 3.6e6_r8)

  end subroutine tm_swm_run

  ! tm_swm_run on the host CPU: the same print intervals and lines, the steps
  ! of the Fortran FP64 CPU path, the wall time of the steps as run_ms.
  subroutine cpu_run(nsteps)

    integer, intent(in) :: nsteps
    integer :: done, n, i
    real(c_double) :: inv0(4), t0
    real(c_double), allocatable :: inv(:,:), beta(:)

    call cpu_invariants(gd, iap_u, ghs, inv0)
    write(*, '(A,I8,4(A,ES24.16))') 'step=', 0, ' total_mass=', inv0(1), ' total_energy=', inv0(2), &
      ' total_potential_enstropy=', inv0(3), ' total_absolute_vorticity=', inv0(4)
    done = 0
    run_ms = 0
    do while (done < nsteps)
      n = min(max(print_every, 1), nsteps - done)
      allocate(inv(4, n), beta(n))
      t0 = omp_get_wtime()
      call cpu_steps(n, dt, gd, iap_u, ghs, ntracers, tracers, inv, beta)
      run_ms = run_ms + 1000.0_r8 * (omp_get_wtime() - t0)
      run_inv(:, done+1:done+n) = inv
      run_beta(done+1:done+n) = beta
      do i = 1, n
        if (mod(done + i, max(print_every, 1)) == 0 .or. done + i == nsteps) then
          write(*, '(A,I8,5(A,ES24.16))') 'step=', done + i, ' total_mass=', inv(1,i), ' total_energy=', inv(2,i), &
            ' total_potential_enstropy=', inv(3,i), ' total_absolute_vorticity=', inv(4,i), ' beta=', beta(i)
        end if
      end do
      deallocate(inv, beta)
      done = done + n
    end do
    call cpu_velocity(gd, iap_u, u)
    write(*, '(A,I0,A,F12.3,A,F14.3)') '[Notice]: ', nsteps, ' steps on the host CPU, time ', run_ms, &
      ' ms, sim-days/hour ', (nsteps * dt / 86400.0_r8) / (max(run_ms, 1.0e-9_r8) / 3.6e6_r8)

  end subroutine cpu_run

  ! The learned closure of a raw weights file (see tm_params_mod closure_file),
  ! applied in every following step.
  subroutine attach_closure(path)

    character(*), intent(in) :: path
    real(c_float), allocatable :: w(:)
    integer :: un, ios
    integer, parameter :: nparams = 38530

    allocate(w(nparams + 82))
    open(newunit=un, file=path, access='stream', form='unformatted', status='old', iostat=ios)
    if (ios == 0) read(un, iostat=ios) w
    if (ios /= 0) then
      write(*, '(A)') '[Error]: cannot read closure_file ' // path
      stop 1
    end if
    close(un)
    call check(tm_mlp_create(mlp, device, 0_c_int64_t), 'tm_mlp_create')
    call check(tm_mlp_set_params(mlp, w(1:nparams)), 'tm_mlp_set_params')
    call check(tm_mlp_set_scales(mlp, w(nparams + 1:nparams + 40), w(nparams + 41:nparams + 80), &
                                 w(nparams + 81:nparams + 82)), 'tm_mlp_set_scales')
    call check(tm_closure_attach(ctx, mlp), 'tm_closure_attach')
    call check(tm_closure_set_interval(ctx, closure_interval, closure_lumped), 'tm_closure_set_interval')
    call check(tm_closure_set_gain(ctx, closure_gain), 'tm_closure_set_gain')
    write(*, '(A)') '[Notice]: learned closure attached from ' // path

  end subroutine attach_closure

  ! history_init (history_mod.F90:27-98): the interval '<value> <units>' in
  ! whole steps, the file <output_file_prefix or case_name>.h0.nc and the
  ! reconstruction coefficients on the device.
  subroutine history_init()

    character(max_name_len) :: value_s, units_s
    real(r8) :: seconds
    real(c_double), allocatable :: coeffs(:)
    character(max_file_path_len) :: path
    integer :: ios

    read(history_interval(1), *, iostat=ios) value_s, units_s
    if (ios == 0) read(value_s, *, iostat=ios) seconds
    if (ios /= 0) then
      write(*, '(A)') '[Error]: invalid history interval ' // trim(history_interval(1))
      stop 1
    end if
    select case (trim(units_s))
    case ('days')
      seconds = seconds * 86400
    case ('hours')
      seconds = seconds * 3600
    case ('minutes')
      seconds = seconds * 60
    case ('seconds')
    case default
      write(*, '(A)') '[Error]: invalid history interval ' // trim(history_interval(1))
      stop 1
    end select
    hist_every = max(1, nint(seconds / dt))
    if (trim(output_file_prefix) /= 'N/A') then
      path = trim(output_file_prefix) // '.h0.nc'
    else
      path = trim(case_name) // '.h0.nc'
    end if
    allocate(coeffs(3 * maxEdges * nCells))
    call check_file(tm_reconstruct_coeffs(mesh, radius, coeffs), 'tm_reconstruct_coeffs')
    call check(tm_history_setup(ctx, coeffs, latCell, lonCell), 'tm_history_setup')
    call check_file(tm_history_create(hist, trim(path) // c_null_char, mesh, radius, dt, int(time_scheme, c_int), &
                                      trim(case_name) // c_null_char), 'tm_history_create')
    allocate(h_u(nEdges), h_zonal(nCells), h_merid(nCells), h_h(nCells), h_pv(nVertices), h_vor(nVertices), &
             h_div(nCells))
    write(*, '(A,I0,A)') '[Notice]: history every ', hist_every, ' steps to ' // trim(path)

  end subroutine history_init

  ! history_write_state (history_mod.F90:106-172) of state(old) after `step` steps
  subroutine history_write(step)

    integer, intent(in) :: step
    real(c_double) :: inv(4)

    call check(tm_history_fields(ctx, h_u, h_zonal, h_merid, h_h, h_pv, h_vor, h_div), 'tm_history_fields')
    call check(tm_invariants(ctx, inv), 'tm_invariants')
    call check_file(tm_history_write(hist, step * dt, h_u, h_zonal, h_merid, h_h, h_pv, h_vor, h_div, inv(1), inv(2)), &
                    'tm_history_write')

  end subroutine history_write

  subroutine tm_swm_final()

    integer(c_int) :: rc
    if (c_associated(hist)) rc = tm_history_close(hist)
    hist = c_null_ptr
    if (c_associated(ctx)) rc = tm_ctx_destroy(ctx)
    ctx = c_null_ptr
    if (c_associated(mlp)) rc = tm_mlp_destroy(mlp)
    mlp = c_null_ptr
    call cpu_final()

  end subroutine tm_swm_final

  ! ---------------------------------------------------------------- operators
  ! operators_mod.F90 names and argument order, computed on the device.
  subroutine div_operator(f_edge, div)
    real(c_double), intent(in) :: f_edge(:)
    real(c_double), intent(out) :: div(:)
    call check(tm_op_div(ctx, f_edge, div), 'tm_op_div')
  end subroutine div_operator

  subroutine curl_operator(f_edge, curl)
    real(c_double), intent(in) :: f_edge(:)
    real(c_double), intent(out) :: curl(:)
    call check(tm_op_curl(ctx, f_edge, curl), 'tm_op_curl')
  end subroutine curl_operator

  subroutine scalar_c2e_interp_operator(f_cell, f_edge)
    real(c_double), intent(in) :: f_cell(:)
    real(c_double), intent(out) :: f_edge(:)
    call check(tm_op_c2e(ctx, f_cell, f_edge), 'tm_op_c2e')
  end subroutine scalar_c2e_interp_operator

  subroutine scalar_c2v_interp_operator(f_cell, f_vertex)
    real(c_double), intent(in) :: f_cell(:)
    real(c_double), intent(out) :: f_vertex(:)
    call check(tm_op_c2v(ctx, f_cell, f_vertex), 'tm_op_c2v')
  end subroutine scalar_c2v_interp_operator

  subroutine scalar_v2c_interp_operator(f_vertex, f_cell)
    real(c_double), intent(in) :: f_vertex(:)
    real(c_double), intent(out) :: f_cell(:)
    call check(tm_op_v2c(ctx, f_vertex, f_cell), 'tm_op_v2c')
  end subroutine scalar_v2c_interp_operator

  subroutine scalar_v2e_interp_operator(f_vertex, f_edge)
    real(c_double), intent(in) :: f_vertex(:)
    real(c_double), intent(out) :: f_edge(:)
    call check(tm_op_v2e(ctx, f_vertex, f_edge), 'tm_op_v2e')
  end subroutine scalar_v2e_interp_operator

  subroutine calc_kinetic_energy(u_edge, ke_cell)
    real(c_double), intent(in) :: u_edge(:)
    real(c_double), intent(out) :: ke_cell(:)
    call check(tm_op_ke(ctx, u_edge, ke_cell), 'tm_op_ke')
  end subroutine calc_kinetic_energy

  subroutine calc_tangent_wind(u_edge, v_edge)
    real(c_double), intent(in) :: u_edge(:)
    real(c_double), intent(out) :: v_edge(:)
    call check(tm_op_tangent_wind(ctx, u_edge, v_edge), 'tm_op_tangent_wind')
  end subroutine calc_tangent_wind

  subroutine calc_tangent_vor_flux(u_edge, gd_edge, pv_edge, tangent_vor_flux)
    real(c_double), intent(in) :: u_edge(:), gd_edge(:), pv_edge(:)
    real(c_double), intent(out) :: tangent_vor_flux(:)
    call check(tm_op_tangent_vor_flux(ctx, u_edge, gd_edge, pv_edge, tangent_vor_flux), 'tm_op_tangent_vor_flux')
  end subroutine calc_tangent_vor_flux

  subroutine add_upwind_correction_on_cell(f_cell, u_edge, f_edge)
    real(c_double), intent(in) :: f_cell(:), u_edge(:)
    real(c_double), intent(inout) :: f_edge(:)
    call check(tm_op_upwind_cell(ctx, f_cell, u_edge, f_edge), 'tm_op_upwind_cell')
  end subroutine add_upwind_correction_on_cell

  subroutine calc_pv_on_vertex(vor_vertex, gd_vertex, pv_vertex)
    real(c_double), intent(in) :: vor_vertex(:), gd_vertex(:)
    real(c_double), intent(out) :: pv_vertex(:)
    call check(tm_op_pv_vertex(ctx, vor_vertex, gd_vertex, pv_vertex), 'tm_op_pv_vertex')
  end subroutine calc_pv_on_vertex

  subroutine calc_pv_on_edge(u_edge, v_edge, gd_edge, gd_tend_vertex, pv_vertex, pv_cell, pv_edge)
    real(c_double), intent(in) :: u_edge(:), v_edge(:), gd_edge(:), gd_tend_vertex(:), pv_vertex(:), pv_cell(:)
    real(c_double), intent(out) :: pv_edge(:)
    call check(tm_op_pv_edge(ctx, u_edge, v_edge, gd_edge, gd_tend_vertex, pv_vertex, pv_cell, pv_edge), &
               'tm_op_pv_edge')
  end subroutine calc_pv_on_edge

  subroutine calc_u_tend_on_edge(u_edge, ke_cell, gd_cell, gd_edge, ghs_cell, pv_flx_edge, gd_tend_edge, &
                                 u_tend_edge, iap_u_tend_edge)
    real(c_double), intent(in) :: u_edge(:), ke_cell(:), gd_cell(:), gd_edge(:), ghs_cell(:), pv_flx_edge(:)
    real(c_double), intent(in) :: gd_tend_edge(:)
    real(c_double), intent(out) :: u_tend_edge(:), iap_u_tend_edge(:)
    call check(tm_op_u_tend(ctx, u_edge, ke_cell, gd_cell, gd_edge, ghs_cell, pv_flx_edge, gd_tend_edge, &
                            u_tend_edge, iap_u_tend_edge), 'tm_op_u_tend')
  end subroutine calc_u_tend_on_edge

end module tm_swm_gpu_mod
