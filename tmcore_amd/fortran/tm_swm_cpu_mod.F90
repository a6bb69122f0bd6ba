! FP64 CPU path of the shallow-water step in Fortran: the operators of
! operators_mod.F90 that the benchmark configuration runs (pv_scheme 3 at
! interp_pv_order 2-4, adv_order 2), the stage of tmcore_swm_mod.F90:92-125,
! the CRK4 step with the energy-conserving beta (time_scheme_mod.F90:109-154),
! the tracer flux advection (adv_scheme_sg11_mod.F90:33-48) and the invariants
! of diag_run (tmcore_swm_mod.F90:215-255).  It is the CPU baseline the GPU
! path is timed against: element loops run over OpenMP threads, each
! element's arithmetic in the reference's association order, and the global
! sums stay serial in element order, as the reference's are.
!
! A Fortran host uses it through cpu_setup / cpu_initial_iap / cpu_invariants
! / cpu_steps / cpu_final (one mesh at a time: the module holds the mesh and
! the work arrays, so the procedures are not re-entrant).  tm_cpu_run
! (bind(c)) does the same for a C caller: it takes the reference-layout mesh
! and fits of include/tmcore_host.h (1-based connectivity, unit-sphere
! geometry) and the state in reference order, advances it nsteps steps and
! returns the invariants and beta of every step and the wall time of the step
! loop.
module tm_swm_cpu_mod

  use iso_c_binding
  use omp_lib
  use tm_desc_mod, only: tm_mesh_desc, tm_fit_desc

  implicit none

  private
  public tm_cpu_run, cpu_setup, cpu_initial_iap, cpu_velocity, cpu_invariants, cpu_steps, cpu_final

  real(8), parameter :: omega = 7.292d-5     ! const_mod.F90:13

  ! the mesh as mesh_init leaves it (radius-scaled), with the derived arrays
  ! of mesh_mod.F90:169-224
  type cpu_mesh_t
    integer :: nC, nE, nV, me, me2, mf
    integer, pointer, contiguous :: nEoC(:), nEoE(:), eoc(:,:), coe(:,:), voe(:,:), eoe(:,:), cov(:,:), eov(:,:)
    real(8), pointer, contiguous :: woe(:,:), kite(:,:)
    real(8), allocatable :: dv(:), dc(:), aE(:), aC(:), aT(:), fV(:)
    integer, allocatable :: nsgn(:,:), tsgn(:,:)
    real(8) :: total_area
    integer, pointer, contiguous :: nfit(:), fit(:,:)
    real(8), pointer, contiguous :: d2v(:,:,:)
  end type cpu_mesh_t

  ! the mesh of cpu_setup, its PV interpolation order, and the work arrays
  type(cpu_mesh_t) :: cm
  integer :: cm_order = 0
  logical :: cm_ready = .false.
  real(8), allocatable :: gd_e(:), u(:), tgd_e(:), v(:), pv_e(:), pv_flx(:), q_e(:), flx(:)
  real(8), allocatable :: tgd_v(:), gd_v(:), vor(:), pv_v(:), ke(:), pv_c(:), en(:)
  ! state and tendency slots of the CRK4 step (state_mod.F90:63, tend_mod.F90:44)
  real(8), allocatable :: sg(:,:), si(:,:), sq(:,:,:), tg(:,:), ti(:,:), tq(:,:,:)

contains

  integer(c_int) function tm_cpu_run(md, fd, radius, dt, interp_pv_order, ntracers, nsteps, nthreads, &
                                     gd, iap, ghs, q, inv, beta, elapsed) bind(c, name='tm_cpu_run')
    type(tm_mesh_desc), intent(in) :: md
    type(tm_fit_desc), intent(in) :: fd
    real(c_double), value :: radius, dt
    integer(c_int), value :: interp_pv_order, ntracers, nsteps, nthreads
    real(c_double), intent(inout) :: gd(md%nCells), iap(md%nEdges), q(md%nCells, *)
    real(c_double), intent(in) :: ghs(md%nCells)
    real(c_double), intent(out) :: inv(4, *), beta(*)
    real(c_double), intent(out) :: elapsed
    real(8) :: t0

    tm_cpu_run = -1
    if (ntracers < 0 .or. nsteps < 0) return
    if (nthreads > 0) call omp_set_num_threads(nthreads)
    tm_cpu_run = cpu_setup(md, fd, radius, interp_pv_order, ntracers)
    if (tm_cpu_run /= 0) return
    t0 = omp_get_wtime()
    call cpu_steps(nsteps, dt, gd, iap, ghs, ntracers, q, inv, beta)
    elapsed = omp_get_wtime() - t0
    call cpu_final()
  end function tm_cpu_run

  ! mesh_init + poly_fit_init's tables for the CPU path (tmcore_swm_mod.F90:37-41):
  ! 0, or -1 when the order is not 2-4, the fits are missing or a mesh is set up
  integer function cpu_setup(md, fd, radius, interp_pv_order, ntracers) result(rc)
    type(tm_mesh_desc), intent(in) :: md
    type(tm_fit_desc), intent(in) :: fd
    real(8), intent(in) :: radius
    integer, intent(in) :: interp_pv_order, ntracers

    rc = -1
    if (cm_ready .or. interp_pv_order < 2 .or. interp_pv_order > 4 .or. ntracers < 0) return
    if (interp_pv_order >= 3 .and. .not. c_associated(fd%deriv2OnVertex)) return
    call mesh_setup(cm, md, fd, radius, interp_pv_order >= 3)
    allocate(sg(cm%nC, 4), si(cm%nE, 4), sq(cm%nC, max(ntracers, 1), 4), tg(cm%nC, 0:4), ti(cm%nE, 0:4), &
             tq(cm%nC, max(ntracers, 1), 0:4))
    cm_order = interp_pv_order
    cm_ready = .true.
    rc = 0
  end function cpu_setup

  subroutine cpu_final()
    if (.not. cm_ready) return
    call work_free()
    call mesh_free(cm)
    deallocate(sg, si, sq, tg, ti, tq)
    cm_ready = .false.
  end subroutine cpu_final

  ! iap_u = sqrt(c2e(gd)) * u, the run prologue (tmcore_swm_mod.F90:62-64)
  subroutine cpu_initial_iap(gd, uu, iap)
    real(8), intent(in), contiguous :: gd(:), uu(:)
    real(8), intent(out), contiguous :: iap(:)
    integer :: e
    call cell_to_edge(cm, gd, gd_e)
    !$omp parallel do
    do e = 1, cm%nE
      iap(e) = sqrt(gd_e(e)) * uu(e)
    end do
  end subroutine cpu_initial_iap

  ! u = iap_u / sqrt(c2e(gd)), the refresh of tmcore_swm_mod.F90:78-80
  subroutine cpu_velocity(gd, iap, uu)
    real(8), intent(in), contiguous :: gd(:), iap(:)
    real(8), intent(out), contiguous :: uu(:)
    integer :: e
    call cell_to_edge(cm, gd, gd_e)
    !$omp parallel do
    do e = 1, cm%nE
      uu(e) = iap(e) / sqrt(gd_e(e))
    end do
  end subroutine cpu_velocity

  ! the invariants of a state (the refresh and diag_run of tmcore_swm_mod.F90:78-85)
  subroutine cpu_invariants(gd, iap, ghs, out)
    real(8), intent(in), contiguous :: gd(:), iap(:), ghs(:)
    real(8), intent(out) :: out(4)
    call invariants(cm, gd, iap, ghs, out)
  end subroutine cpu_invariants

  ! nsteps CRK4 steps with beta (time_scheme_mod.F90:109-154); the tracers
  ! q(:, 1:ntracers) with plain RK4 weights; inv(:, s) and beta(s) of step s
  subroutine cpu_steps(nsteps, dt, gd, iap, ghs, nT, q, inv, beta)
    integer, intent(in) :: nsteps, nT
    real(8), intent(in) :: dt
    real(8), intent(inout) :: gd(cm%nC), iap(cm%nE), q(cm%nC, *)
    real(8), intent(in) :: ghs(cm%nC)
    real(8), intent(out) :: inv(4, *), beta(*)
    real(8) :: cf(3), b, r12, r23, r34, r44
    integer :: step, k, i, t

    cf = [0.5d0, 0.5d0, 1.0d0]
    do step = 1, nsteps
      sg(:,1) = gd
      si(:,1) = iap
      if (nT > 0) sq(:,1:nT,1) = q(:,1:nT)
      do k = 1, 4
        call stage(cm, cm_order, sg(:,k), si(:,k), ghs, nT, sq(:,:,k), tg(:,k), ti(:,k), tq(:,:,k))
        if (k < 4) then
          call axpy(sg(:,1), cf(k) * dt, tg(:,k), sg(:,k+1))
          call axpy(si(:,1), cf(k) * dt, ti(:,k), si(:,k+1))
          do t = 1, nT
            call axpy(sq(:,t,1), cf(k) * dt, tq(:,t,k), sq(:,t,k+1))
          end do
        end if
      end do
      !$omp parallel do
      do i = 1, cm%nC
        tg(i,0) = (((tg(i,1) + 2.0d0 * tg(i,2)) + 2.0d0 * tg(i,3)) + tg(i,4)) / 6.0d0
      end do
      !$omp parallel do
      do i = 1, cm%nE
        ti(i,0) = (((ti(i,1) + 2.0d0 * ti(i,2)) + 2.0d0 * ti(i,3)) + ti(i,4)) / 6.0d0
      end do
      do t = 1, nT
        !$omp parallel do
        do i = 1, cm%nC
          tq(i,t,0) = (((tq(i,t,1) + 2.0d0 * tq(i,t,2)) + 2.0d0 * tq(i,t,3)) + tq(i,t,4)) / 6.0d0
        end do
      end do
      r44 = inner(cm, tg(:,0), ti(:,0), tg(:,0), ti(:,0))
      r12 = inner(cm, tg(:,1), ti(:,1), tg(:,2), ti(:,2))
      r23 = inner(cm, tg(:,2), ti(:,2), tg(:,3), ti(:,3))
      r34 = inner(cm, tg(:,3), ti(:,3), tg(:,4), ti(:,4))
      b = (r12 + r23 + r34) / (3.0d0 * r44)
      call add_scaled(gd, b * dt, tg(:,0))
      call add_scaled(iap, b * dt, ti(:,0))
      do t = 1, nT
        call add_scaled(q(:,t), dt, tq(:,t,0))
      end do
      beta(step) = b
      call invariants(cm, gd, iap, ghs, inv(:,step))
    end do
  end subroutine cpu_steps

  ! mesh_init's scaling and the derived arrays (mesh_mod.F90:169-224); the
  ! caller's connectivity is viewed in place
  subroutine mesh_setup(m, md, fd, radius, want_fit)
    type(cpu_mesh_t), intent(out) :: m
    type(tm_mesh_desc), intent(in) :: md
    type(tm_fit_desc), intent(in) :: fd
    real(8), intent(in) :: radius
    logical, intent(in) :: want_fit
    real(8), pointer, contiguous :: p1(:), lat(:)
    integer :: i, k, e, vx

    m%nC = md%nCells; m%nE = md%nEdges; m%nV = md%nVertices
    m%me = md%maxEdges; m%me2 = md%maxEdges2
    call c_f_pointer(md%nEdgesOnCell, m%nEoC, [m%nC])
    call c_f_pointer(md%nEdgesOnEdge, m%nEoE, [m%nE])
    call c_f_pointer(md%edgesOnCell, m%eoc, [m%me, m%nC])
    call c_f_pointer(md%cellsOnEdge, m%coe, [2, m%nE])
    call c_f_pointer(md%verticesOnEdge, m%voe, [2, m%nE])
    call c_f_pointer(md%edgesOnEdge, m%eoe, [m%me2, m%nE])
    call c_f_pointer(md%weightsOnEdge, m%woe, [m%me2, m%nE])
    call c_f_pointer(md%cellsOnVertex, m%cov, [3, m%nV])
    call c_f_pointer(md%edgesOnVertex, m%eov, [3, m%nV])
    call c_f_pointer(md%kiteAreasOnVertex, p1, [3 * m%nV])
    allocate(m%dv(m%nE), m%dc(m%nE), m%aE(m%nE), m%aC(m%nC), m%aT(m%nV), m%fV(m%nV))
    ! kites are scaled into a copy (the caller's array stays on the unit sphere)
    allocate(m%kite(3, m%nV))
    m%kite = reshape(p1, [3, m%nV]) * (radius * radius)
    call c_f_pointer(md%dvEdge, p1, [m%nE])
    m%dv = p1
    call c_f_pointer(md%dcEdge, lat, [m%nE])
    m%dc = lat
    do e = 1, m%nE
      m%aE(e) = (m%dv(e) * m%dc(e)) * (radius * radius)
    end do
    m%dv = m%dv * radius
    m%dc = m%dc * radius
    call c_f_pointer(md%areaCell, p1, [m%nC])
    m%aC = p1 * (radius * radius)
    call c_f_pointer(md%areaTriangle, p1, [m%nV])
    m%aT = p1 * (radius * radius)
    call c_f_pointer(md%latVertex, lat, [m%nV])
    do vx = 1, m%nV
      m%fV(vx) = 2.0d0 * omega * sin(lat(vx))
    end do
    allocate(m%nsgn(m%me, m%nC), m%tsgn(3, m%nV))
    m%nsgn = 0
    do i = 1, m%nC
      do k = 1, m%nEoC(i)
        e = m%eoc(k,i)
        if (i == m%coe(1,e)) m%nsgn(k,i) = 1
        if (i == m%coe(2,e)) m%nsgn(k,i) = -1
      end do
    end do
    m%tsgn = 0
    do vx = 1, m%nV
      do k = 1, 3
        e = m%eov(k,vx)
        if (vx == m%voe(1,e)) m%tsgn(k,vx) = 1
        if (vx == m%voe(2,e)) m%tsgn(k,vx) = -1
      end do
    end do
    m%total_area = 0.0d0
    do i = 1, m%nC
      m%total_area = m%total_area + m%aC(i)
    end do
    m%mf = fd%maxFit
    nullify(m%nfit, m%fit, m%d2v)
    if (want_fit) then
      call c_f_pointer(fd%nFit2Vertices, m%nfit, [m%nV])
      call c_f_pointer(fd%fit2Vertices, m%fit, [m%mf, m%nV])
      call c_f_pointer(fd%deriv2OnVertex, m%d2v, [m%mf, 2, m%nE])
    end if
    call work_alloc(m)
  end subroutine mesh_setup

  subroutine mesh_free(m)
    type(cpu_mesh_t), intent(inout) :: m
    deallocate(m%dv, m%dc, m%aE, m%aC, m%aT, m%fV, m%nsgn, m%tsgn, m%kite)
  end subroutine mesh_free

  subroutine work_alloc(m)
    type(cpu_mesh_t), intent(in) :: m
    allocate(gd_e(m%nE), u(m%nE), tgd_e(m%nE), v(m%nE), pv_e(m%nE), pv_flx(m%nE), q_e(m%nE), flx(m%nE))
    allocate(tgd_v(m%nV), gd_v(m%nV), vor(m%nV), pv_v(m%nV), ke(m%nC), pv_c(m%nC), en(m%nC))
  end subroutine work_alloc

  subroutine work_free()
    deallocate(gd_e, u, tgd_e, v, pv_e, pv_flx, q_e, flx, tgd_v, gd_v, vor, pv_v, ke, pv_c, en)
  end subroutine work_free

  ! out = old + a * t (update_state, tmcore_swm_mod.F90:203-213)
  subroutine axpy(old, a, t, out)
    real(8), intent(in), contiguous :: old(:), t(:)
    real(8), intent(in) :: a
    real(8), intent(inout), contiguous :: out(:)
    integer :: i
    !$omp parallel do
    do i = 1, size(old)
      out(i) = old(i) + a * t(i)
    end do
  end subroutine axpy

  ! x = x + a * t (the final update of update_state on state(old))
  subroutine add_scaled(x, a, t)
    real(8), intent(inout), contiguous :: x(:)
    real(8), intent(in), contiguous :: t(:)
    real(8), intent(in) :: a
    integer :: i
    !$omp parallel do
    do i = 1, size(x)
      x(i) = x(i) + a * t(i)
    end do
  end subroutine add_scaled

  ! div_operator (operators_mod.F90:40-55): minus the divergence
  subroutine divergence(m, f, d)
    type(cpu_mesh_t), intent(in) :: m
    real(8), intent(in), contiguous :: f(:)
    real(8), intent(out), contiguous :: d(:)
    integer :: i, k, e
    real(8) :: s
    !$omp parallel do private(k, e, s)
    do i = 1, m%nC
      s = 0.0d0
      do k = 1, m%nEoC(i)
        e = m%eoc(k,i)
        s = s + (dble(m%nsgn(k,i)) * f(e)) * m%dv(e)
      end do
      d(i) = (-s) / m%aC(i)
    end do
  end subroutine divergence

  ! curl_operator (operators_mod.F90:76-91)
  subroutine curl(m, f, r)
    type(cpu_mesh_t), intent(in) :: m
    real(8), intent(in), contiguous :: f(:)
    real(8), intent(out), contiguous :: r(:)
    integer :: vx, k, e
    real(8) :: s
    !$omp parallel do private(k, e, s)
    do vx = 1, m%nV
      s = 0.0d0
      do k = 1, 3
        e = m%eov(k,vx)
        s = s + (dble(m%tsgn(k,vx)) * f(e)) * m%dc(e)
      end do
      r(vx) = (-s) / m%aT(vx)
    end do
  end subroutine curl

  ! scalar_c2e_interp_operator (operators_mod.F90:93-104)
  subroutine cell_to_edge(m, f, fe)
    type(cpu_mesh_t), intent(in) :: m
    real(8), intent(in), contiguous :: f(:)
    real(8), intent(out), contiguous :: fe(:)
    integer :: e
    !$omp parallel do
    do e = 1, m%nE
      fe(e) = 0.5d0 * (f(m%coe(1,e)) + f(m%coe(2,e)))
    end do
  end subroutine cell_to_edge

  ! scalar_c2v_interp_operator (operators_mod.F90:106-123)
  subroutine cell_to_vertex(m, f, fv)
    type(cpu_mesh_t), intent(in) :: m
    real(8), intent(in), contiguous :: f(:)
    real(8), intent(out), contiguous :: fv(:)
    integer :: vx, k
    real(8) :: s
    !$omp parallel do private(k, s)
    do vx = 1, m%nV
      s = 0.0d0
      do k = 1, 3
        s = s + f(m%cov(k,vx)) * m%kite(k,vx)
      end do
      fv(vx) = s / m%aT(vx)
    end do
  end subroutine cell_to_vertex

  ! scalar_v2c_interp_operator (operators_mod.F90:125-141): a scatter over
  ! the vertices' cells, serial in vertex order as the reference's loop
  subroutine vertex_to_cell(m, fv, f)
    type(cpu_mesh_t), intent(in) :: m
    real(8), intent(in), contiguous :: fv(:)
    real(8), intent(out), contiguous :: f(:)
    integer :: vx, k, i
    f = 0.0d0
    do vx = 1, m%nV
      do k = 1, 3
        i = m%cov(k,vx)
        f(i) = f(i) + (fv(vx) * m%kite(k,vx)) / m%aC(i)
      end do
    end do
  end subroutine vertex_to_cell

  ! calc_kinetic_energy (operators_mod.F90:206-221)
  subroutine kinetic_energy(m, ue, kc)
    type(cpu_mesh_t), intent(in) :: m
    real(8), intent(in), contiguous :: ue(:)
    real(8), intent(out), contiguous :: kc(:)
    integer :: i, k, e
    real(8) :: s
    !$omp parallel do private(k, e, s)
    do i = 1, m%nC
      s = 0.0d0
      do k = 1, m%nEoC(i)
        e = m%eoc(k,i)
        s = s + (ue(e) * ue(e)) * m%aE(e)
      end do
      kc(i) = (s * 0.25d0) / m%aC(i)
    end do
  end subroutine kinetic_energy

  ! calc_tangent_wind (operators_mod.F90:223-237)
  subroutine tangent_wind(m, ue, ve)
    type(cpu_mesh_t), intent(in) :: m
    real(8), intent(in), contiguous :: ue(:)
    real(8), intent(out), contiguous :: ve(:)
    integer :: e, k
    real(8) :: s
    !$omp parallel do private(k, s)
    do e = 1, m%nE
      s = 0.0d0
      do k = 1, m%nEoE(e)
        s = s + ue(m%eoe(k,e)) * m%woe(k,e)
      end do
      ve(e) = s
    end do
  end subroutine tangent_wind

  ! calc_tangent_vor_flux (operators_mod.F90:239-260)
  subroutine tangent_vor_flux(m, ue, ge, pe, fl)
    type(cpu_mesh_t), intent(in) :: m
    real(8), intent(in), contiguous :: ue(:), ge(:), pe(:)
    real(8), intent(out), contiguous :: fl(:)
    integer :: e, k, j
    real(8) :: s
    !$omp parallel do private(k, j, s)
    do e = 1, m%nE
      s = 0.0d0
      do k = 1, m%nEoE(e)
        j = m%eoe(k,e)
        s = s + (((m%woe(k,e) * ue(j)) * ge(j)) * 0.5d0) * (pe(e) + pe(j))
      end do
      fl(e) = s
    end do
  end subroutine tangent_vor_flux

  ! calc_pv_on_edge, scheme 3 (operators_mod.F90:359-412) at orders 2-4
  subroutine pv_on_edge(m, order, ve, pv, pe)
    type(cpu_mesh_t), intent(in) :: m
    integer, intent(in) :: order
    real(8), intent(in), contiguous :: ve(:), pv(:)
    real(8), intent(out), contiguous :: pe(:)
    integer :: e, v1, v2, j
    real(8) :: d1, d2, coef2, dv
    coef2 = merge(1.0d0, 0.0d0, order == 3)
    !$omp parallel do private(v1, v2, j, d1, d2, dv)
    do e = 1, m%nE
      v1 = m%voe(1,e)
      v2 = m%voe(2,e)
      pe(e) = 0.5d0 * (pv(v1) + pv(v2))
      if (order >= 3) then
        d1 = 0.0d0
        do j = 2, m%nfit(v1)
          d1 = d1 + m%d2v(j,1,e) * (pv(m%fit(j,v1)) - pv(v1))
        end do
        d2 = 0.0d0
        do j = 2, m%nfit(v2)
          d2 = d2 + m%d2v(j,2,e) * (pv(m%fit(j,v2)) - pv(v2))
        end do
        dv = m%dv(e)
        pe(e) = pe(e) - ((dv * dv) * ((d1 + d2) - (sign(1.0d0, ve(e)) * coef2) * (d2 - d1))) / 12.0d0
      end if
    end do
  end subroutine pv_on_edge

  ! spatial_operators (tmcore_swm_mod.F90:92-125) without the vector
  ! reconstruction of :101, which no tendency reads
  subroutine stage(m, order, gd, iap, ghs, nT, q, tend_gd, tend_iap, tend_q)
    type(cpu_mesh_t), intent(in) :: m
    integer, intent(in) :: order, nT
    real(8), intent(in), contiguous :: gd(:), iap(:), ghs(:), q(:,:)
    real(8), intent(out), contiguous :: tend_gd(:), tend_iap(:), tend_q(:,:)
    integer :: e, i, t
    real(8) :: sg, dEdx
    call cell_to_edge(m, gd, gd_e)
    !$omp parallel do
    do e = 1, m%nE
      u(e) = iap(e) / sqrt(gd_e(e))
    end do
    !$omp parallel do
    do e = 1, m%nE
      flx(e) = u(e) * gd_e(e)
    end do
    call divergence(m, flx, tend_gd)
    call cell_to_edge(m, tend_gd, tgd_e)
    call cell_to_vertex(m, tend_gd, tgd_v)
    call cell_to_vertex(m, gd, gd_v)
    call curl(m, u, vor)
    call tangent_wind(m, u, v)
    call kinetic_energy(m, u, ke)
    !$omp parallel do
    do i = 1, m%nV
      pv_v(i) = (m%fV(i) + vor(i)) / gd_v(i)
    end do
    call vertex_to_cell(m, pv_v, pv_c)
    call pv_on_edge(m, order, v, pv_v, pv_e)
    call tangent_vor_flux(m, u, gd_e, pv_e, pv_flx)
    ! calc_u_tend_on_edge (tmcore_swm_mod.F90:155-201, no 4th-order flux)
    !$omp parallel do
    do i = 1, m%nC
      en(i) = (ke(i) + gd(i)) + ghs(i)
    end do
    !$omp parallel do private(sg, dEdx)
    do e = 1, m%nE
      sg = sqrt(gd_e(e))
      dEdx = (en(m%coe(2,e)) - en(m%coe(1,e))) / m%dc(e)
      tend_iap(e) = sg * (pv_flx(e) - dEdx) + ((0.5d0 * u(e)) / sg) * tgd_e(e)
    end do
    ! adv_scheme_sg11_run at adv_order 2: c2e, flux, div_operator
    do t = 1, nT
      call cell_to_edge(m, q(:,t), q_e)
      !$omp parallel do
      do e = 1, m%nE
        flx(e) = u(e) * q_e(e)
      end do
      call divergence(m, flx, tend_q(:,t))
    end do
  end subroutine stage

  ! inner_product_tend (operators_mod.F90:187-194): serial sums
  real(8) function inner(m, ag, ai, bg, bi)
    type(cpu_mesh_t), intent(in) :: m
    real(8), intent(in), contiguous :: ag(:), ai(:), bg(:), bi(:)
    real(8) :: s1, s2
    integer :: i
    s1 = 0.0d0
    do i = 1, m%nE
      s1 = s1 + (ai(i) * bi(i)) * m%aE(i)
    end do
    s2 = 0.0d0
    do i = 1, m%nC
      s2 = s2 + (ag(i) * bg(i)) * m%aC(i)
    end do
    inner = (s1 + s2) / m%total_area
  end function inner

  ! the post-step refresh and diag_run (tmcore_swm_mod.F90:78-85, 215-255):
  ! mass, total energy, potential enstrophy, mean absolute vorticity
  subroutine invariants(m, gd, iap, ghs, out)
    type(cpu_mesh_t), intent(in) :: m
    real(8), intent(in), contiguous :: gd(:), iap(:), ghs(:)
    real(8), intent(out) :: out(4)
    real(8) :: mass, e1, e2, ens, av, at
    integer :: i, e
    call cell_to_edge(m, gd, gd_e)
    !$omp parallel do
    do e = 1, m%nE
      u(e) = iap(e) / sqrt(gd_e(e))
    end do
    call cell_to_vertex(m, gd, gd_v)
    call curl(m, u, vor)
    !$omp parallel do
    do i = 1, m%nV
      pv_v(i) = (m%fV(i) + vor(i)) / gd_v(i)
    end do
    mass = 0.0d0; e1 = 0.0d0; e2 = 0.0d0; ens = 0.0d0; av = 0.0d0; at = 0.0d0
    do i = 1, m%nC
      mass = mass + gd(i) * m%aC(i)
    end do
    do e = 1, m%nE
      e1 = e1 + (iap(e) * iap(e)) * m%aE(e)
    end do
    do i = 1, m%nC
      e2 = e2 + ((gd(i) * gd(i)) + (2.0d0 * gd(i)) * ghs(i)) * m%aC(i)
    end do
    do i = 1, m%nV
      ens = ens + (gd_v(i) * (pv_v(i) * pv_v(i))) * m%aT(i)
    end do
    do i = 1, m%nV
      av = av + (vor(i) + m%fV(i)) * m%aT(i)
    end do
    do i = 1, m%nV
      at = at + m%aT(i)
    end do
    out(1) = mass
    out(2) = e1 + e2
    out(3) = ens
    out(4) = av / at
  end subroutine invariants

end module tm_swm_cpu_mod
