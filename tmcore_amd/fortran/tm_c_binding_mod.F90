! iso_c_binding declarations of the C-ABI in include/tmcore_host.h and
! include/tmcore_gpu.h.  Host arrays keep the reference's Fortran layout
! (1-based int32 connectivity, (slot, n) column-major), so they are passed as
! they are allocated here.
module tm_c_binding_mod

  use iso_c_binding
  use tm_desc_mod

  implicit none

  integer(c_int), parameter :: TM_OK = 0

  type, bind(c) :: tm_opts
    integer(c_int) :: time_scheme, pv_scheme, interp_pv_order, adv_order, adv_mono, conserve_energy
    real(c_double) :: apvm_weight, dt, radius
    integer(c_int) :: ntracers, precision, device, renumber, geometry_scaled
  end type tm_opts

  interface

    integer(c_int) function tm_mesh_icos_dims(level, nC, nE, nV) bind(c, name='tm_mesh_icos_dims')
      import :: c_int, c_int32_t
      integer(c_int), value :: level
      integer(c_int32_t) :: nC, nE, nV
    end function tm_mesh_icos_dims

    integer(c_int) function tm_mesh_icos_generate(level, m) bind(c, name='tm_mesh_icos_generate')
      import :: c_int, tm_mesh_desc
      integer(c_int), value :: level
      type(tm_mesh_desc) :: m
    end function tm_mesh_icos_generate

    integer(c_int) function tm_mesh_mpas_dims(path, dims, sphere_radius) bind(c, name='tm_mesh_mpas_dims')
      import :: c_int, c_int32_t, c_double, c_char
      character(kind=c_char) :: path(*)
      integer(c_int32_t) :: dims(6)
      real(c_double) :: sphere_radius
    end function tm_mesh_mpas_dims

    integer(c_int) function tm_mesh_mpas_read(path, m) bind(c, name='tm_mesh_mpas_read')
      import :: c_int, c_char, tm_mesh_desc
      character(kind=c_char) :: path(*)
      type(tm_mesh_desc) :: m
    end function tm_mesh_mpas_read

    type(c_ptr) function tm_mesh_mpas_error() bind(c, name='tm_mesh_mpas_error')
      import :: c_ptr
    end function tm_mesh_mpas_error

    integer(c_int) function tm_poly_fit(m, radius, f) bind(c, name='tm_poly_fit')
      import :: c_int, c_double, tm_mesh_desc, tm_fit_desc
      type(tm_mesh_desc) :: m
      real(c_double), value :: radius
      type(tm_fit_desc) :: f
    end function tm_poly_fit

    integer(c_int) function tm_case_seeded(m, radius, seed, ntracers, tracer_seeds, gd, u, ghs, tracers) &
        bind(c, name='tm_case_seeded')
      import :: c_int, c_int64_t, c_double, c_ptr, tm_mesh_desc
      type(tm_mesh_desc) :: m
      real(c_double), value :: radius
      integer(c_int64_t), value :: seed
      integer(c_int), value :: ntracers
      integer(c_int64_t) :: tracer_seeds(*)
      real(c_double) :: gd(*), u(*), ghs(*)
      type(c_ptr), value :: tracers
    end function tm_case_seeded

    integer(c_int) function tm_case_steady_geostrophic(m, radius, gd, u, ghs) bind(c, name='tm_case_steady_geostrophic')
      import :: c_int, c_double, tm_mesh_desc
      type(tm_mesh_desc) :: m
      real(c_double), value :: radius
      real(c_double) :: gd(*), u(*), ghs(*)
    end function tm_case_steady_geostrophic

    subroutine tm_opts_default(o) bind(c, name='tm_opts_default')
      import :: tm_opts
      type(tm_opts) :: o
    end subroutine tm_opts_default

    integer(c_int) function tm_ctx_create(ctx, m, f, o) bind(c, name='tm_ctx_create')
      import :: c_int, c_ptr, tm_mesh_desc, tm_fit_desc, tm_opts
      type(c_ptr) :: ctx
      type(tm_mesh_desc) :: m
      type(tm_fit_desc) :: f
      type(tm_opts) :: o
    end function tm_ctx_create

    integer(c_int) function tm_ctx_destroy(ctx) bind(c, name='tm_ctx_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function tm_ctx_destroy

    type(c_ptr) function tm_last_error(ctx) bind(c, name='tm_last_error')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function tm_last_error

    integer(c_int) function tm_state_upload(ctx, gd, u, iap_u, ghs, tracers) bind(c, name='tm_state_upload')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, gd, u, iap_u, ghs, tracers
    end function tm_state_upload

    integer(c_int) function tm_state_download(ctx, gd, u, iap_u, tracers) bind(c, name='tm_state_download')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, gd, u, iap_u, tracers
    end function tm_state_download

    integer(c_int) function tm_step(ctx, nsteps, invariants, beta) bind(c, name='tm_step')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps
      real(c_double) :: invariants(*), beta(*)
    end function tm_step

    integer(c_int) function tm_invariants(ctx, out) bind(c, name='tm_invariants')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: out(4)
    end function tm_invariants

    integer(c_int) function tm_last_elapsed_ms(ctx, ms) bind(c, name='tm_last_elapsed_ms')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: ms
    end function tm_last_elapsed_ms

    integer(c_int) function tm_op_div(ctx, f_in, f_out) bind(c, name='tm_op_div')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: f_in(*), f_out(*)
    end function tm_op_div

    integer(c_int) function tm_op_curl(ctx, f_in, f_out) bind(c, name='tm_op_curl')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: f_in(*), f_out(*)
    end function tm_op_curl

    integer(c_int) function tm_op_c2e(ctx, f_in, f_out) bind(c, name='tm_op_c2e')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: f_in(*), f_out(*)
    end function tm_op_c2e

    integer(c_int) function tm_op_c2v(ctx, f_in, f_out) bind(c, name='tm_op_c2v')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: f_in(*), f_out(*)
    end function tm_op_c2v

    integer(c_int) function tm_op_v2c(ctx, f_in, f_out) bind(c, name='tm_op_v2c')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: f_in(*), f_out(*)
    end function tm_op_v2c

    integer(c_int) function tm_op_v2e(ctx, f_in, f_out) bind(c, name='tm_op_v2e')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: f_in(*), f_out(*)
    end function tm_op_v2e

    integer(c_int) function tm_op_ke(ctx, f_in, f_out) bind(c, name='tm_op_ke')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: f_in(*), f_out(*)
    end function tm_op_ke

    integer(c_int) function tm_op_tangent_wind(ctx, f_in, f_out) bind(c, name='tm_op_tangent_wind')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: f_in(*), f_out(*)
    end function tm_op_tangent_wind

    integer(c_int) function tm_op_tangent_vor_flux(ctx, u, gd, pv, flux) bind(c, name='tm_op_tangent_vor_flux')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: u(*), gd(*), pv(*), flux(*)
    end function tm_op_tangent_vor_flux

    integer(c_int) function tm_op_upwind_cell(ctx, f_cell, u_edge, f_edge) bind(c, name='tm_op_upwind_cell')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: f_cell(*), u_edge(*), f_edge(*)
    end function tm_op_upwind_cell

    integer(c_int) function tm_op_pv_vertex(ctx, vor_vertex, gd_vertex, pv_vertex) bind(c, name='tm_op_pv_vertex')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: vor_vertex(*), gd_vertex(*), pv_vertex(*)
    end function tm_op_pv_vertex

    integer(c_int) function tm_op_pv_edge(ctx, u_edge, v_edge, gd_edge, gd_tend_vertex, pv_vertex, pv_cell, pv_edge) &
        bind(c, name='tm_op_pv_edge')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: u_edge(*), v_edge(*), gd_edge(*), gd_tend_vertex(*), pv_vertex(*), pv_cell(*), pv_edge(*)
    end function tm_op_pv_edge

    integer(c_int) function tm_op_u_tend(ctx, u_edge, ke_cell, gd_cell, gd_edge, ghs_cell, pv_flx_edge, gd_tend_edge, &
        u_tend_edge, iap_u_tend_edge) bind(c, name='tm_op_u_tend')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: u_edge(*), ke_cell(*), gd_cell(*), gd_edge(*), ghs_cell(*), pv_flx_edge(*), gd_tend_edge(*)
      real(c_double) :: u_tend_edge(*), iap_u_tend_edge(*)
    end function tm_op_u_tend

    ! learned closure (include/tmcore_gpu.h): the per-edge MLP and its attachment
    integer(c_int) function tm_mlp_create(mlp, device, seed) bind(c, name='tm_mlp_create')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr) :: mlp
      integer(c_int), value :: device
      integer(c_int64_t), value :: seed
    end function tm_mlp_create

    integer(c_int) function tm_mlp_destroy(mlp) bind(c, name='tm_mlp_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: mlp
    end function tm_mlp_destroy

    integer(c_int) function tm_mlp_set_params(mlp, params) bind(c, name='tm_mlp_set_params')
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: mlp
      real(c_float) :: params(*)
    end function tm_mlp_set_params

    integer(c_int) function tm_mlp_set_scales(mlp, in_shift, in_scale, out_scale) bind(c, name='tm_mlp_set_scales')
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: mlp
      real(c_float) :: in_shift(40), in_scale(40), out_scale(2)
    end function tm_mlp_set_scales

    integer(c_int) function tm_closure_attach(ctx, mlp) bind(c, name='tm_closure_attach')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, mlp
    end function tm_closure_attach

    integer(c_int) function tm_closure_set_interval(ctx, nsteps, lumped) bind(c, name='tm_closure_set_interval')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps, lumped
    end function tm_closure_set_interval

    ! history output (include/tmcore_host.h, include/tmcore_gpu.h)
    integer(c_int) function tm_reconstruct_coeffs(m, radius, coeffs) bind(c, name='tm_reconstruct_coeffs')
      import :: c_int, c_double, tm_mesh_desc
      type(tm_mesh_desc) :: m
      real(c_double), value :: radius
      real(c_double) :: coeffs(*)
    end function tm_reconstruct_coeffs

    integer(c_int) function tm_history_setup(ctx, coeffs, latCell, lonCell) bind(c, name='tm_history_setup')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: coeffs(*), latCell(*), lonCell(*)
    end function tm_history_setup

    integer(c_int) function tm_history_fields(ctx, u, zonal, meridional, h, pv, vor, div) &
        bind(c, name='tm_history_fields')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: u(*), zonal(*), meridional(*), h(*), pv(*), vor(*), div(*)
    end function tm_history_fields

    integer(c_int) function tm_history_create(hist, path, m, radius, dt, time_scheme, case_name) &
        bind(c, name='tm_history_create')
      import :: c_int, c_ptr, c_char, c_double, tm_mesh_desc
      type(c_ptr) :: hist
      character(kind=c_char) :: path(*), case_name(*)
      type(tm_mesh_desc) :: m
      real(c_double), value :: radius, dt
      integer(c_int), value :: time_scheme
    end function tm_history_create

    integer(c_int) function tm_history_write(hist, time_seconds, u, zonal, meridional, h, pv, vor, div, &
        total_mass, total_energy) bind(c, name='tm_history_write')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: hist
      real(c_double), value :: time_seconds, total_mass, total_energy
      real(c_double) :: u(*), zonal(*), meridional(*), h(*), pv(*), vor(*), div(*)
    end function tm_history_write

    integer(c_int) function tm_history_close(hist) bind(c, name='tm_history_close')
      import :: c_int, c_ptr
      type(c_ptr), value :: hist
    end function tm_history_close

    integer(c_int) function tm_closure_set_gain(ctx, gain) bind(c, name='tm_closure_set_gain')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: gain
    end function tm_closure_set_gain

    integer(c_size_t) function c_strlen(s) bind(c, name='strlen')
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function c_strlen

  end interface

contains

  ! Error text of the last failed call on ctx, as a Fortran string.
  function tm_error_text(ctx) result(res)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: res
    res = c_text(tm_last_error(ctx))
  end function tm_error_text

  ! A NUL-terminated C string as a Fortran string ('' for NULL).
  function c_text(p) result(res)
    type(c_ptr), intent(in) :: p
    character(len=:), allocatable :: res
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    if (.not. c_associated(p)) then
      res = ''
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: res)
    do i = 1, n
      res(i:i) = chars(i)
    end do
  end function c_text

end module tm_c_binding_mod
