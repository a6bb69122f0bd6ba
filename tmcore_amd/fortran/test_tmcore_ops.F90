! The operators of operators_mod.F90 (reference names, argument order and
! assumed-shape dummies, tm_swm_gpu_mod) on arrays read from a raw file:
!   test_tmcore_ops namelist inputs outputs
! inputs:  f_cell(nC) f_edge(nE) f_vertex(nV) u(nE) v(nE) gd_edge(nE)
!          gd_tend_vertex(nV) pv_vertex(nV) pv_cell(nC) ke(nC) gd_cell(nC)
!          ghs(nC) pv_flx(nE) gd_tend_edge(nE), FP64, native byte order;
! outputs: div(nC) curl(nV) c2e(nE) c2v(nV) v2c(nC) v2e(nE) ke(nC)
!          tangent_wind(nE) tangent_vor_flux(nE) upwind(nE) pv_vertex(nV)
!          pv_edge(nE) u_tend(nE) iap_u_tend(nE).
program test_tmcore_ops

  use tm_swm_gpu_mod

  implicit none

  character(512) :: nml, fin, fout
  real(8), allocatable :: fc(:), fe(:), fv(:), ue(:), ve(:), gde(:), tgv(:), pvv(:), pvc(:)
  real(8), allocatable :: kec(:), gdc(:), ghsc(:), pvf(:), tge(:)
  real(8), allocatable :: o_div(:), o_curl(:), o_c2e(:), o_c2v(:), o_v2c(:), o_v2e(:), o_ke(:), o_tw(:)
  real(8), allocatable :: o_tvf(:), o_upw(:), o_pvv(:), o_pve(:), o_ut(:), o_it(:)
  integer :: io

  call get_command_argument(1, nml)
  call get_command_argument(2, fin)
  call get_command_argument(3, fout)
  call tm_swm_init(trim(nml))

  allocate(fc(nCells), fe(nEdges), fv(nVertices), ue(nEdges), ve(nEdges), gde(nEdges), tgv(nVertices))
  allocate(pvv(nVertices), pvc(nCells), kec(nCells), gdc(nCells), ghsc(nCells), pvf(nEdges), tge(nEdges))
  open(newunit=io, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
  read(io) fc, fe, fv, ue, ve, gde, tgv, pvv, pvc, kec, gdc, ghsc, pvf, tge
  close(io)

  allocate(o_div(nCells), o_curl(nVertices), o_c2e(nEdges), o_c2v(nVertices), o_v2c(nCells), o_v2e(nEdges))
  allocate(o_ke(nCells), o_tw(nEdges), o_tvf(nEdges), o_upw(nEdges), o_pvv(nVertices), o_pve(nEdges))
  allocate(o_ut(nEdges), o_it(nEdges))
  call div_operator(fe, o_div)
  call curl_operator(fe, o_curl)
  call scalar_c2e_interp_operator(fc, o_c2e)
  call scalar_c2v_interp_operator(fc, o_c2v)
  call scalar_v2c_interp_operator(fv, o_v2c)
  call scalar_v2e_interp_operator(fv, o_v2e)
  call calc_kinetic_energy(ue, o_ke)
  call calc_tangent_wind(ue, o_tw)
  call calc_tangent_vor_flux(ue, gde, fe, o_tvf)
  o_upw = fe
  call add_upwind_correction_on_cell(fc, ue, o_upw)
  call calc_pv_on_vertex(fv, abs(pvv) + 1, o_pvv)
  call calc_pv_on_edge(ue, ve, gde, tgv, pvv, pvc, o_pve)
  call calc_u_tend_on_edge(ue, kec, gdc, gde, ghsc, pvf, tge, o_ut, o_it)

  open(newunit=io, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
  write(io) o_div, o_curl, o_c2e, o_c2v, o_v2c, o_v2e, o_ke, o_tw, o_tvf, o_upw, o_pvv, o_pve, o_ut, o_it
  close(io)
  call tm_swm_final()
  write(*, '(A)') 'test_tmcore_ops: done'

end program test_tmcore_ops
