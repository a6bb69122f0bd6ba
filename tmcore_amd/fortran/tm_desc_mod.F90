! The mesh and fit descriptors of include/tmcore_host.h as bind(c) types:
! shared by the C-ABI bindings (tm_c_binding_mod) and the Fortran CPU path
! (tm_swm_cpu_mod), which needs no other part of the C-ABI.
module tm_desc_mod

  use iso_c_binding

  implicit none

  type, bind(c) :: tm_mesh_desc
    integer(c_int32_t) :: nCells, nEdges, nVertices, maxEdges, maxEdges2, vertexDegree
    type(c_ptr) :: latCell, lonCell, xCell, yCell, zCell
    type(c_ptr) :: latEdge, lonEdge, xEdge, yEdge, zEdge
    type(c_ptr) :: latVertex, lonVertex, xVertex, yVertex, zVertex
    type(c_ptr) :: nEdgesOnCell, nEdgesOnEdge
    type(c_ptr) :: cellsOnCell, edgesOnCell, verticesOnCell
    type(c_ptr) :: cellsOnEdge, verticesOnEdge
    type(c_ptr) :: edgesOnEdge, weightsOnEdge
    type(c_ptr) :: cellsOnVertex, edgesOnVertex, kiteAreasOnVertex
    type(c_ptr) :: dvEdge, dv1Edge, dv2Edge, dcEdge, angleEdge
    type(c_ptr) :: areaCell, areaTriangle
  end type tm_mesh_desc

  type, bind(c) :: tm_fit_desc
    integer(c_int32_t) :: maxFit
    type(c_ptr) :: nFit2Cells, fit2Cells, deriv2OnCell
    type(c_ptr) :: nFit4Cells, fit4Cells, deriv4OnCell
    type(c_ptr) :: nFit2Vertices, fit2Vertices, deriv2OnVertex
    type(c_ptr) :: nFit4Vertices, fit4Vertices, deriv4OnVertex
  end type tm_fit_desc

end module tm_desc_mod
