"""Mesh generator: element counts, MPAS conventions and TRiSK weight identities.

The reference only reads MPAS meshes (reference src/core/mesh_mod.F90:79-167);
these tests pin the generated arrays by what the mathematics fixes.
"""
import numpy as np
import pytest

from tmcore_amd.mesh import Mesh


def frames(m):
    xe = np.stack([m.xEdge, m.yEdge, m.zEdge], 1)
    xc = np.stack([m.xCell, m.yCell, m.zCell], 1)
    xv = np.stack([m.xVertex, m.yVertex, m.zVertex], 1)
    c1, c2 = m.cellsOnEdge[:, 0] - 1, m.cellsOnEdge[:, 1] - 1
    d = xc[c2] - xc[c1]
    d -= (d * xe).sum(1)[:, None] * xe
    n = d / np.linalg.norm(d, axis=1)[:, None]
    return xe, xc, xv, n, np.cross(xe, n)


@pytest.mark.parametrize("level", [0, 1, 2, 3, 4])
def test_counts_and_euler(level):
    m = Mesh(level)
    assert (m.nCells, m.nEdges, m.nVertices) == (10 * 4**level + 2, 30 * 4**level, 20 * 4**level)
    assert m.nCells - m.nEdges + m.nVertices == 2
    assert np.bincount(m.nEdgesOnCell)[5] == 12  # twelve pentagons
    assert m.nEdgesOnCell.sum() == 2 * m.nEdges


@pytest.mark.parametrize("level", [1, 3, 5])
def test_areas_and_kites_partition(level):
    m = Mesh(level)
    assert abs(m.areaCell.sum() - 4 * np.pi) < 1e-12
    assert abs(m.areaTriangle.sum() - 4 * np.pi) < 1e-12
    assert m.kiteAreasOnVertex.min() > 0
    np.testing.assert_allclose(m.kiteAreasOnVertex.sum(1), m.areaTriangle, rtol=1e-13)
    acc = np.zeros(m.nCells)
    np.add.at(acc, (m.cellsOnVertex - 1).ravel(), m.kiteAreasOnVertex.ravel())
    np.testing.assert_allclose(acc, m.areaCell, rtol=1e-12)


def test_orientation_conventions():
    m = Mesh(3)
    xe, xc, xv, n, t = frames(m)
    v1, v2 = m.verticesOnEdge[:, 0] - 1, m.verticesOnEdge[:, 1] - 1
    # tangent k x n points from verticesOnEdge(1) to verticesOnEdge(2)
    assert ((xv[v2] - xv[v1]) * t).sum(1).min() > 0
    # cellsOnCell(i) lies across edgesOnCell(i)
    for c in range(m.nCells):
        for i in range(m.nEdgesOnCell[c]):
            e = m.edgesOnCell[c, i] - 1
            assert set(m.cellsOnEdge[e]) == {c + 1, m.cellsOnCell[c, i]}
    # counter-clockwise cells: consecutive neighbour directions turn left
    for c in range(0, m.nCells, 7):
        k = m.nEdgesOnCell[c]
        p = xc[c]
        for i in range(k):
            a = xc[m.cellsOnCell[c, i] - 1] - p
            b = xc[m.cellsOnCell[c, (i + 1) % k] - 1] - p
            assert np.dot(np.cross(a, b), p) > 0
    # edgesOnVertex(i) joins cellsOnVertex(i) and cellsOnVertex(i+1)
    for v in range(m.nVertices):
        for i in range(3):
            e = m.edgesOnVertex[v, i] - 1
            assert set(m.cellsOnEdge[e]) == {m.cellsOnVertex[v, i], m.cellsOnVertex[v, (i + 1) % 3]}


def test_trisk_antisymmetry():
    m = Mesh(3)
    A = m.dvEdge * m.dcEdge
    for e in range(m.nEdges):
        for s in range(m.nEdgesOnEdge[e]):
            j = m.edgesOnEdge[e, s] - 1
            k = list(m.edgesOnEdge[j, :m.nEdgesOnEdge[j]] - 1).index(e)
            assert abs(A[e] * m.weightsOnEdge[e, s] + A[j] * m.weightsOnEdge[j, k]) < 1e-15 * A[e]


def _tangent(m, u):
    idx = np.maximum(m.edgesOnEdge - 1, 0)
    return (m.weightsOnEdge * u[idx]).sum(1)


def test_trisk_curl_of_perp_is_kite_divergence():
    """curl(W u) at a vertex equals minus the kite-weighted divergence of u
    (Thuburn et al. 2009): holds to round-off for any u only with correct
    weights, orderings and signs."""
    m = Mesh(3)
    rng = np.random.default_rng(3)
    u = rng.standard_normal(m.nEdges)
    w = _tangent(m, u)
    c1 = m.cellsOnEdge[:, 0] - 1
    eoc = np.maximum(m.edgesOnCell - 1, 0)
    nsign = np.where(c1[eoc] == np.arange(m.nCells)[:, None], 1.0, -1.0)
    mask = np.arange(6)[None, :] < m.nEdgesOnCell[:, None]
    div = (nsign * u[eoc] * m.dvEdge[eoc] * mask).sum(1) / m.areaCell
    eov = m.edgesOnVertex - 1
    tsign = np.where(m.verticesOnEdge[eov, 0] - 1 == np.arange(m.nVertices)[:, None], 1.0, -1.0)
    curl = -(tsign * w[eov] * m.dcEdge[eov]).sum(1) / m.areaTriangle
    kdiv = (m.kiteAreasOnVertex * div[m.cellsOnVertex - 1]).sum(1) / m.areaTriangle
    np.testing.assert_allclose(curl, -kdiv, atol=1e-12 * np.abs(kdiv).max())


def test_trisk_tangent_reconstruction_converges_in_rms():
    errs = []
    for L in (3, 4, 5):
        m = Mesh(L)
        xe, _, _, n, t = frames(m)
        V = np.stack([-xe[:, 1], xe[:, 0], 0 * xe[:, 0]], 1) * 20.0
        err = _tangent(m, (V * n).sum(1)) - (V * t).sum(1)
        errs.append(np.sqrt((err**2).mean()))
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 0.02 * 20.0
