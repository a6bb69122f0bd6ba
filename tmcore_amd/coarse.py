"""Coarse-graining of a fine-mesh shallow-water state onto a coarser
icosahedral mesh (host side of the learned closure's data pipeline).

The fine cells are partitioned twice by nearest generator: among the coarse
cells (the coarse Voronoi cells, for thickness and topography) and among the
coarse edge points (for velocity).  With A_i the fine cell areas:
  gd_c  = sum_{i in c} A_i gd_i / sum_{i in c} A_i             (mass-weighted)
  u_e   = n_e . sum_{i in e} A_i U_i / sum_{i in e} A_i
where U_i is the fine cell's velocity vector from the Perot (2000)
reconstruction U_i = (1/A_i) sum_{e' in i} (x_e' - x_i) s(e', i) u_e' dv_e'
(s = +1 when i = cellsOnEdge(1, e'), the outward normal velocity), projected
on the tangent plane, and n_e the coarse edge normal cellsOnEdge(1) -> (2).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
from scipy.spatial import cKDTree

from .mesh import Mesh


def _xyz(m: Mesh, where: str) -> np.ndarray:
    p = np.stack([getattr(m, f"x{where}"), getattr(m, f"y{where}"), getattr(m, f"z{where}")], 1)
    return p / np.linalg.norm(p, axis=1, keepdims=True)


class CoarseGrainer:
    """Both maps as sparse matrices: gd_c = Mc gd, u_c = Gu u (velocity=False
    builds Mc only, for fine meshes too close to the coarse one to give every
    coarse edge a fine cell)."""

    def __init__(self, fine: Mesh, coarse: Mesh, velocity: bool = True):
        self.fine, self.coarse = fine, coarse
        nCf, nEf, nCc, nEc = fine.nCells, fine.nEdges, coarse.nCells, coarse.nEdges
        fc = _xyz(fine, "Cell")
        cell_of = cKDTree(_xyz(coarse, "Cell")).query(fc, workers=-1)[1]
        A = fine.areaCell
        Ac = np.bincount(cell_of, weights=A, minlength=nCc)
        if np.any(Ac <= 0):
            raise ValueError("a coarse cell received no fine cell: the fine mesh is too coarse")
        cols = np.arange(nCf)
        self.Mc = sp.csr_matrix((A / Ac[cell_of], (cell_of, cols)), shape=(nCc, nCf))
        self.Gu = None
        if not velocity:
            return
        edge_of = cKDTree(_xyz(coarse, "Edge")).query(fc, workers=-1)[1]
        Ae = np.bincount(edge_of, weights=A, minlength=nEc)
        if np.any(Ae <= 0):
            raise ValueError("a coarse edge received no fine cell: the fine mesh is too coarse")
        # coarse edge normals cellsOnEdge(1) -> (2) at the edge points
        cc = _xyz(coarse, "Cell")
        ccoe = coarse.cellsOnEdge - 1
        ce = _xyz(coarse, "Edge")
        n = cc[ccoe[:, 1]] - cc[ccoe[:, 0]]
        n -= np.sum(n * ce, 1, keepdims=True) * ce
        n /= np.linalg.norm(n, axis=1, keepdims=True)
        # Perot reconstruction on the fine mesh, component j: P_j [nCf, nEf]
        coe = fine.cellsOnEdge - 1
        c1, c2 = coe[:, 0], coe[:, 1]
        xe = _xyz(fine, "Edge")
        e_idx = np.arange(nEf)
        r1 = (xe - fc[c1]) * (fine.dvEdge / A[c1])[:, None]
        r2 = -(xe - fc[c2]) * (fine.dvEdge / A[c2])[:, None]
        rows = np.concatenate([c1, c2])
        ecol = np.concatenate([e_idx, e_idx])
        # u_c[e] = sum_i w_ie sum_j m_eij U_ij with w_ie = A_i / Ae, m_eij =
        # n_ej - (n_e . x_i) x_ij (normal component of the tangent projection)
        w = A / Ae[edge_of]
        ndotx = np.sum(n[edge_of] * fc, 1)
        Gu = None
        for j in range(3):
            m_j = n[edge_of, j] - ndotx * fc[:, j]
            W = sp.csr_matrix((w * m_j, (edge_of, cols)), shape=(nEc, nCf))
            P = sp.csr_matrix((np.concatenate([r1[:, j], r2[:, j]]), (rows, ecol)), shape=(nCf, nEf))
            Gu = W @ P if Gu is None else Gu + W @ P
        self.Gu = Gu.tocsr()

    def cell_mean(self, f_cell: np.ndarray) -> np.ndarray:
        return self.Mc @ f_cell

    def edge_velocity(self, u_edge: np.ndarray) -> np.ndarray:
        return self.Gu @ u_edge

    def state(self, gd, u, ghs):
        """(gd, u, ghs) on the coarse mesh."""
        return self.Mc @ gd, self.Gu @ u, self.Mc @ ghs
