"""TEST INFRASTRUCTURE: FP64 numpy restatement of the learned closure.

Only tests/ import this module, as the checker of
the HIP closure (tmcore_amd/csrc/gpu/mlp.hip, closure_kernels.h).  The closure
has no counterpart in the reference; what it follows is the definition in
include/tmcore_gpu.h:
  * the MLP 40 -> 128 -> 128 -> 128 -> 2 with exact GELU x*Phi(x)
    (Hendrycks & Gimpel 2016), forward and backward;
  * optionally the rounding points of the device path (weights,
    activations, dY, dZ rounded to nearest-even BF16, products accumulated
    exactly here, in FP32 on the device) and its forward GELU, the chord
    interpolant of GELU on 1024 equal intervals of [-6, 6] (gelu_chord);
  * the 40 inputs of an edge from the stage fields, the state correction and
    the training loss with its gradient with respect to the outputs.
Pinned by: torch float64 autograd (forward, backward), torch's bfloat16
conversion (rounding), the interpolation error bound h^2/8 max|GELU''| and
exactness at the nodes (gelu_chord), central finite differences (loss gradient), the
edge-orientation invariance of the inputs and exact mass conservation of the
correction (tests/test_closure_oracle.py).
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

NIN, NHID, NOUT = 40, 128, 2
SHAPES = (("W1", (NIN, NHID)), ("b1", (NHID,)), ("W2", (NHID, NHID)), ("b2", (NHID,)),
          ("W3", (NHID, NHID)), ("b3", (NHID,)), ("W4", (NHID, NOUT)), ("b4", (NOUT,)))
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + n], np.float64).reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SHAPES])


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.ascontiguousarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


GELU_N, GELU_LO, GELU_HI = 1024, -6.0, 6.0


def gelu_chord(z):
    """The device's forward GELU: on interval j = clamp(trunc((z + 6) / h),
    0, N - 1) of [-6, 6] (h = 12 / N), the line through GELU at its two
    ends, continued beyond the end intervals."""
    h = (GELU_HI - GELU_LO) / GELU_N
    x0 = GELU_LO + np.arange(GELU_N) * h
    g0, g1 = gelu(x0), gelu(x0 + h)
    b = (g1 - g0) / h
    a = g0 - b * x0
    z = np.asarray(z, np.float64)
    t = np.trunc((z - GELU_LO) / h)
    j = np.clip(np.nan_to_num(t, nan=0.0), 0, GELU_N - 1).astype(np.int64)
    return a[j] + b[j] * z


def gelu_d(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(params, X, bf=False):
    """Y[n,2] and the cache for backward.  bf=True rounds where the device
    does and takes the device's GELU (gelu_chord)."""
    P = unpack(params)
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    A = r(X)
    cache = {"A": [A], "Z": []}
    for k in (1, 2, 3):
        Z = A @ r(P[f"W{k}"]) + P[f"b{k}"]
        H = gelu_chord(Z) if bf else gelu(Z)
        cache["Z"].append(Z)
        A = r(H)
        cache["A"].append(A)
    Y = A @ r(P["W4"]) + P["b4"]
    return Y, cache


def backward(params, X, dY, bf=False):
    """Gradient of sum(dY * Y) w.r.t. the flat parameters and X."""
    P = unpack(params)
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    _, cache = forward(params, X, bf)
    A, Z = cache["A"], cache["Z"]
    g = {}
    D = r(dY)
    g["W4"] = A[3].T @ D
    g["b4"] = D.sum(0)
    dH = D @ r(P["W4"]).T
    for k in (3, 2, 1):
        dZ = r(dH * r(gelu_d(Z[k - 1])))
        g[f"W{k}"] = A[k - 1].T @ dZ
        g[f"b{k}"] = dZ.sum(0)
        dH = dZ @ r(P[f"W{k}"]).T
    return pack(g), dH


# ----------------------------------------------------------------- closure
def kappa(mesh):
    """+1 when the normal cellsOnEdge(1) -> (2) points east (north if exactly
    meridional), -1 otherwise."""
    coe = mesh.cellsOnEdge - 1
    c1, c2 = coe[:, 0], coe[:, 1]
    nx = mesh.xCell[c2] - mesh.xCell[c1]
    ny = mesh.yCell[c2] - mesh.yCell[c1]
    nz = mesh.zCell[c2] - mesh.zCell[c1]
    de = -nx * mesh.yEdge + ny * mesh.xEdge
    return np.where(de > 0, 1.0, np.where(de < 0, -1.0, np.where(nz >= 0, 1.0, -1.0)))


def orient(mesh, e, j):
    """nSign(e, c) * nSign(j, c) for the cell c that edges e and j share
    (0-based edge ids; nSign = +1 when c = cellsOnEdge(1, .))."""
    coe = mesh.cellsOnEdge - 1
    a, b = coe[e, 0], coe[e, 1]
    c, d = coe[j, 0], coe[j, 1]
    on_a = (a == c) | (a == d)
    shared = np.where(on_a, a, b)
    return np.where(on_a, 1.0, -1.0) * np.where(shared == c, 1.0, -1.0)


def slot_edges(mesh):
    """[nE, 10] 0-based input edges: the edge, then the west cell's other edges
    counter-clockwise from it, then the east cell's (9 of up to 10)."""
    nE = mesh.nEdges
    kap = kappa(mesh)
    n = mesh.nEdgesOnEdge.astype(np.int64)
    n1 = mesh.nEdgesOnCell[mesh.cellsOnEdge[:, 0] - 1].astype(np.int64) - 1
    nb = np.empty((nE, 10), np.int64)
    nb[:, 0] = np.arange(nE)
    for s in range(1, 10):
        q = s - 1
        slot = np.where(kap > 0, q, (q + n1) % n)
        j = mesh.edgesOnEdge[np.arange(nE), slot] - 1
        nb[:, s] = np.where(q < n, j, np.arange(nE))
    return nb


def divergence(mesh, u, radius):
    """Velocity divergence of the cells, (1/A_c) sum_i nSign(i,c) u(e_i) dv(e_i)."""
    dv = mesh.dvEdge * radius
    A = mesh.areaCell * radius ** 2
    coe = mesh.cellsOnEdge - 1
    s = np.bincount(coe[:, 0], weights=u * dv, minlength=mesh.nCells) \
        - np.bincount(coe[:, 1], weights=u * dv, minlength=mesh.nCells)
    return s / A


def features(mesh, u, gd_edge, vor_vertex, div_cell, in_shift, in_scale):
    """The 40 normalised inputs of every edge, [nEdges, 40] float32: normal
    velocity (canonical frame), thickness, vorticity and divergence averaged
    to the edge, on the edge and its 9 neighbours."""
    nE = mesh.nEdges
    e = np.arange(nE)
    nb = slot_edges(mesh)
    kap = kappa(mesh)
    c1 = mesh.cellsOnEdge[:, 0] - 1
    c2 = mesh.cellsOnEdge[:, 1] - 1
    v1 = mesh.verticesOnEdge[:, 0] - 1
    v2 = mesh.verticesOnEdge[:, 1] - 1
    vor_e = 0.5 * (vor_vertex[v1] + vor_vertex[v2])
    div_e = 0.5 * (div_cell[c1] + div_cell[c2])
    f = np.empty((nE, 40))
    for s in range(10):
        j = nb[:, s]
        f[:, s] = kap * orient(mesh, e, j) * u[j]
        f[:, 10 + s] = gd_edge[j]
        f[:, 20 + s] = vor_e[j]
        f[:, 30 + s] = div_e[j]
    return ((f - np.asarray(in_shift, np.float64)) * np.asarray(in_scale, np.float64)).astype(np.float32)


def flux_div(mesh, y1, s1, radius):
    """P_c = -(1/A_c) sum_i nSign(i,c) F(e_i) dv(e_i), F = kappa s1 y1 along the
    mesh normal, radius-scaled geometry."""
    dv = mesh.dvEdge * radius
    A = mesh.areaCell * radius ** 2
    F = kappa(mesh) * s1 * np.asarray(y1, np.float64) * dv
    coe = mesh.cellsOnEdge - 1
    s = np.zeros(mesh.nCells)
    np.add.at(s, coe[:, 0], F)
    np.add.at(s, coe[:, 1], -F)
    return -s / A


def apply(mesh, gd, iap, gd_edge_start, Y, out_scale, dt, radius):
    """State after the CRK4 update -> corrected state (include/tmcore_gpu.h)."""
    du = kappa(mesh) * out_scale[0] * np.asarray(Y[:, 0], np.float64)
    iap2 = iap + dt * np.sqrt(gd_edge_start) * du
    gd2 = gd + dt * flux_div(mesh, Y[:, 1], out_scale[1], radius)
    return gd2, iap2


def loss_and_dy(mesh, Y, t_u, t_gd, out_scale, tg_scale, radius):
    """(velocity loss, thickness loss, dL/dY [nE, 2])."""
    Y = np.asarray(Y, np.float64)
    kap = kappa(mesh)
    d = Y[:, 0] - kap * t_u / out_scale[0]
    lu = np.mean(d * d)
    r = (flux_div(mesh, Y[:, 1], out_scale[1], radius) - t_gd) / tg_scale
    lg = np.mean(r * r)
    nE, nC = mesh.nEdges, mesh.nCells
    dv = mesh.dvEdge * radius
    A = mesh.areaCell * radius ** 2
    coe = mesh.cellsOnEdge - 1
    dY = np.zeros((nE, 2))
    dY[:, 0] = 2.0 * d / nE
    g = kap * 2.0 * out_scale[1] * dv / (nC * tg_scale)
    dY[:, 1] = g * (r[coe[:, 1]] / A[coe[:, 1]] - r[coe[:, 0]] / A[coe[:, 0]])
    return lu, lg, dY
