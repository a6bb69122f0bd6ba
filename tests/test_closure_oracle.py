"""Pins the FP64 closure oracle (oracle/mlp_oracle.py) before it checks the
device: forward/backward against torch float64 autograd, BF16 rounding
against torch's conversion, the loss gradient against central differences,
orientation invariance of the inputs, and exact mass conservation of the
thickness correction."""
import types

import numpy as np
import pytest
import torch

from oracle import mlp_oracle as mo
from tmcore_amd.mesh import RADIUS, Mesh


def rand_params(rng, scale=1.0):
    d = {}
    for name, shape in mo.SHAPES:
        fan_in = shape[0] if name.startswith("W") else 1
        d[name] = rng.standard_normal(shape) * (scale / np.sqrt(fan_in) if name.startswith("W") else 0.1)
    return mo.pack(d)


def torch_net(params):
    P = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in mo.unpack(params).items()}

    def f(X):
        A = X
        for k in (1, 2, 3):
            A = torch.nn.functional.gelu(A @ P[f"W{k}"] + P[f"b{k}"])  # exact (erf) GELU
        return A @ P["W4"] + P["b4"]
    return P, f


def test_forward_backward_match_torch_autograd():
    rng = np.random.default_rng(11)
    p = rand_params(rng)
    X = rng.standard_normal((37, 40))
    dY = rng.standard_normal((37, 2))
    Y, _ = mo.forward(p, X)
    g, dX = mo.backward(p, X, dY)
    P, f = torch_net(p)
    Xt = torch.tensor(X, requires_grad=True)
    Yt = f(Xt)
    np.testing.assert_allclose(Y, Yt.detach().numpy(), rtol=1e-12, atol=1e-12)
    (Yt * torch.tensor(dY)).sum().backward()
    gt = mo.pack({k: v.grad.numpy() for k, v in P.items()})
    np.testing.assert_allclose(g, gt, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(dX, Xt.grad.numpy(), rtol=1e-10, atol=1e-12)


def test_gelu_special_values():
    # Phi(1) = 0.8413447460685429 (standard normal CDF), GELU(x) -> x and 0 in the tails
    assert abs(mo.gelu(np.array(1.0)) - 0.8413447460685429) < 1e-15
    assert mo.gelu(np.array(0.0)) == 0.0
    assert abs(mo.gelu(np.array(10.0)) - 10.0) < 1e-12
    assert abs(mo.gelu(np.array(-10.0))) < 1e-20
    # derivative against a central difference
    z = np.linspace(-4, 4, 17)
    h = 1e-6
    np.testing.assert_allclose(mo.gelu_d(z), (mo.gelu(z + h) - mo.gelu(z - h)) / (2 * h), atol=1e-8)


def test_gelu_chord_is_the_interpolant_within_its_error_bound():
    """The device's forward GELU (gelu_chord): exact at the 1025 nodes of
    [-6, 6], within the linear-interpolation bound h^2/8 max|GELU''| of GELU
    between them (GELU'' = phi(z)(2 - z^2) peaks at z = 0 with 2 phi(0)), the
    bound attained next to z = 0; beyond +-6 the end lines' slopes differ from
    GELU's by at most 6 phi(5.99) + Phi(-6) < 4e-8, so the error grows at most
    that fast."""
    h = 12.0 / mo.GELU_N
    nodes = -6.0 + h * np.arange(mo.GELU_N + 1)
    assert np.abs(mo.gelu_chord(nodes) - mo.gelu(nodes)).max() <= 1e-15
    z = np.linspace(-6, 6, 1_000_001)
    err = np.abs(mo.gelu_chord(z) - mo.gelu(z))
    bound = h * h / 8 * 2 / np.sqrt(2 * np.pi)
    assert err.max() <= bound * (1 + 1e-9)
    assert err.max() >= 0.99 * bound and abs(z[err.argmax()]) < h
    far = np.concatenate([np.linspace(-50, -6, 1001), np.linspace(6, 50, 1001)])
    assert np.all(np.abs(mo.gelu_chord(far) - mo.gelu(far)) <= 1e-8 + 4e-8 * (np.abs(far) - 6))
    assert np.isnan(mo.gelu_chord(np.array([np.nan]))[0])


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.standard_normal(10000) * 10.0 ** rng.integers(-8, 8, 10000),
                        # exact ties: bf16 value + half an ulp
                        np.array([1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -(2.0 + 2.0 ** -7)])]).astype(np.float32)
    ref = torch.tensor(x).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(mo.bf16(x), ref)


def test_bf16_backward_is_close_to_fp64():
    rng = np.random.default_rng(5)
    p = rand_params(rng)
    X = rng.standard_normal((64, 40))
    dY = rng.standard_normal((64, 2))
    g, dX = mo.backward(p, X, dY)
    gb, dXb = mo.backward(p, X, dY, bf=True)
    assert np.linalg.norm(gb - g) / np.linalg.norm(g) < 3e-2
    assert np.linalg.norm(dXb - dX) / np.linalg.norm(dX) < 3e-2


@pytest.fixture(scope="module")
def m2():
    return Mesh(2)


def test_loss_gradient_matches_central_differences(m2):
    rng = np.random.default_rng(9)
    Y = rng.standard_normal((m2.nEdges, 2))
    t_u = rng.standard_normal(m2.nEdges) * 3e-5
    t_gd = rng.standard_normal(m2.nCells) * 1e-3
    # scales that make both loss terms O(1) on this mesh
    s = (2e-5, 1e-3 * np.mean(m2.areaCell * RADIUS ** 2) / np.mean(m2.dvEdge * RADIUS))
    tgs = 1e-3
    lu, lg, dY = mo.loss_and_dy(m2, Y, t_u, t_gd, s, tgs, RADIUS)
    for _ in range(12):
        e, k = int(rng.integers(m2.nEdges)), int(rng.integers(2))
        h = 1e-5
        Yp, Ym = Y.copy(), Y.copy()
        Yp[e, k] += h
        Ym[e, k] -= h
        fp = sum(mo.loss_and_dy(m2, Yp, t_u, t_gd, s, tgs, RADIUS)[:2])
        fm = sum(mo.loss_and_dy(m2, Ym, t_u, t_gd, s, tgs, RADIUS)[:2])
        assert abs((fp - fm) / (2 * h) - dY[e, k]) <= 1e-6 * max(1e-8, abs(dY[e, k])) + 1e-12


def test_thickness_correction_conserves_mass(m2):
    rng = np.random.default_rng(2)
    Y = rng.standard_normal((m2.nEdges, 2))
    P = mo.flux_div(m2, Y[:, 1], 3.0e7, RADIUS)
    A = m2.areaCell * RADIUS ** 2
    assert abs(np.sum(A * P)) <= 1e-13 * np.sum(np.abs(A * P))


def flipped_mesh(m, flip):
    """The same mesh with the orientation of the edges in `flip` reversed:
    cellsOnEdge swapped and, as the edgesOnEdge convention requires
    (cellsOnEdge(1)'s edges first), the two edgesOnEdge groups swapped."""
    coe = m.cellsOnEdge.copy()
    coe[flip] = coe[flip][:, ::-1]
    eoe = m.edgesOnEdge.copy()
    for e in np.nonzero(flip)[0]:
        n = m.nEdgesOnEdge[e]
        n1 = m.nEdgesOnCell[m.cellsOnEdge[e, 0] - 1] - 1
        eoe[e, :n] = np.concatenate([m.edgesOnEdge[e, n1:n], m.edgesOnEdge[e, :n1]])
    # the vertices of a reversed edge swap too (verticesOnEdge follows the normal)
    voe = m.verticesOnEdge.copy()
    voe[flip] = voe[flip][:, ::-1]
    keys = ("nEdges", "nCells", "nEdgesOnEdge", "nEdgesOnCell", "xCell", "yCell", "zCell", "xEdge", "yEdge",
            "dvEdge", "areaCell")
    return types.SimpleNamespace(cellsOnEdge=coe, edgesOnEdge=eoe, verticesOnEdge=voe,
                                 **{k: getattr(m, k) for k in keys})


def test_inputs_and_correction_do_not_depend_on_edge_orientation(m2):
    """Reversing edges' orientation (cellsOnEdge swapped, u negated) leaves
    every edge's inputs, and the thickness correction, unchanged."""
    rng = np.random.default_rng(4)
    u = rng.standard_normal(m2.nEdges)
    gde = rng.standard_normal(m2.nEdges)
    vor, div = rng.standard_normal(m2.nVertices), rng.standard_normal(m2.nCells)
    sh, sc = rng.standard_normal(40), 1 + rng.random(40)
    x0 = mo.features(m2, u, gde, vor, div, sh, sc)
    flip = rng.random(m2.nEdges) < 0.4
    mf = flipped_mesh(m2, flip)
    x1 = mo.features(mf, np.where(flip, -u, u), gde, vor, div, sh, sc)
    np.testing.assert_array_equal(x0, x1)
    # the divergence itself is orientation independent
    np.testing.assert_allclose(mo.divergence(mf, np.where(flip, -u, u), RADIUS), mo.divergence(m2, u, RADIUS),
                               rtol=1e-12, atol=1e-12 * np.abs(mo.divergence(m2, u, RADIUS)).max())
    # same outputs -> same thickness tendency; the velocity correction flips with the normal
    y1 = rng.standard_normal(m2.nEdges)
    np.testing.assert_allclose(mo.flux_div(mf, y1, 5.0, RADIUS), mo.flux_div(m2, y1, 5.0, RADIUS), rtol=1e-13,
                               atol=1e-13 * np.abs(mo.flux_div(m2, y1, 5.0, RADIUS)).max())
    np.testing.assert_array_equal(mo.kappa(mf), np.where(flip, -1.0, 1.0) * mo.kappa(m2))
    # the orientation of slot 0 (the edge itself) is +1
    assert np.all(mo.orient(m2, np.arange(m2.nEdges), np.arange(m2.nEdges)) == 1.0)
    # planted mistakes are caught: without kappa, or without the group swap
    k0 = mo.kappa
    try:
        mo.kappa = lambda m: np.ones(m.nEdges)
        assert not np.array_equal(mo.features(m2, u, gde, vor, div, sh, sc),
                                  mo.features(mf, np.where(flip, -u, u), gde, vor, div, sh, sc))
    finally:
        mo.kappa = k0
    mf_noswap = flipped_mesh(m2, flip)
    mf_noswap.edgesOnEdge = m2.edgesOnEdge
    assert not np.array_equal(x0, mo.features(mf_noswap, np.where(flip, -u, u), gde, vor, div, sh, sc))


def test_canonical_slots_are_a_permutation_of_edges_on_edge(m2):
    nb = mo.slot_edges(m2)
    for e in range(0, m2.nEdges, 7):
        n = m2.nEdgesOnEdge[e]
        assert set(nb[e, 1:]) <= set(m2.edgesOnEdge[e, :n] - 1)
        assert len(set(nb[e, 1:])) == 9


def test_divergence_matches_the_c_oracle_div_operator(m2):
    """div = -div_operator(u) (the reference's div_operator returns -div F,
    operators_mod.F90:40-55), computed by the independent C oracle."""
    from oracle.oracle import Oracle
    rng = np.random.default_rng(6)
    u = rng.standard_normal(m2.nEdges)
    o = Oracle(m2, m2.fit(), dt=600.0)
    ref = -o.op("div", u, out_where="c")
    np.testing.assert_allclose(mo.divergence(m2, u, RADIUS), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
