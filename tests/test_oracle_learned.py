"""Pins of the CPU restatements used by the learned-warm-start tests:
the graph network (oracle/gnn.py) and the semi-Lagrangian update
(oracle/advect.py), plus the host-side vertex graph."""
import numpy as np
import pytest
import torch

from conftest import relerr
from oracle import advect as A
from oracle import gnn as G


@pytest.fixture(scope="module")
def graph(tiny):
    from m3tet_sph_amd import learned as L
    return L.vertex_graph(tiny)


def test_bf16_round_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-30, 30, 20000),
                        np.array([0.0, -0.0, 1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 3e38,
                                  np.inf, -np.inf], dtype=np.float32)]).astype(np.float32)
    ours = G.bf16_round(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(ours.view(np.uint32), ref.view(np.uint32))
    assert G.bf16_round(np.array([1.0 + 2 ** -8], dtype=np.float32))[0] == 1.0  # tie to even
    assert np.isnan(G.bf16_round(np.array([np.nan], dtype=np.float32))[0])


def test_vertex_graph_is_symmetric_and_edge_parents_are_edges(tiny, graph):
    ptr, col, par, rad = graph
    nV = tiny.nVnod
    rows = np.repeat(np.arange(nV), np.diff(ptr))
    fwd = set(zip(rows.tolist(), col.tolist()))
    assert all((j, i) in fwd for i, j in fwd)
    assert all(np.all(np.diff(col[ptr[i]:ptr[i + 1]]) > 0) for i in range(nV))
    assert not np.any(rows == col)
    # each edge node lies near the midpoint of its parents (curved edges bow out slightly)
    X = tiny.GCOORD
    mid = 0.5 * (X[par[:, 0]] + X[par[:, 1]])
    elen = np.linalg.norm(X[par[:, 0]] - X[par[:, 1]], axis=1)
    assert np.all(np.linalg.norm(X[nV:] - mid, axis=1) < 0.1 * elen)
    assert np.allclose(rad, np.linalg.norm(X[:nV], axis=1))


def test_mean_aggregation_known_answer():
    # path graph 0-1-2-3
    ptr = np.array([0, 1, 3, 5, 6])
    col = np.array([1, 0, 2, 1, 3, 2])
    H = np.array([[1.0], [2.0], [4.0], [8.0]])
    s = G._neighbour_sum(ptr, col, H, np.float64)
    assert np.allclose(s[:, 0], [2.0, 5.0, 10.0, 4.0])


def _problem(graph, seed=3):
    ptr, col, _, _ = graph
    n = len(ptr) - 1
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, G.NF))
    Yt = rng.standard_normal((n, G.NO))
    p = rng.standard_normal(G.NPARAM) * 0.2
    norm = np.concatenate([rng.standard_normal(G.NF) * 0.1, 1 + rng.random(G.NF), 0.5 + rng.random(G.NO)])
    return X, Yt, p, norm


def test_gnn_gradient_matches_finite_differences(graph):
    ptr, col, _, _ = graph
    X, Yt, p, norm = _problem(graph)
    loss, g, _ = G.loss_and_grad(X, Yt, p, norm, ptr, col, exact=False)
    rng = np.random.default_rng(5)
    idx = np.concatenate([rng.choice(G.NPARAM, 60, replace=False),
                          np.arange(G.OFF_BD, G.NPARAM), np.arange(G.OFF_BE, G.OFF_BE + 4)])
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (G.loss_and_grad(X, Yt, pp, norm, ptr, col, exact=False)[0] -
              G.loss_and_grad(X, Yt, pm, norm, ptr, col, exact=False)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-7, (i, fd, g[i])


def test_gnn_planted_backward_error_is_caught(graph, monkeypatch):
    """A transposed aggregation in the backward pass (sum over i's
    neighbours without the 1/deg of the sender) fails the check."""
    ptr, col, _, _ = graph
    X, Yt, p, norm = _problem(graph)
    orig = G._neighbour_sum
    calls = {"n": 0}

    def wrong(ptr_, col_, V, dtype):
        calls["n"] += 1
        out = orig(ptr_, col_, V, dtype)
        # the first 3 calls are the forward aggregations; later ones are backward
        if calls["n"] > 3:
            out = out * (np.diff(ptr_)[:, None] / np.diff(ptr_).mean())
        return out

    monkeypatch.setattr(G, "_neighbour_sum", wrong)
    _, g, _ = G.loss_and_grad(X, Yt, p, norm, ptr, col, exact=False)
    monkeypatch.setattr(G, "_neighbour_sum", orig)
    _, g_ok, _ = G.loss_and_grad(X, Yt, p, norm, ptr, col, exact=False)
    assert relerr(g[G.OFF_WE:G.OFF_L], g_ok[G.OFF_WE:G.OFF_L]) > 1e-3


def test_gnn_device_arithmetic_mode_is_the_same_function(graph):
    ptr, col, _, _ = graph
    X, Yt, p, norm = _problem(graph)
    l64, g64, y64 = G.loss_and_grad(X, Yt, p, norm, ptr, col, exact=False)
    l32, g32, y32 = G.loss_and_grad(X, Yt, p, norm, ptr, col, exact=True)
    assert relerr(y32, y64) < 2e-2  # BF16 activations
    assert abs(l32 - l64) < 2e-2 * l64
    assert relerr(g32, g64) < 5e-2


def _shell_p1(tiny):
    nV = tiny.nVnod
    X = tiny.GCOORD[:nV]
    tets = tiny.levels[0].EL2NOD[:, :4].astype(np.int64)
    return X, tets


def _interior(X):
    r = np.linalg.norm(X, axis=1)
    return (r > r.min() + 1e-6) & (r < r.max() - 1e-6)


def test_extrapolation_is_exact_for_polynomials_of_its_degree():
    """The base state of degree d through d+1 past states reproduces any
    polynomial in time of degree <= d at a new time (uneven steps), order 0
    returns the last state, and a degree-(d+1) polynomial is missed by the
    leading term of the interpolation error, prod (t - t_j) times the
    coefficient."""
    rng = np.random.default_rng(3)
    ts = np.cumsum(rng.uniform(0.5, 1.5, 5))
    t = ts[-1] + 0.8
    c = rng.standard_normal((5, 7))
    for order in range(4):
        poly = lambda tt, d=order: sum(c[i] * tt ** i for i in range(d + 1))
        states = [poly(tj) for tj in ts]
        assert np.allclose(G.extrapolate(ts, states, t, order), poly(t), rtol=1e-11, atol=1e-11)
        # one degree more: error = c_{d+1} prod over the used points (t - t_j)
        poly2 = lambda tt, d=order: sum(c[i] * tt ** i for i in range(d + 2))
        err = poly2(t) - G.extrapolate(ts, [poly2(tj) for tj in ts], t, order)
        assert np.allclose(err, c[order + 1] * np.prod(t - ts[-order - 1:]), rtol=1e-9, atol=1e-9)
    assert np.array_equal(G.extrapolate(ts, [c[i] for i in range(5)], t, 0), c[4])
    # fewer states than order + 1: the degree drops to what the history holds
    assert np.allclose(G.extrapolate(ts[:2], [c[0], c[1]], t, 3),
                       c[1] + (c[1] - c[0]) * (t - ts[1]) / (ts[1] - ts[0]), rtol=1e-13)


def test_sl_zero_velocity_is_identity_with_boundary_values(tiny):
    X, tets = _shell_p1(tiny)
    T = np.random.default_rng(0).random(len(X))
    out = A.sl_advect(X, tets, T, np.zeros_like(X), 0.1, T_cmb=1.0, T_surf=0.0)
    inner = _interior(X)
    r = np.linalg.norm(X, axis=1)
    assert np.allclose(out[inner], T[inner], atol=1e-13)
    assert np.all(out[np.isclose(r, r.min())] == 1.0) and np.all(out[np.isclose(r, r.max())] == 0.0)


def _rotation(omega):
    return lambda P: np.cross(omega[None, :], P)


def test_sl_reproduces_linear_fields_at_the_departure_points(tiny):
    """With a velocity linear in x (rigid rotation) and T linear in x, P1
    interpolation is exact, so the update equals T at the scheme's departure
    points."""
    X, tets = _shell_p1(tiny)
    omega = np.array([0.3, -0.2, 0.9])
    uc = _rotation(omega)(X)
    a, b = np.array([0.7, -1.1, 0.4]), 0.25
    T = X @ a + b
    elen = np.linalg.norm(X[tets[:, 0]] - X[tets[:, 1]], axis=1).min()
    dt = 0.2 * elen / np.linalg.norm(uc, axis=1).max()
    out = A.sl_advect(X, tets, T, uc, dt)
    xd = A.departure_points(X, tets, _rotation(omega), dt)
    inner = _interior(X)
    assert np.max(np.abs(out[inner] - (xd[inner] @ a + b))) < 1e-12


def test_sl_backtracking_is_second_order(tiny):
    X, tets = _shell_p1(tiny)
    omega = np.array([0.3, -0.2, 0.9])
    th = np.linalg.norm(omega)
    k = omega / th

    def exact(dt):  # rotation by -dt*|omega| about k (Rodrigues)
        c, s = np.cos(-th * dt), np.sin(-th * dt)
        return X * c + np.cross(k, X) * s + np.outer(X @ k, k) * (1 - c)

    errs = []
    for dt in (0.04, 0.02, 0.01):
        xd = A.departure_points(X, tets, _rotation(omega), dt)
        errs.append(np.abs(xd - exact(dt)).max())
    assert errs[0] / errs[1] > 6 and errs[1] / errs[2] > 6, errs
