"""Pins of the diffusion-step oracle (oracle/diffusion.py) on the CPU.

* every integration rule integrates the monomials of its degree exactly;
* the element matrices of one element of each class frame equal the closed
  form of the spherical metric: for N linear in (theta, phi, r),
  grad N_i . grad N_j = a_i a_j / r^2 + b_i b_j / (r sin theta)^2 + c_i c_j
  with (a, b, c) = d(theta, phi, r) N, and dV = r^2 sin theta |J_SL| dxi;
* assembled conductivity rows sum to zero, the matrix is symmetric positive
  semi-definite, the lumped and consistent masses have the same total;
* a diffusion step equals a direct sparse solve of the same system, and a
  constant field with matching fixed values stays constant.
"""
import math

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import diffusion as D
from tests.conftest import relerr

DEGREE = {1: 1, 4: 2, 5: 3, 8: 3, 10: 3, 11: 4, 14: 5}


@pytest.mark.parametrize("nip", sorted(DEGREE))
def test_ip_rules_are_exact_to_their_degree(nip):
    x, w = D.ip_rule(nip)
    deg = DEGREE[nip]
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            for c in range(deg + 1 - a - b):
                exact = math.factorial(a) * math.factorial(b) * math.factorial(c) / math.factorial(a + b + c + 3)
                got = np.sum(w * x[0] ** a * x[1] ** b * x[2] ** c)
                assert abs(got - exact) < 1e-14, (nip, a, b, c, got, exact)


def _sph2cart(th, ph, r):
    return np.stack([r * np.sin(th) * np.cos(ph), r * np.sin(th) * np.sin(ph), r * np.cos(th)], -1)


def _cart2sph(X):
    r = np.linalg.norm(X, axis=-1)
    th = np.arctan2(np.hypot(X[..., 0], X[..., 1]), X[..., 2])
    ph = np.arctan2(X[..., 1], X[..., 0])
    return np.stack([th, np.where(ph < 0, ph + 2 * np.pi, ph), r], -1)


def _closed_form(P4, h, kappa, nip):
    """Packed CC, MM of the element with vertices P4 + h_k e_k (k = 1..3), P4
    in its own spherical frame."""
    x, w = D.ip_rule(nip)
    grad = np.array([[1 / h[0], 0, 0], [0, 1 / h[1], 0], [0, 0, 1 / h[2]], [0, 0, 0]])
    grad[3] = -grad[:3].sum(0)
    C = np.zeros((4, 4))
    M = np.zeros((4, 4))
    for q in range(x.shape[1]):
        xi = x[:, q]
        th, ph, r = P4 + h * xi
        N = np.array([xi[0], xi[1], xi[2], 1 - xi.sum()])
        wq = w[q] * r ** 2 * np.sin(th) * np.prod(h)
        G = grad[:, 0:1] * grad[:, 0:1].T / r ** 2 + grad[:, 1:2] * grad[:, 1:2].T / (r * np.sin(th)) ** 2 \
            + grad[:, 2:3] * grad[:, 2:3].T
        C += kappa * wq * G
        M += wq * np.outer(N, N)
    iu = [(i, j) for i in range(4) for j in range(i, 4)]
    return np.array([C[i, j] for i, j in iu]), np.array([M[i, j] for i, j in iu])


def _element(frame_pts, rotated):
    """(X, SPH) of the 4 vertices whose frame coordinates are frame_pts; in
    the rotated frame x' = x, y' = -z, z' = y when ``rotated``."""
    Xf = _sph2cart(frame_pts[:, 0], frame_pts[:, 1], frame_pts[:, 2])
    X = np.stack([Xf[:, 0], Xf[:, 2], -Xf[:, 1]], 1) if rotated else Xf
    return X, _cart2sph(X)


@pytest.mark.parametrize("case", ["out_cone", "cross_2pi", "in_cone", "in_cone_iso"])
@pytest.mark.parametrize("nip", [4, 14])
def test_element_matrices_closed_form(case, nip):
    h = np.array([0.05, 0.07, 0.03])
    kappa = 0.37
    if case == "out_cone":
        P4, rotated, want = np.array([1.0, 0.7, 0.8]), False, 0
    elif case == "cross_2pi":
        P4, rotated, want = np.array([1.2, 2 * np.pi - 0.03, 0.8]), False, 1
    elif case == "in_cone":
        P4, rotated, want = np.array([np.pi / 2, 4.71, 0.8]), True, 2
    else:
        P4, rotated, want = None, True, 3
        for ph0 in np.linspace(5.40, 5.60, 81):   # straddle theta = 45 degrees
            cand = np.array([np.pi / 2, ph0, 0.8])
            pts = np.stack([cand + np.diag(h)[k] for k in range(3)] + [cand])
            X, S = _element(pts, True)
            if D.classes_linear(S, np.arange(4)[None, :])[0] == 3:
                P4 = cand
                break
        assert P4 is not None
    pts = np.stack([P4 + np.diag(h)[k] for k in range(3)] + [P4])
    X, S = _element(pts, rotated)
    EL = np.arange(4)[None, :]
    assert D.classes_linear(S, EL)[0] == want
    Cv, Mv = D.element_matrices(X, S, EL, kappa, nip)
    Cr, Mr = _closed_form(P4, h, kappa, nip)
    assert relerr(Cv[0], Cr) < 1e-12
    assert relerr(Mv[0], Mr) < 1e-12


def test_closed_form_catches_a_transposed_jacobian():
    """The closed-form check is sharp: swapping theta and phi scalings (as a
    transposed J_SL would) changes the conductivity far beyond tolerance."""
    P4, h = np.array([1.0, 0.7, 0.8]), np.array([0.05, 0.07, 0.03])
    Cr, _ = _closed_form(P4, h, 1.0, 4)
    Cw, _ = _closed_form(P4, h[[1, 0, 2]], 1.0, 4)
    assert relerr(Cw, Cr) > 1e-2


@pytest.fixture(scope="module")
def c0v(c0):
    EL4 = c0.levels[0].EL2NOD[:, :4]
    return c0, EL4, c0.nVnod


def test_classes_match_the_mesh_classes(c0v):
    m, EL4, _ = c0v
    cls = D.classes_linear(m.GCOORD_SPH, EL4, m.theta_cone)
    for k, name in enumerate(("out_cone_no_cross_2pi", "out_cone_cross_2pi", "in_cone_no_iso", "in_cone_iso")):
        assert np.array_equal(np.nonzero(cls == k)[0], np.sort(m.classes[name]))


def test_assembled_operators(c0v):
    m, EL4, n = c0v
    Cv, Mv = D.element_matrices(m.GCOORD, m.GCOORD_SPH, EL4, 1.0, 4)
    CC = D.assemble(EL4, Cv, n)
    MM = D.assemble(EL4, Mv, n)
    A = CC.toarray()
    assert np.abs(A.sum(1)).max() < 1e-13 * np.abs(A).max()
    assert np.abs(A - A.T).max() < 1e-14 * np.abs(A).max()
    assert np.linalg.eigvalsh(A).min() > -1e-12 * np.abs(A).max()
    assert abs(MM.sum() - np.asarray(MM.sum(axis=1)).sum()) < 1e-12
    # total mass is the shell volume up to the mesh's geometric error
    vol = 4 / 3 * np.pi * (m.r_surf ** 3 - m.r_cmb ** 3)
    assert abs(MM.sum() / vol - 1) < 0.02


def _fixed(m):
    ifix = np.nonzero(np.isin(m.PointID[:m.nVnod], (301, 306)))[0]
    vfix = np.where(m.PointID[ifix] == 301, 1.0, 0.0)
    return ifix, vfix


@pytest.mark.parametrize("theta,lumping", [(1.0, True), (0.5, True), (1.0, False)])
def test_step_equals_direct_solve(c0v, theta, lumping):
    m, EL4, n = c0v
    rng = np.random.default_rng(3)
    T0 = rng.random(n)
    ifix, vfix = _fixed(m)
    T, info = D.diffusion3d_p(T0, 1e-2, m.GCOORD, m.GCOORD_SPH, EL4, ifix, vfix, 0.5, theta=theta,
                              lumping=lumping, tol=1e-14)
    ifree = info["ifree"]
    A, rhs = info["A"], info["rhs"]
    x = spla.spsolve(A[ifree][:, ifree].tocsc(), rhs[ifree])
    assert relerr(T[ifree], x) < 1e-10
    assert np.all(T[ifix] == vfix)
    assert info["hist"][-1] < info["tol"]


def test_constant_field_stays_constant(c0v):
    m, EL4, n = c0v
    ifix, _ = _fixed(m)
    T, info = D.diffusion3d_p(np.full(n, 0.3), 5e-2, m.GCOORD, m.GCOORD_SPH, EL4, ifix,
                              np.full(len(ifix), 0.3), 1.0)
    assert np.abs(T - 0.3).max() < 1e-12


def test_kappa_zero_returns_input(c0v):
    m, EL4, n = c0v
    T0 = np.linspace(0, 1, n)
    T, info = D.diffusion3d_p(T0, 0.0, m.GCOORD, m.GCOORD_SPH, EL4, [], [], 1.0)
    assert np.array_equal(T, T0) and info["nit"] == 0
