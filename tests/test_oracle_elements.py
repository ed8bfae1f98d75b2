"""Pins of the oracle's element numerics (known answers, closed forms)."""
import math

import numpy as np
import pytest

from oracle import elements as E


def test_elblk_known_answer():
    # elblk_AtxBt.m:31-43,57-60 and elblk_AxB.m:30-42,56-59: A=[1 2 3;4 5 6;7 8 9],
    # B=100+A replicated over 4 elements; the blocked product minus the
    # plain product is exactly zero.
    A = np.array([[1.0, 2, 3], [4, 5, 6], [7, 8, 9]])
    B = 100 + A
    nn = 4
    cols = lambda M, j: np.tile(M[:, j], (nn, 1))
    Cx, Cy, Cz = E.elblk_AtxBt(cols(A, 0), cols(A, 1), cols(A, 2), cols(B, 0), cols(B, 1), cols(B, 2))
    C = A.T @ B.T
    for e in (0, nn - 1):
        got = np.stack([Cx[e], Cy[e], Cz[e]], axis=1)
        assert np.array_equal(got, C)
    Dx, Dy, Dz = E.elblk_AxB(cols(A, 0), cols(A, 1), cols(A, 2), cols(B, 0), cols(B, 1), cols(B, 2))
    assert np.array_equal(np.stack([Dx[0], Dy[0], Dz[0]], axis=1), A @ B)


def test_elblk_det_inv_transpose():
    rng = np.random.default_rng(0)
    M = rng.standard_normal((50, 3, 3))
    Jx, Jy, Jz = M[:, :, 0], M[:, :, 1], M[:, :, 2]     # columns
    d = E.elblk_detA(Jx, Jy, Jz)
    assert np.allclose(d, np.linalg.det(M), rtol=1e-12, atol=1e-12)
    iJx, iJy, iJz = E.elblk_invert_A(Jx, Jy, Jz, d)
    inv = np.linalg.inv(M)
    # rows of inv(A) when A has columns Jx, Jy, Jz
    assert np.allclose(np.stack([iJx, iJy, iJz], axis=1), inv, rtol=1e-10, atol=1e-12)
    Tx, Ty, Tz = E.elblk_transpose_A(Jx, Jy, Jz)
    assert np.array_equal(np.stack([Tx, Ty, Tz], axis=2), np.transpose(M, (0, 2, 1)))


def test_quadrature_constants_and_exactness():
    ipx, ipw = E.ip_tetrahedron(14)
    # decimal constants quoted in ip_tetrahedron.m:109-117
    a, b, c, d = ipx[0, 0], ipx[0, 1], ipx[0, 4], ipx[0, 5]
    e, f = ipx[0, 8], ipx[0, 11]
    assert abs(b - 0.3108859192633005) < 1e-15 and abs(a - 0.0673422422100983) < 1e-15
    assert abs(d - 0.0927352503108912) < 1e-15 and abs(c - 0.7217942490673264) < 1e-15
    assert abs(e - 0.4544962958743506) < 1e-15 and abs(f - 0.0455037041256494) < 1e-15
    assert abs(6 * ipw[0] - 0.1126879257180162) < 1e-15
    assert abs(6 * ipw[4] - 0.0734930431163619) < 1e-15
    assert abs(6 * ipw[8] - 0.0425460207770812) < 1e-15
    assert abs(ipw.sum() - 1 / 6) < 1e-16
    # exact for all monomials r^i s^j t^k of degree <= 5 (degree of precision 5)
    for i in range(6):
        for j in range(6 - i):
            for k in range(6 - i - j):
                exact = math.factorial(i) * math.factorial(j) * math.factorial(k) / math.factorial(i + j + k + 3)
                q = np.sum(ipw * ipx[0] ** i * ipx[1] ** j * ipx[2] ** k)
                assert abs(q - exact) < 1e-15, (i, j, k)
    # degree 6 is not integrated exactly (the rule is not over-fitted)
    q6 = np.sum(ipw * ipx[0] ** 6)
    assert abs(q6 - math.factorial(6) / math.factorial(9)) > 1e-9


def _nodes10():
    return np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, 0], [.5, .5, 0], [0, .5, .5],
                     [0, 0, .5], [.5, 0, 0], [.5, 0, .5], [0, .5, 0]])


def _nodes20():
    r = [1, 0, 0, 0, 2 / 3, 1 / 3, 0, 0, 0, 0, 1 / 3, 2 / 3, 2 / 3, 1 / 3, 0, 0, 0, 1 / 3, 1 / 3, 1 / 3]
    s = [0, 1, 0, 0, 1 / 3, 2 / 3, 2 / 3, 1 / 3, 0, 0, 0, 0, 0, 0, 1 / 3, 2 / 3, 1 / 3, 0, 1 / 3, 1 / 3]
    t = [0, 0, 1, 0, 0, 0, 1 / 3, 2 / 3, 2 / 3, 1 / 3, 0, 0, 1 / 3, 2 / 3, 0, 0, 1 / 3, 1 / 3, 0, 1 / 3]
    return np.stack([r, s, t], axis=1)


@pytest.mark.parametrize("fn,nodes", [(E.sf_tet10, _nodes10), (E.sf_tet20, _nodes20)])
def test_shape_functions_kronecker_unity_and_derivatives(fn, nodes):
    X = nodes()
    n = len(X)
    # Kronecker delta at the nodes (sf_dsf_tet.m:829-840 node table)
    Nm = np.array([fn(*x)[0] for x in X])
    assert np.allclose(Nm, np.eye(n), atol=1e-13)
    rng = np.random.default_rng(1)
    for _ in range(20):
        p = rng.dirichlet(np.ones(4))[:3]
        N, dN = fn(*p)
        assert abs(N.sum() - 1) < 1e-13          # partition of unity
        assert np.allclose(dN.sum(axis=0), 0, atol=1e-12)
        h = 1e-6                                  # derivatives vs central differences
        for k in range(3):
            dp = np.zeros(3)
            dp[k] = h
            fd = (fn(*(p + dp))[0] - fn(*(p - dp))[0]) / (2 * h)
            assert np.allclose(dN[:, k], fd, atol=1e-7)
