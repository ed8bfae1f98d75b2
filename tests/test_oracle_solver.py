"""Pins of the oracle's solver restatement (CPU): Cholesky, MGPCG, Patera
against direct solves of the same discrete systems."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import assembly as A, stokes as S
from tests.conftest import relerr


@pytest.fixture(scope="module")
def problem(c0, c0_inputs):
    i = c0_inputs
    KK, Fb, GG, MM, _ = A.assemble_opt(c0, i["Visc"], i["Dens"], i["phys"])
    return KK, Fb, GG, MM


@pytest.fixture(scope="module")
def prepared(c0, c0_inputs, problem):
    KK, Fb, GG, MM = problem
    i = c0_inputs
    RR = S.make_rotation_matrix(c0.GCOORD)
    St = S.rotate_and_lift(KK, GG, Fb, RR, i["ifix"], i["vfix"], np.zeros(c0.nVnod))
    ip, ci, w = c0.transfer(0)
    Ic = S.ic2f_dof(ip, ci, w, c0.levels[1].nnod)
    KKl, IcU = S.restrict_stiffmat(St["K1"], St["ifree"], [Ic], [3 * L.nnod for L in c0.levels])
    return St, KKl, IcU


def test_normdf():
    x = np.array([3.0, 4.0])
    assert S.normdf(x) == 5.0 / np.sqrt(2.0)


def test_rotation_matrix_orthonormal(c0):
    RR = S.make_rotation_matrix(c0.GCOORD[:50])
    R = RR.toarray()
    assert np.allclose(R.T @ R, np.eye(150), atol=1e-14)
    # third column of every node block is the radial unit vector
    X = c0.GCOORD[:50]
    rhat = X / np.linalg.norm(X, axis=1)[:, None]
    assert np.allclose(R[np.arange(150).reshape(50, 3)[:, :, None],
                         np.arange(150).reshape(50, 3)[:, None, :]][:, :, 2], rhat, atol=1e-14)


def test_bc_lift_equals_full_product(prepared):
    St, _, _ = prepared
    assert relerr(St["bc_rhs"], St["K1_full"] @ St["U0"]) < 1e-15


def test_cholesky_matches_dense_solve(prepared):
    _, KKl, _ = prepared
    Kc = KKl[-1].toarray()
    b = np.random.default_rng(0).standard_normal(len(Kc))
    x = S.Chol(Kc).solve(b)
    assert relerr(x, np.linalg.solve(Kc, b)) < 1e-12


def test_mgpcg_converges_to_direct_solution(prepared):
    St, KKl, IcU = prepared
    ch = S.Chol(KKl[-1])
    x, it, hist = S.mgpcg_solver_p(St["K1"], KKl, ch, St["Rhs"], np.zeros(len(St["ifree"])), IcU,
                                   1e-12, False, 60)
    xd = spla.spsolve(St["K1"].tocsc(), St["Rhs"])
    assert hist[-1] < 1e-12 * hist[0]
    assert relerr(x, xd) < 1e-9
    assert it < 40


def test_mgpcg_rejects_nan(prepared):
    St, KKl, IcU = prepared
    b = St["Rhs"].copy()
    b[3] = np.nan
    with pytest.raises(ValueError, match="NaN"):
        S.mgpcg_solver_p(St["K1"], KKl, S.Chol(KKl[-1]), b, np.zeros_like(b), IcU, 1e-8, True, 5)


def test_patera_matches_direct_saddle_solve(c0, c0_inputs, problem):
    """The reference algorithm iterated tightly reproduces the exact discrete
    Stokes solution (velocity block + mean-zero pressure) to 1e-8."""
    KK, Fb, GG, MM = problem
    i = c0_inputs
    U, P, info, St, _, _ = S.solve_reference(c0, KK, Fb, GG, MM, i["ifix"], i["vfix"],
                                             rtol_Pat=1e-11, rtol_U=1e-3, itmax_U=60)
    Ud, Pd = S.direct_saddle_solve(St["K1"], St["GG"], St["Fb"], St["bc_rhs"], St["ifree"], St["U0"])
    assert relerr(U, Ud) < 1e-8
    assert relerr(P - P.mean(), Pd - Pd.mean()) < 1e-8
