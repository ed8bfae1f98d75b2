"""Patera / MGPCG Stokes solve, restated (test infrastructure only).

Serial branch (nsd==1, 'GPlates' plate model, coarse_solver='chol') of
mfiles_M3TET/flow3d_patera_mg_sph_p.m:225-298 (rotation, BC lift),
:306 restrict_stiffmat (restrict_stiffmat.m:20-64), :358 Cholesky
(cholesky_factorization.m / cholesky_solve.m), :374-384 lumped pressure mass,
:413-681 Patera iteration, with mgpcg_solver_p.m:70-198 (PR-CG),
:388-596 (V-cycle) and :674-682 (smoother weights), normdf.m:22.
"""
import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp


def normdf(x):
    """normdf.m:22."""
    return np.linalg.norm(x) / np.sqrt(x.size)


def make_rotation_matrix(X):
    """make_rotation_matrix.m:69-112 -> sparse (3n x 3n)."""
    x, y, z = X[:, 0], X[:, 1], X[:, 2]
    th = np.arctan2(np.sqrt(x ** 2 + y ** 2), z)
    ph = np.arctan2(y, x)
    ph = np.where(ph < 0, ph + 2 * np.pi, ph)
    n = len(X)
    d = 3 * np.arange(n)
    rows = np.concatenate([d, d + 1, d + 2, d, d, d + 1, d + 1, d + 2, d + 2])
    cols = np.concatenate([d, d + 1, d + 2, d + 1, d + 2, d, d + 2, d, d + 1])
    vals = np.concatenate([np.cos(ph) * np.cos(th), np.cos(ph), np.cos(th), -np.sin(ph),
                           np.cos(ph) * np.sin(th), np.sin(ph) * np.cos(th),
                           np.sin(ph) * np.sin(th), -np.sin(th), np.zeros(n)])
    return sp.csc_matrix((vals, (rows, cols)), shape=(3 * n, 3 * n))


def ic2f_dof(indptr, indices, w, ncoarse):
    """Node-level interpolation -> 3-dof interpolation (intermesh_transfer_matrix.m:50-55)."""
    nf = len(indptr) - 1
    rows = np.repeat(np.arange(nf), np.diff(indptr))
    R = np.concatenate([3 * rows + c for c in range(3)])
    C = np.concatenate([3 * indices + c for c in range(3)])
    V = np.concatenate([w, w, w])
    return sp.csr_matrix((V, (R, C)), shape=(3 * nf, 3 * ncoarse))


def rotate_and_lift(KK, GG, Fb, RR, ifix, vfix, P):
    """flow3d_patera_mg_sph_p.m:233-298.  Returns dict with K1 (free x free,
    full symmetric), GG, Fb, bc_rhs, Rhs (free), ifree."""
    K1 = KK + sp.tril(KK, -1).T
    K1 = (RR.T @ K1 @ RR).tocsc()
    GG = (RR.T @ GG).tocsc()
    Fb = RR.T @ Fb
    nU = K1.shape[0]
    U0 = np.zeros(nU)
    U0[ifix] = vfix
    bc_rhs = K1 @ U0                                    # = lower + upper - diag
    ifree = np.setdiff1d(np.arange(nU), ifix)
    Rhs = Fb + GG @ P - bc_rhs
    return dict(K1=K1[ifree][:, ifree].tocsr(), K1_full=K1, GG=GG, Fb=Fb, bc_rhs=bc_rhs,
                Rhs=Rhs[ifree], ifree=ifree, U0=U0)


def restrict_stiffmat(K1, ifree, Ic2f, nUdof_mg):
    """restrict_stiffmat.m:20-49 (the blockwise product of :70-124 equals
    the plain Galerkin product).  Returns (KK list, Ic2f_U list): KK[0] is the
    diagonal of K1, KK[l] the Galerkin operator on level l's free dofs."""
    KK = [K1.diagonal().copy()]
    Ic2f_U = []
    Kl = K1
    for l in range(len(Ic2f)):
        iUfree_c = ifree[ifree < nUdof_mg[l + 1]]
        I = Ic2f[l][ifree][:, iUfree_c].tocsr()
        Ic2f_U.append(I)
        Kl = (I.T @ Kl @ I).tocsr()
        KK.append(Kl)
        ifree = iUfree_c
    return KK, Ic2f_U


class Chol:
    """cholesky_factorization.m / cholesky_solve.m: L L' = A (dense; the
    fill-reducing permutation does not change the solution)."""

    def __init__(self, A):
        A = A.toarray() if sp.issparse(A) else A
        self.c = sla.cho_factor(A, lower=True)

    def solve(self, b):
        return sla.cho_solve(self.c, b)


SMOOTH_W = dict(down1=[0.9, 0.4, 0.25], up1=[0.25, 0.4, 0.9],
                down=[0.25, 0.4, 0.9], up=[0.9, 0.4, 0.25])


def mgpcg_solver_p(K1, KK, chol, b, x, Ic2f_U, tol, is_abs_tol, itmax=50):
    """mgpcg_solver_p.m:44-207 with MG_V_cycle_v2 (:388-596), Setup 3
    weights (:674-682).  Returns (x, it, rrms_vec)."""
    nmesh = len(KK)
    if np.any(np.isnan(b)):
        raise ValueError("RHS vector contains NaNs")
    KDi = [1.0 / KK[0]] + [1.0 / KK[l].diagonal() for l in range(1, nmesh)]
    Kop = [K1] + KK[1:]
    wd = [SMOOTH_W["down1"]] + [SMOOTH_W["down"]] * (nmesh - 2)
    wu = [SMOOTH_W["up1"]] + [SMOOTH_W["up"]] * (nmesh - 2)

    def vcycle(r):
        r0 = r
        dd, rr = [], []
        for l in range(nmesh - 1):
            rr.append(r0)
            r1 = r0
            d1 = np.zeros_like(r0)
            for w in wd[l]:
                d1 = d1 + w * r1 * KDi[l]
                r1 = r0 - Kop[l] @ d1
            dd.append(d1)
            r0 = Ic2f_U[l].T @ r1
        d = chol.solve(r0)
        for l in range(nmesh - 2, -1, -1):
            d = dd[l] + Ic2f_U[l] @ d
            for w in wu[l]:
                r1 = rr[l] - Kop[l] @ d
                d = d + w * r1 * KDi[l]
        return d

    r = b - K1 @ x
    rrms = normdf(r)
    if not is_abs_tol:
        tol = tol * rrms
    rr_vec = [rrms]
    q = rd = rlast = None
    it = 0
    for it in range(1, itmax + 1):
        d = vcycle(r)
        if it == 1:
            q = d
        else:
            rdlast = rd
            beta = np.dot(r - rlast, d) / rdlast
            q = d + beta * q
        rd = np.dot(r, d)
        Kq = K1 @ q
        qKq = np.dot(q, Kq)
        rlast = r
        alpha = rd / qKq
        x = x + alpha * q
        r = r - alpha * Kq
        rrms = normdf(r)
        rr_vec.append(rrms)
        if rrms < tol:
            break
    return x, it, np.array(rr_vec)


def patera(K1, KK, chol, Ic2f_U, GG, MM_tril, Fb, bc_rhs, ifree, U0, P, Fpdil=0.0,
           rtol_Pat=1e-4, rtol_U=5e-3, itmax_U=20, itmin_Pat=3, itmax_Pat=120,
           wght_MM=(1.0,), info=None):
    """flow3d_patera_mg_sph_p.m:374-620 + final solve :629-632.
    Returns (U (rotated frame, full), P, info)."""
    rtol_restart = 0.8 * rtol_Pat
    MMf = MM_tril + sp.tril(MM_tril, -1).T
    LMM = np.asarray(MMf.sum(axis=1)).ravel()
    LiMM = 1.0 / LMM
    U = U0.copy()
    nU = len(U)

    def precondition_P(r):
        d = r * LiMM
        for w in wght_MM:
            d = d + w * LiMM * (r - MMf @ d)
        return d

    def rm0(a):
        return a - np.mean(a)

    Rhs = Fb + GG @ P - bc_rhs
    U[ifree], nit, _ = mgpcg_solver_p(K1, KK, chol, Rhs[ifree], U[ifree], Ic2f_U,
                                      rtol_U * rtol_restart, False, itmax_U)
    nitU = nit
    r = rm0(-GG.T @ U + Fpdil)
    norm_rP = normdf(r)
    tol_Pat = rtol_Pat * norm_rP
    U[ifree], nit, _ = mgpcg_solver_p(K1, KK, chol, Rhs[ifree], U[ifree], Ic2f_U,
                                      rtol_U * rtol_restart * norm_rP, True, itmax_U)
    nitU += nit
    tol_U_abs = rtol_U * tol_Pat
    nitZ = 0
    restart = True
    conv_problem = False
    q = rd = rlast = None
    tolZ = None
    itPat = 0
    hist = []
    for itPat in range(1, itmax_Pat + 1):
        d = precondition_P(r)
        if restart:
            q = d
            restart = False
            norm_rP0 = norm_rP
            tolZ = rtol_U * max(rtol_restart * norm_rP0, tol_Pat)
        else:
            rdlast = rd
            beta = np.dot(r - rlast, d) / rdlast
            q = d + beta * q
        q = rm0(q)
        rd = np.dot(r, d)
        y = GG @ q
        z = np.zeros(nU)
        z[ifree], nit, _ = mgpcg_solver_p(K1, KK, chol, y[ifree], z[ifree], Ic2f_U, tolZ, True, itmax_U)
        nitZ += nit
        Sq = GG.T @ z
        qSq = np.dot(q, Sq)
        rlast = r
        alpha = rd / qSq
        P = P + alpha * q
        U = U + alpha * z
        r = rm0(r - alpha * Sq)
        norm_rP_last = norm_rP
        norm_rP = normdf(r)
        norm_rP_min = min(norm_rP_last, norm_rP)
        hist.append(norm_rP)
        if norm_rP < tol_Pat and itPat >= itmin_Pat:
            break
        if norm_rP < rtol_restart * norm_rP0 or norm_rP > 10 * norm_rP0:
            Rhs = Fb + GG @ P - bc_rhs
            U[ifree], nit, _ = mgpcg_solver_p(K1, KK, chol, Rhs[ifree], U[ifree], Ic2f_U,
                                              tol_U_abs, True, itmax_U)
            nitU += nit
            r = -GG.T @ U + Fpdil
            norm_rP_true = normdf(r)
            if norm_rP_true > 1.1 * norm_rP:
                rtol_U = rtol_U * 0.75 * (norm_rP / norm_rP_true)
            if norm_rP > 10 * norm_rP0 or norm_rP > 1.2 * norm_rP_min:
                if conv_problem:
                    break
                rtol_U = 0.1 * rtol_U
                conv_problem = True
            restart = True
    Rhs = Fb + GG @ P - bc_rhs
    U[ifree], nit, _ = mgpcg_solver_p(K1, KK, chol, Rhs[ifree], U[ifree], Ic2f_U,
                                      tol_U_abs, True, itmax_U)
    nitU += nit
    if np.any(np.isnan(U)) or np.any(np.isnan(P)):
        raise ValueError("Solution contains NaNs")
    return U, P, dict(nitP=itPat, nitU=nitU, nitZ=nitZ, norm_rP=norm_rP, hist=hist)


def direct_saddle_solve(K1_free, GG, Fb, bc_rhs, ifree, U0):
    """Exact discrete solution (velocity block + mean-zero pressure) by a
    sparse direct solve of [K -G; -G' 0] with the pressure mean fixed by a
    bordered Lagrange multiplier; defines the reference solution the
    iterative solve is compared with."""
    import scipy.sparse.linalg as spla
    Gf = GG[ifree]
    nf = len(ifree)
    nP = GG.shape[1]
    f = (Fb - bc_rhs)[ifree]
    one = sp.csr_matrix(np.ones((1, nP)))
    A = sp.bmat([[K1_free, -Gf, None], [-Gf.T, None, one.T], [None, one, None]], format="csc")
    rhs = np.concatenate([f, GG.T @ U0, [0.0]])
    x = spla.spsolve(A, rhs)
    U = U0.copy()
    U[ifree] = x[:nf]
    return U, x[nf:nf + nP]


def solve_reference(mesh, KK, Fb, GG, MM, ifix, vfix, P0=None, **kw):
    """End-to-end oracle Stokes solve (flow3d_patera_mg_sph_p serial branch)."""
    RR = make_rotation_matrix(mesh.GCOORD)
    P = np.zeros(mesh.nVnod) if P0 is None else P0.copy()
    S = rotate_and_lift(KK, GG, Fb, RR, ifix, vfix, P)
    Ic2f = []
    for l in range(mesh.nmg - 1):
        ip, ci, w = mesh.transfer(l)
        Ic2f.append(ic2f_dof(ip, ci, w, mesh.levels[l + 1].nnod))
    nUdof_mg = [3 * L.nnod for L in mesh.levels]
    KKl, Ic2f_U = restrict_stiffmat(S["K1"], S["ifree"], Ic2f, nUdof_mg)
    chol = Chol(KKl[-1])
    U, P, info = patera(S["K1"], KKl, chol, Ic2f_U, S["GG"], MM, S["Fb"], S["bc_rhs"],
                        S["ifree"], S["U0"], P, **kw)
    return U, P, info, S, KKl, Ic2f_U
