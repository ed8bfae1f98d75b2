"""Thermal diffusion step of the reference, restated (test infrastructure only).

* ``ip_rule``: ip_tetrahedron.m (rules 1, 4, 5, 8, 10, 11 at :23-97, 14 at
  :100-122), weights / 6 (:189-191).
* ``classes_linear``: check_els_in_cone.m (cone count on the 4 vertices) and
  check_phi.m (a vertex pair more than pi apart in phi) on linear tetrahedra,
  as assembly_diffusion_3d_double_jacobian.m:73-93 recomputes them.
* ``element_matrices``: the MILAMIN class kernels
  mfiles_SPH/assembly_diffusion_els_out_cone_no_cross_2pi_opt.m:54-149
  (vertex (theta, phi, r) from GCOORD_SPH), ..._out_cone_cross_2pi_opt.m:
  59-61 (phi + pi, wrapped), ..._in_cone_no_iso_opt.m:61-160 and
  ..._in_cone_iso_opt.m:56-160 (GCOORD rotated 90 degrees about X, then
  spherical; for 4-node elements both in-cone kernels reduce to the same
  arithmetic): J_SL = (theta, phi, r)' dN4/dxi, J_SC at each integration
  point, dN/dx = (J_SC inv(J_SL))' dN/dxi, CC += kappa w dN/dx . dN/dx,
  MM += w N N', w = detJ_CS detJ_SL w_ip.  Packed per element as the upper
  triangle row by row (the i <= j loop of the kernels).
* ``assemble``: sparse_create of both halves (assembly_diffusion_3d_double_
  jacobian.m:195-199).
* ``diffusion3d_p``: mfiles_M3TET/diffusion3d_p.m:88-116 (lumping, theta
  rule, fixed temperatures) with ``pcg_solver_p`` (pcg_solver_p.m:49-186,
  Jacobi 'diag' preconditioner, Polak-Ribiere beta, RMS norms).
"""
import numpy as np
import scipy.sparse as sp

from . import elements as E


def ip_rule(nip):
    """(ipx (3, nip), ipw (nip,)) of ip_tetrahedron(nip), weights / 6."""
    if nip == 14:
        return E.ip_tetrahedron(14)
    if nip == 1:
        x = np.full((3, 1), 0.25)
        w = np.array([1.0])
    elif nip == 4:
        a, b = (5 + 3 * np.sqrt(5)) / 20, (5 - np.sqrt(5)) / 20
        x = np.array([[a, b, b, b], [b, a, b, b], [b, b, a, b]])
        w = np.full(4, 0.25)
    elif nip == 5:
        x = np.array([[1 / 4, 1 / 2, 1 / 6, 1 / 6, 1 / 6],
                      [1 / 4, 1 / 6, 1 / 2, 1 / 6, 1 / 6],
                      [1 / 4, 1 / 6, 1 / 6, 1 / 2, 1 / 6]])
        w = np.array([-4 / 5, 9 / 20, 9 / 20, 9 / 20, 9 / 20])
    elif nip == 8:
        x = np.array([[0, 1 / 3, 1 / 3, 1 / 3, 1, 0, 0, 0],
                      [1 / 3, 0, 1 / 3, 1 / 3, 0, 1, 0, 0],
                      [1 / 3, 1 / 3, 0, 1 / 3, 0, 0, 1, 0]])
        w = np.array([9 / 40] * 4 + [1 / 40] * 4)
    elif nip == 10:
        a, b = 0.7784952948213300, 0.0738349017262234
        c, d = 0.4062443438840510, 0.0937556561159491
        w1, w2 = 0.0476331348432089, 0.1349112434378610
        x = np.array([[a, b, b, b, c, c, c, d, d, d],
                      [b, a, b, b, c, d, d, c, c, d],
                      [b, b, a, b, d, c, d, c, d, c]])
        w = np.array([w1] * 4 + [w2] * 6)
    elif nip == 11:
        a, b = (1 + np.sqrt(5 / 14)) / 4, (1 - np.sqrt(5 / 14)) / 4
        w1, w2, w3 = -148 / 1875, 343 / 7500, 56 / 375
        x = np.array([[1 / 4, 11 / 14, 1 / 14, 1 / 14, 1 / 14, a, a, a, b, b, b],
                      [1 / 4, 1 / 14, 11 / 14, 1 / 14, 1 / 14, a, b, b, a, a, b],
                      [1 / 4, 1 / 14, 1 / 14, 11 / 14, 1 / 14, b, a, b, a, b, a]])
        w = np.array([w1] + [w2] * 4 + [w3] * 6)
    else:
        raise ValueError("unknown integration rule")
    return x, w / 6.0


def classes_linear(SPH, EL4, theta_cone=45.0):
    """Class per linear element: 0 out-cone, 1 out-cone crossing phi = 2pi,
    2 in-cone (3 or 4 vertices in the cone), 3 in-cone iso (1 or 2)."""
    tc = theta_cone * np.pi / 180.0
    th = SPH[EL4, 0]
    cnt = ((th <= tc) | (th >= np.pi - tc)).sum(axis=1)
    ph = SPH[EL4, 1]
    cross = np.zeros(len(EL4), dtype=bool)
    for a, b in ((0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)):
        cross |= np.abs(ph[:, a] - ph[:, b]) > np.pi
    cls = np.zeros(len(EL4), dtype=np.int32)
    cls[(cnt == 0) & cross] = 1
    cls[cnt >= 3] = 2
    cls[(cnt > 0) & (cnt < 3)] = 3
    return cls


def _vertex_coords(cls, X, SPH, EL4):
    """(theta, phi, r) of the 4 vertices in the frame of the element's class."""
    if cls == 0:
        S = SPH[EL4]
        return S[..., 0], S[..., 1], S[..., 2]
    if cls == 1:
        S = SPH[EL4]
        ph = S[..., 1] + np.pi
        ph = np.where(ph > 2 * np.pi, ph - 2 * np.pi, ph)
        return S[..., 0], ph, S[..., 2]
    Xe = X[EL4]
    return E.elblk_cartesian2spherical(Xe[..., 0], -Xe[..., 2], Xe[..., 1])


def element_matrices(X, SPH, EL4, kappa, nip=4, theta_cone=45.0, cls=None):
    """Packed upper-triangle (row by row, 10 values) CC and MM per element."""
    if cls is None:
        cls = classes_linear(SPH, EL4, theta_cone)
    ipx, ipw = ip_rule(nip)
    dN4ds = E.sf_tet4(0.25, 0.25, 0.25)[1]          # (4, 3), constant
    nel = len(EL4)
    CC = np.zeros((nel, 10))
    MM = np.zeros((nel, 10))
    for k in range(4):
        els = np.nonzero(cls == k)[0]
        if len(els) == 0:
            continue
        n = len(els)
        th, ph, r = _vertex_coords(k, X, SPH, EL4[els])
        Jth, Jph, Jr = th @ dN4ds, ph @ dN4ds, r @ dN4ds
        detJ_SL = E.elblk_detA(Jth, Jph, Jr)
        if np.any(detJ_SL < 0):
            raise ValueError("negative Jacobi")
        iJth, iJph, iJr = E.elblk_invert_A(Jth, Jph, Jr, detJ_SL)
        C = np.zeros((n, 10))
        M = np.zeros((n, 10))
        for ip in range(ipx.shape[1]):
            N4 = E.sf_tet4(*ipx[:, ip])[0]
            th_ip, ph_ip, r_ip = th @ N4, ph @ N4, r @ N4
            if np.any(th_ip == 0):
                raise ValueError("ip has theta = 0")
            s_th, c_th, s_ph, c_ph = np.sin(th_ip), np.cos(th_ip), np.sin(ph_ip), np.cos(ph_ip)
            detJ_CS = r_ip ** 2 * s_th
            J1 = np.stack([r_ip * s_th * c_th * c_ph, -r_ip * s_ph, r_ip ** 2 * s_th ** 2 * c_ph], 1) / detJ_CS[:, None]
            J2 = np.stack([r_ip * s_th * c_th * s_ph, r_ip * c_ph, r_ip ** 2 * s_th ** 2 * s_ph], 1) / detJ_CS[:, None]
            J3 = np.stack([-r_ip * s_th ** 2, np.zeros(n), r_ip ** 2 * s_th * c_th], 1) / detJ_CS[:, None]
            Jx, Jy, Jz = E.elblk_AtxBt(J1, J2, J3, iJth, iJph, iJr)
            Jx, Jy, Jz = E.elblk_transpose_A(Jx, Jy, Jz)
            dx, dy, dz = Jx @ dN4ds.T, Jy @ dN4ds.T, Jz @ dN4ds.T
            weight = detJ_CS * detJ_SL * ipw[ip]
            idx = 0
            for i in range(4):
                for j in range(i, 4):
                    C[:, idx] += kappa * weight * (dx[:, i] * dx[:, j] + dy[:, i] * dy[:, j] + dz[:, i] * dz[:, j])
                    M[:, idx] += weight * N4[i] * N4[j]
                    idx += 1
        CC[els] = C
        MM[els] = M
    return CC, MM


_UPPER = [(i, j) for i in range(4) for j in range(i, 4)]


def assemble(EL4, Cv, n):
    """Global symmetric matrix (both halves) from packed element values."""
    I, J, V = [], [], []
    for idx, (i, j) in enumerate(_UPPER):
        I.append(EL4[:, i]); J.append(EL4[:, j]); V.append(Cv[:, idx])
        if i != j:
            I.append(EL4[:, j]); J.append(EL4[:, i]); V.append(Cv[:, idx])
    return sp.csr_matrix((np.concatenate(V), (np.concatenate(I), np.concatenate(J))), shape=(n, n))


def normdf(x):
    return np.linalg.norm(x) / np.sqrt(x.size)


def pcg_solver_p(A, b, x, tol=1e-8, is_abs_tol=False, itmax=500):
    """Jacobi-preconditioned CG of pcg_solver_p.m:49-186 ('diag')."""
    ADi = 1.0 / A.diagonal()
    x = x.copy()
    r = b - A @ x
    norm_r = normdf(r)
    tol = tol if is_abs_tol else tol * norm_r
    hist = [norm_r]
    if norm_r < tol or norm_r == 0:     # zero residual: nothing to solve (0/0 otherwise)
        return x, 0, np.array(hist), tol
    rd = rlast = q = None
    it = 0
    for it in range(1, itmax + 1):
        d = r * ADi
        if it == 1:
            q = d
        else:
            rdlast = rd
            beta = np.dot(r - rlast, d) / rdlast
            q = d + beta * q
        rd = np.dot(r, d)
        Aq = A @ q
        alpha = rd / np.dot(q, Aq)
        rlast = r
        x = x + alpha * q
        r = r - alpha * Aq
        norm_r = normdf(r)
        hist.append(norm_r)
        if norm_r < tol:
            break
    return x, it, np.array(hist), tol


def diffusion3d_p(T, kappa, X, SPH, EL4, ifix, vfix, dt, theta=1.0, lumping=True, nip=4,
                  tol=1e-8, is_abs_tol=False, itmax=500, theta_cone=45.0):
    """T_new, info (diffusion3d_p.m:45-116) on linear tetrahedra EL4 over the
    len(T) nodes; ifix 0-based, vfix values."""
    T = np.array(T, dtype=float)
    if kappa == 0:
        return T, dict(nit=0)
    n = len(T)
    Cv, Mv = element_matrices(X, SPH, EL4, kappa, nip, theta_cone)
    CC = assemble(EL4, Cv, n)
    MM = assemble(EL4, Mv, n)
    if lumping:
        MM = sp.diags(np.asarray(MM.sum(axis=1)).ravel()).tocsr()
    rhs = MM @ T - dt * (1 - theta) * (CC @ T)
    A = (dt * theta * CC + MM).tocsr()
    ifix = np.asarray(ifix, dtype=np.int64)
    ifree = np.setdiff1d(np.arange(n), ifix)
    if len(ifix):
        rhs = rhs - A[:, ifix] @ np.asarray(vfix, dtype=float)
    T[ifix] = vfix
    Aff = A[ifree][:, ifree].tocsr()
    x, nit, hist, tol_abs = pcg_solver_p(Aff, rhs[ifree], T[ifree], tol, is_abs_tol, itmax)
    T[ifree] = x
    return T, dict(nit=nit, hist=hist, tol=tol_abs, CC=CC, MM=MM, A=A, rhs=rhs, ifree=ifree)
