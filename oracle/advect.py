"""CPU restatement of the semi-Lagrangian temperature update (test
infrastructure only).

Follows the scheme of libm3stokes' m3_sl_advect, which restates the
predictor-corrector back-tracking of mfiles_SPH/advect3d_slm_p_sph.m:165-243
(backtrack_PC2, with the spherical projections of
predictor_step_with_projections / corrector_step_with_projections :298-440)
and the boundary shift of find_pts_outside_domain_sph: back-tracking along
the great circle of the tangential velocity, radial change -dt*u_r (u_r at
the predicted point in its own radial direction, so that a rigid rotation
keeps the radius), radius
clamped to the shell, linear interpolation in the straight P1 tetrahedron
found around the vertex.  Pure Python per vertex: small meshes only.
"""
import numpy as np


def tet_geometry(X, tets):
    """Inverse of [x1-x4, x2-x4, x3-x4] and x4 per tetrahedron."""
    P = X[tets]
    M = np.stack([P[:, 0] - P[:, 3], P[:, 1] - P[:, 3], P[:, 2] - P[:, 3]], axis=2)
    return np.linalg.inv(M), P[:, 3]


def stars(tets, nV):
    lists = [[] for _ in range(nV)]
    for e, t in enumerate(tets):
        for v in t:
            lists[v].append(e)
    return lists


def _bary(Minv, x4, p):
    l = Minv @ (p - x4)
    return np.array([l[0], l[1], l[2], 1.0 - l.sum()])


def locate(p, v, star, tets, Minv, x4):
    best, bt, bl = -np.inf, -1, None

    def scan(w):
        nonlocal best, bt, bl
        for t in star[w]:
            l = _bary(Minv[t], x4[t], p)
            if l.min() > best:
                best, bt, bl = l.min(), t, l

    scan(v)
    if best < -1e-12:
        for w in tets[bt]:
            if w != v:
                scan(w)
    if best < 0:
        bl = np.maximum(bl, 0.0)
        bl = bl / bl.sum()
    return bt, bl


def gc_back(xh, r, u, tau, r_ang, rh, r_in, r_out):
    """Back-track r*xh over tau: arc tau*|u_t|/r_ang along the tangential part
    of u at xh, radius change -tau*(u . rh)."""
    t = u - (u @ xh) * xh
    ur = u @ rh
    s = np.linalg.norm(t)
    y = xh.copy()
    if s > 0:
        al = tau * s / r_ang
        y = np.cos(al) * xh - np.sin(al) * (t / s)
    rn = min(max(r - tau * ur, r_in), r_out)
    return rn * y


def sl_advect(X, tets, T, uc, dt, T_cmb=1.0, T_surf=0.0):
    """X: vertex coordinates (nV x 3); tets: P1 tetrahedra (vertex ids);
    T: vertex temperatures; uc: Cartesian vertex velocities (nV x 3)."""
    nV = len(X)
    Minv, x4 = tet_geometry(X, tets)
    star = stars(tets, nV)
    rad = np.linalg.norm(X, axis=1)
    r_in, r_out = rad.min(), rad.max()
    out = np.empty(nV)
    tol = 1e-9 * r_out
    for v in range(nV):
        r = rad[v]
        if abs(r - r_in) < tol:
            out[v] = T_cmb
            continue
        if abs(r - r_out) < tol:
            out[v] = T_surf
            continue
        xh = X[v] / r
        xm = gc_back(xh, r, uc[v], 0.5 * dt, r, xh, r_in, r_out)
        t, l = locate(xm, v, star, tets, Minv, x4)
        um = l @ uc[tets[t]]
        rm = np.linalg.norm(xm)
        xd = gc_back(xh, r, um, dt, rm, xm / rm, r_in, r_out)
        t, l = locate(xd, v, star, tets, Minv, x4)
        out[v] = l @ T[tets[t]]
    return out


def departure_points(X, tets, uc_fn, dt):
    """Departure points of the scheme for an analytic velocity field
    ``uc_fn(points) -> velocities`` (interpolation replaced by evaluation),
    used to check the back-tracking's order of accuracy."""
    rad = np.linalg.norm(X, axis=1)
    out = np.empty_like(X)
    for v in range(len(X)):
        r = rad[v]
        xh = X[v] / r
        xm = gc_back(xh, r, uc_fn(X[v][None])[0], 0.5 * dt, r, xh, -np.inf, np.inf)
        rm = np.linalg.norm(xm)
        out[v] = gc_back(xh, r, uc_fn(xm[None])[0], dt, rm, xm / rm, -np.inf, np.inf)
    return out
