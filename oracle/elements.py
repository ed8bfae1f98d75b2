"""Element numerics of the reference, restated (test infrastructure only).

* ``ip_tetrahedron(14)``: mfiles_M3TET/ip_tetrahedron.m:100-122,189-191
* ``sf_dsf_tet`` 4/10/20 nodes: mfiles_M3TET/sf_dsf_tet.m:909-951, 821-905, 247-374
* MILAMIN block 3x3 algebra: elblk_detA.m:3-8, elblk_invert_A.m:6-20,
  elblk_AtxBt.m:45-55, elblk_transpose_A.m:24-26,
  mfiles_SPH/elblk_cartesian2spherical.m:45-48
"""
import numpy as np


def ip_tetrahedron(nip=14):
    """Gellert & Harbord degree-5 rule (ip_tetrahedron.m:100-122); weights /6
    (ip_tetrahedron.m:189-191).  Returns (ipx (3,nip), ipw (nip,))."""
    if nip != 14:
        raise ValueError("only the 14-point rule is on the hot path")
    g = 1 / (46 * np.sqrt(46))
    h = np.arccos(g) + (2 / 3) * np.arcsin(g)
    k = (104 + 8 * np.sqrt(46) * np.cos(h)) / 3
    s = np.sqrt(49 - k)
    b = (7 + s) / k
    a = 1 - 3 * b
    d = (7 - s) / k
    c = 1 - 3 * d
    p = (98 - k - 14 * s) / (1680 * s * (b - a) ** 3)
    q = (98 - k + 14 * s) / (1680 * s * (c - d) ** 3)
    r = (1 - 4 * (p + q)) / 6
    e = (1 + (2 / (105 * r)) ** 0.25) / 4
    f = (1 - 2 * e) / 2
    ipx = np.array([[a, b, b, b, c, d, d, d, e, e, e, f, f, f],
                    [b, a, b, b, d, c, d, d, e, f, f, e, e, f],
                    [b, b, a, b, d, d, c, d, f, e, f, e, f, e]])
    ipw = np.array([p, p, p, p, q, q, q, q, r, r, r, r, r, r]) / 6.0
    return ipx, ipw


def sf_tet4(r, s, t):
    """(sf_dsf_tet.m:909-951) -> N (4,), dN (4,3)."""
    u = 1 - r - s - t
    N = np.array([r, s, t, u])
    dN = np.array([[1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0], [-1.0, -1.0, -1.0]])
    return N, dN


def sf_tet10(r, s, t):
    """(sf_dsf_tet.m:821-905) -> N (10,), dN (10,3)."""
    u = 1 - r - s - t
    N = np.array([r * (2 * r - 1), s * (2 * s - 1), t * (2 * t - 1), u * (2 * u - 1),
                  4 * r * s, 4 * s * t, 4 * t * u, 4 * r * u, 4 * r * t, 4 * s * u])
    dN = np.array([[4 * r - 1, 0, 0], [0, 4 * s - 1, 0], [0, 0, 4 * t - 1],
                   [1 - 4 * u, 1 - 4 * u, 1 - 4 * u],
                   [4 * s, 4 * r, 0], [0, 4 * t, 4 * s], [-4 * t, -4 * t, 4 * (u - t)],
                   [4 * (u - r), -4 * r, -4 * r], [4 * t, 0, 4 * r], [-4 * s, 4 * (u - s), -4 * s]])
    return N, dN


def sf_tet20(r, s, t):
    """(sf_dsf_tet.m:247-374) -> N (20,), dN (20,3)."""
    u = 1 - r - s - t
    N = np.array([
        r * (3 * r - 2) * ((3 * r) / 2 - 1 / 2), s * (3 * s - 2) * ((3 * s) / 2 - 1 / 2),
        t * (3 * t - 2) * ((3 * t) / 2 - 1 / 2), u * (3 * u - 2) * ((3 * u) / 2 - 1 / 2),
        r * s * ((27 * r) / 2 - 9 / 2), r * s * ((27 * s) / 2 - 9 / 2),
        s * t * ((27 * s) / 2 - 9 / 2), s * t * ((27 * t) / 2 - 9 / 2),
        t * u * ((27 * t) / 2 - 9 / 2), t * u * ((27 * u) / 2 - 9 / 2),
        r * u * ((27 * u) / 2 - 9 / 2), r * u * ((27 * r) / 2 - 9 / 2),
        r * t * ((27 * r) / 2 - 9 / 2), r * t * ((27 * t) / 2 - 9 / 2),
        s * u * ((27 * u) / 2 - 9 / 2), s * u * ((27 * s) / 2 - 9 / 2),
        27 * s * t * u, 27 * r * t * u, 27 * r * s * u, 27 * r * s * t])
    d4 = (25 / 2 - 27 * r * s - 27 * r * t - 27 * s * t - (27 * r ** 2) / 2
          - (27 * s ** 2) / 2 - (27 * t ** 2) / 2 - 18 * u)
    dN = np.array([
        [(27 * r ** 2) / 2 - 9 * r + 1, 0, 0],
        [0, (27 * s ** 2) / 2 - 9 * s + 1, 0],
        [0, 0, (27 * t ** 2) / 2 - 9 * t + 1],
        [d4, d4, d4],
        [(9 * s * (6 * r - 1)) / 2, (9 * r * (3 * r - 1)) / 2, 0],
        [(9 * s * (3 * s - 1)) / 2, (9 * r * (6 * s - 1)) / 2, 0],
        [0, (9 * t * (6 * s - 1)) / 2, (9 * s * (3 * s - 1)) / 2],
        [0, (9 * t * (3 * t - 1)) / 2, (9 * s * (6 * t - 1)) / 2],
        [-(9 * t * (3 * t - 1)) / 2, -(9 * t * (3 * t - 1)) / 2,
         (9 * r) / 2 + (9 * s) / 2 + 36 * t - 27 * r * t - 27 * s * t - (81 * t ** 2) / 2 - 9 / 2],
        [-(9 * t * (6 * u - 1)) / 2, -(9 * t * (6 * u - 1)) / 2,
         (9 * t) / 2 + (9 * u * (3 * u - 1)) / 2 - 27 * t * u],
        [(9 * r) / 2 + (9 * u * (3 * u - 1)) / 2 - 27 * r * u, -(9 * r * (6 * u - 1)) / 2,
         -(9 * r * (6 * u - 1)) / 2],
        [36 * r + (9 * s) / 2 + (9 * t) / 2 - 27 * r * s - 27 * r * t - (81 * r ** 2) / 2 - 9 / 2,
         -(9 * r * (3 * r - 1)) / 2, -(9 * r * (3 * r - 1)) / 2],
        [(9 * t * (6 * r - 1)) / 2, 0, (9 * r * (3 * r - 1)) / 2],
        [(9 * t * (3 * t - 1)) / 2, 0, (9 * r * (6 * t - 1)) / 2],
        [-(9 * s * (6 * u - 1)) / 2, (9 * s) / 2 + (9 * u * (3 * u - 1)) / 2 - 27 * s * u,
         -(9 * s * (6 * u - 1)) / 2],
        [-(9 * s * (3 * s - 1)) / 2,
         (9 * r) / 2 + 36 * s + (9 * t) / 2 - 27 * r * s - 27 * s * t - (81 * s ** 2) / 2 - 9 / 2,
         -(9 * s * (3 * s - 1)) / 2],
        [-27 * s * t, -27 * t * (s - u), -27 * s * (t - u)],
        [-27 * t * (r - u), -27 * r * t, -27 * r * (t - u)],
        [-27 * s * (r - u), -27 * r * (s - u), -27 * r * s],
        [27 * s * t, 27 * r * t, 27 * r * s]])
    return N, dN


def sf_tables(nip=14):
    """Shape-function tables at the IPs: dict of NU (nip,10), dNUds (nip,10,3),
    NP (nip,4), dN4ds (4,3), NU20 (nip,20), dNUds20 (nip,20,3), IP_w (nip,)
    (assembly_stokes_sph_3d_double_jacobian.m:342-360)."""
    ipx, ipw = ip_tetrahedron(nip)
    T = dict(IP_w=ipw, NU=[], dNUds=[], NP=[], NU20=[], dNUds20=[])
    for ip in range(nip):
        r, s, t = ipx[:, ip]
        N, dN = sf_tet10(r, s, t)
        T["NU"].append(N)
        T["dNUds"].append(dN)
        N4, dN4 = sf_tet4(r, s, t)
        T["NP"].append(N4)
        N20, dN20 = sf_tet20(r, s, t)
        T["NU20"].append(N20)
        T["dNUds20"].append(dN20)
    for k in ("NU", "dNUds", "NP", "NU20", "dNUds20"):
        T[k] = np.array(T[k])
    T["dN4ds"] = sf_tet4(0.25, 0.25, 0.25)[1]
    return T


# ---- MILAMIN-style block 3x3 algebra: Ax[:, :] is column 1 of A -----------

def elblk_detA(Jx, Jy, Jz):
    """elblk_detA.m:3-8."""
    return (Jx[:, 0] * Jy[:, 1] * Jz[:, 2] + Jx[:, 1] * Jy[:, 2] * Jz[:, 0]
            + Jx[:, 2] * Jy[:, 0] * Jz[:, 1] - Jx[:, 2] * Jy[:, 1] * Jz[:, 0]
            - Jx[:, 0] * Jy[:, 2] * Jz[:, 1] - Jx[:, 1] * Jy[:, 0] * Jz[:, 2])


def elblk_invert_A(Jx, Jy, Jz, detJ):
    """elblk_invert_A.m:6-20."""
    idet = 1.0 / detJ
    iJx = np.stack([idet * (Jy[:, 1] * Jz[:, 2] - Jy[:, 2] * Jz[:, 1]),
                    idet * (Jy[:, 2] * Jz[:, 0] - Jy[:, 0] * Jz[:, 2]),
                    idet * (Jy[:, 0] * Jz[:, 1] - Jy[:, 1] * Jz[:, 0])], axis=1)
    iJy = np.stack([idet * (Jx[:, 2] * Jz[:, 1] - Jx[:, 1] * Jz[:, 2]),
                    idet * (Jx[:, 0] * Jz[:, 2] - Jx[:, 2] * Jz[:, 0]),
                    idet * (Jx[:, 1] * Jz[:, 0] - Jx[:, 0] * Jz[:, 1])], axis=1)
    iJz = np.stack([idet * (Jx[:, 1] * Jy[:, 2] - Jx[:, 2] * Jy[:, 1]),
                    idet * (Jx[:, 2] * Jy[:, 0] - Jx[:, 0] * Jy[:, 2]),
                    idet * (Jx[:, 0] * Jy[:, 1] - Jx[:, 1] * Jy[:, 0])], axis=1)
    return iJx, iJy, iJz


def elblk_AtxBt(Ax, Ay, Az, Bx, By, Bz):
    """C = A' * B' in block format (elblk_AtxBt.m:45-55)."""
    Cx = np.stack([Ax[:, 0] * Bx[:, 0] + Ax[:, 1] * By[:, 0] + Ax[:, 2] * Bz[:, 0],
                   Ay[:, 0] * Bx[:, 0] + Ay[:, 1] * By[:, 0] + Ay[:, 2] * Bz[:, 0],
                   Az[:, 0] * Bx[:, 0] + Az[:, 1] * By[:, 0] + Az[:, 2] * Bz[:, 0]], axis=1)
    Cy = np.stack([Ax[:, 0] * Bx[:, 1] + Ax[:, 1] * By[:, 1] + Ax[:, 2] * Bz[:, 1],
                   Ay[:, 0] * Bx[:, 1] + Ay[:, 1] * By[:, 1] + Ay[:, 2] * Bz[:, 1],
                   Az[:, 0] * Bx[:, 1] + Az[:, 1] * By[:, 1] + Az[:, 2] * Bz[:, 1]], axis=1)
    Cz = np.stack([Ax[:, 0] * Bx[:, 2] + Ax[:, 1] * By[:, 2] + Ax[:, 2] * Bz[:, 2],
                   Ay[:, 0] * Bx[:, 2] + Ay[:, 1] * By[:, 2] + Ay[:, 2] * Bz[:, 2],
                   Az[:, 0] * Bx[:, 2] + Az[:, 1] * By[:, 2] + Az[:, 2] * Bz[:, 2]], axis=1)
    return Cx, Cy, Cz


def elblk_AxB(Ax, Ay, Az, Bx, By, Bz):
    """C = A * B in block format (the elblk_AxB.m product; used by the
    known-answer test together with elblk_AtxBt)."""
    A = np.stack([Ax, Ay, Az], axis=2)      # A[:, i, j]
    B = np.stack([Bx, By, Bz], axis=2)
    C = np.einsum("eik,ekj->eij", A, B)
    return C[:, :, 0], C[:, :, 1], C[:, :, 2]


def elblk_transpose_A(Ax, Ay, Az):
    """elblk_transpose_A.m:24-26."""
    return (np.stack([Ax[:, 0], Ay[:, 0], Az[:, 0]], axis=1),
            np.stack([Ax[:, 1], Ay[:, 1], Az[:, 1]], axis=1),
            np.stack([Ax[:, 2], Ay[:, 2], Az[:, 2]], axis=1))


def elblk_cartesian2spherical(x, y, z):
    """elblk_cartesian2spherical.m:45-48 (phi in [0, 2pi))."""
    r = np.sqrt(x ** 2 + y ** 2 + z ** 2)
    th = np.arctan2(np.sqrt(x ** 2 + y ** 2), z)
    ph = np.arctan2(y, x)
    ph = np.where(ph < 0, ph + 2 * np.pi, ph)
    return th, ph, r
