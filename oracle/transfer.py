"""Inter-mesh transfer operator, restated literally (test infrastructure only).

intermesh_transfer_matrix.m:8-62 on the straight geometry of
create_multigrid_meshes_3d_sph.m:54-64,79-82: every fine node is located in
the coarse element ceil(el/8) of the last fine element that holds it, its
local coordinates are computed from the straight (affine) coarse element,
clamped to [0,1] and rounded to 1e-3, and the 10 P2 shape functions are
evaluated there.  The product (m3tet_sph_amd.mesh.ShellMesh.transfer) builds
the same operator from the refinement tables; this module pins it.
"""
import numpy as np

from . import elements as E

EDGES = np.array([[0, 1], [1, 2], [2, 3], [3, 0], [0, 2], [3, 1]])


def straight_coords(mesh, l_coarse):
    """Straight coordinates of level l_coarse's P2 nodes and, refined, of
    level l_coarse-1 (vertices = coarse vertices, mid nodes = midpoints)."""
    def straighten(X_vert, EL):
        nn = int(EL.max()) + 1
        X = np.zeros((nn, 3))
        X[:len(X_vert)] = X_vert
        for j, (a, b) in enumerate(EDGES):
            X[EL[:, 4 + j]] = 0.5 * (X[EL[:, a]] + X[EL[:, b]])
        return X
    top = mesh.nmg - 1
    Xs = straighten(mesh.GCOORD[:mesh.levels[top].nVnod], mesh.levels[top].EL2NOD)
    for l in range(top - 1, l_coarse - 1, -1):
        Xs = straighten(Xs, mesh.levels[l].EL2NOD)
    Xf = straighten(Xs, mesh.levels[l_coarse - 1].EL2NOD)
    return Xs, Xf


def local_coords(Xc, ELc, els_c, P):
    """Barycentric (r,s,t) of points P in straight coarse elements."""
    V = Xc[ELc[els_c, :4]]                 # (n,4,3)
    A = np.stack([V[:, 0] - V[:, 3], V[:, 1] - V[:, 3], V[:, 2] - V[:, 3]], axis=2)
    return np.linalg.solve(A, (P - V[:, 3])[..., None])[..., 0]


def intermesh_transfer_matrix(mesh, l):
    """Node-level interpolation fine level l <- coarse level l+1 as a dense
    (nnod_f, 10) weight table and (nnod_f, 10) coarse node ids."""
    Xc, Xf = straight_coords(mesh, l + 1)
    EL = mesh.levels[l].EL2NOD
    ELc = mesh.levels[l + 1].EL2NOD
    nf = mesh.levels[l].nnod
    flat = EL.ravel()
    loc = np.zeros(nf, dtype=np.int64)
    loc[flat] = np.arange(len(flat))       # last occurrence wins (column-major 'last')
    els = loc // 10
    els_c = els // 8
    lc = local_coords(Xc, ELc, els_c, Xf[:nf])
    tol = 1e-5
    if lc.min() < -tol or lc.max() > 1 + tol or lc.sum(axis=1).max() > 1 + tol:
        raise ValueError("Calculation of interpolation matrix failed.")
    lc = np.clip(lc, 0, 1)
    lc = np.round(1000 * lc) / 1000
    N = np.array([E.sf_tet10(*x)[0] for x in lc])
    return N, ELc[els_c]
