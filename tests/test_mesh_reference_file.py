"""The mesh set-up layer against the reference's own mesh file
(tests/golden/EarthShell_n37k.mat, SETUP_TEST/MESHES, written by the
reference's MATLAB tools): promoting its vertex mesh EL2NOD(1:4,:) to curved
P2 with the cone half-aperture 45 degrees reproduces the file's 10-node
connectivity (find_unique_bars numbering, local edge order), its P2 node
coordinates (curved mid-side nodes in the class frames) and PointID."""
import os

import numpy as np

from m3tet_sph_amd import mesh as M

GOLD = os.path.join(os.path.dirname(__file__), "golden", "EarthShell_n37k.mat")


def test_p2_promotion_reproduces_the_reference_file():
    X, EL, pid, r_cmb, r_surf = M.load_mesh_sph(GOLD)
    assert (r_cmb, r_surf) == (3471.0, 6371.0) and EL.shape == (26360, 10) and X.shape == (37092, 3)
    nV = int(EL[:, :4].max()) + 1
    X4, EL4 = X[:nV], np.ascontiguousarray(EL[:, :4])
    Xp, SPH, EL10, pidp, classes, inc, bars = M.p1_to_p2_sph(X4, M.cart2sph(X4), EL4, pid[:nV], r_cmb, r_surf,
                                                             45.0)
    assert np.array_equal(EL10, EL)
    assert np.array_equal(pidp, pid)
    assert np.abs(Xp - X).max() < 1e-12 * r_surf
    # a different cone moves the mid-side nodes of the elements near the poles
    Xq = M.p1_to_p2_sph(X4, M.cart2sph(X4), EL4, pid[:nV], r_cmb, r_surf, 30.0)[0]
    assert np.abs(Xq - X).max() > 1.0


def test_hierarchy_from_the_reference_file():
    m = M.build_from_file(GOLD, nmg=2)
    assert m.nmg == 2 and m.levels[1].nnod == 37092 and len(m.levels[0].EL2NOD) == 8 * 26360
    assert abs(m.r_cmb - 3471 / 6371) < 1e-15 and m.r_surf == 1.0
    r = np.linalg.norm(m.GCOORD, axis=1)
    assert r.min() > m.r_cmb - 1e-12 and r.max() < 1.0 + 1e-12
    # vertices of the fine level are the coarse P2 nodes, numbered first
    assert m.levels[0].nVnod == m.levels[1].nnod
