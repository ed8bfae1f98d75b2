"""Writes tests/golden/c0_vectors.npz from the oracle (NumPy restatement of
the reference) on the 3,840-tet shell: element-packed K/G/M/Fb of 6 elements
per class, rows of the transfer operator, the Galerkin coarse operator
applied to a seeded vector, and an MGPCG iterate history.  The reference
ships no golden vectors for this path, so these fixtures freeze the oracle
(pinned by the tests in test_oracle_*.py) for the CPU regression test and
the GPU comparison.

Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from m3tet_sph_amd import mesh as M, fields as F  # noqa: E402
from oracle import assembly as A, stokes as S  # noqa: E402


def main():
    m = M.build_config("c0")
    inp = F.make_inputs(m, visc_contrast=1e3)
    els = np.concatenate([m.classes[k][:6] for k in A.CLASS_NAMES])
    K, G, Mm, Fb = A.element_matrices(m, inp["Visc"], inp["Dens"], inp["phys"])
    ip, ci, w = m.transfer(0)
    nodes = np.arange(0, m.nnod, 97)
    rows = [np.stack([ci[ip[n]:ip[n + 1]], w[ip[n]:ip[n + 1]]]) for n in nodes]
    KK, Fbg, GG, MM, _ = A.assemble_opt(m, inp["Visc"], inp["Dens"], inp["phys"])
    RR = S.make_rotation_matrix(m.GCOORD)
    St = S.rotate_and_lift(KK, GG, Fbg, RR, inp["ifix"], inp["vfix"], np.zeros(m.nVnod))
    Ic = S.ic2f_dof(ip, ci, w, m.levels[1].nnod)
    KKl, IcU = S.restrict_stiffmat(St["K1"], St["ifree"], [Ic], [3 * L.nnod for L in m.levels])
    x = np.random.default_rng(7).standard_normal(KKl[1].shape[0])
    _, it, hist = S.mgpcg_solver_p(St["K1"], KKl, S.Chol(KKl[1]), St["Rhs"], np.zeros(len(St["ifree"])),
                                   IcU, 1e-8, False, 50)
    np.savez_compressed(os.path.join(HERE, "c0_vectors.npz"), els=els, K=K[els], G=G[els], M=Mm[els],
                        Fb=Fb[els], nodes=nodes, trans_rows=np.concatenate(rows, axis=1),
                        trans_len=np.array([r.shape[1] for r in rows]), x=x, K2x=KKl[1] @ x,
                        mg_it=it, mg_hist=hist)


if __name__ == "__main__":
    main()
