"""Oracle (and, on a GPU, the HIP path) against the committed c0 fixtures."""
import os

import numpy as np
import pytest

from tests.conftest import ROOT, relerr

GOLD = os.path.join(ROOT, "tests", "golden", "c0_vectors.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


def test_oracle_element_matrices_match_fixture(c0, c0_inputs, gold):
    from oracle import assembly as A, elements as E
    T = E.sf_tables()
    els = gold["els"]
    for kind in A.CLASS_NAMES:
        sel = np.isin(els, c0.classes[kind])
        K, G, M, F = A.class_kernel_opt(kind, els[sel], c0, c0_inputs["Visc"], c0_inputs["Dens"],
                                        c0_inputs["phys"], T)
        assert relerr(K, gold["K"][sel]) < 1e-14
        assert relerr(G, gold["G"][sel]) < 1e-14
        assert relerr(F, gold["Fb"][sel]) < 1e-14


def test_transfer_rows_match_fixture(c0, gold):
    ip, ci, w = c0.transfer(0)
    rows = np.split(gold["trans_rows"], np.cumsum(gold["trans_len"])[:-1], axis=1)
    for n, r in zip(gold["nodes"], rows):
        assert np.array_equal(ci[ip[n]:ip[n + 1]], r[0].astype(np.int64))
        assert np.array_equal(w[ip[n]:ip[n + 1]], r[1])


@pytest.mark.gpu
def test_hip_element_matrices_match_fixture(c0, c0_inputs, gold):
    from m3tet_sph_amd import api
    ctx = api.StokesContext(c0)
    K, G, M, F = ctx.element_matrices(c0_inputs["Visc"], c0_inputs["Dens"], c0_inputs["phys"])
    els = gold["els"]
    assert relerr(K[els], gold["K"]) < 1e-12
    assert relerr(G[els], gold["G"]) < 1e-12
    assert relerr(M[els], gold["M"]) < 1e-12
    assert relerr(F[els], gold["Fb"]) < 1e-12
    ctx.close()
