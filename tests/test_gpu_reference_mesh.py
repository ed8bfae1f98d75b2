"""The Stokes path on the reference's own mesh (tests/golden/
EarthShell_n37k.mat, SETUP_TEST/MESHES: 26,360 elements with an embedded
high-resolution region), promoted and refined once by the set-up layer:
210,880 P2 elements.  The solve must reach rel.res 1e-6 in FP64 and in the
mixed form, and the two solutions must agree to the conditioning."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "EarthShell_n37k.mat")


def log(*a):
    print(*a, flush=True)


def test_solve_on_the_reference_mesh():
    from m3tet_sph_amd import api, fields as F, mesh as M
    m = M.build_from_file(GOLD, nmg=2, device=0)
    log("reference mesh", len(m.levels[0].EL2NOD), "elements,", 3 * m.nnod, "velocity dofs")
    inp = F.make_inputs(m, visc_contrast=1e3)
    ctx = api.StokesContext(m)
    try:
        ctx.set_bc(inp["ifix"], inp["vfix"])
        out = {}
        for form in ("fp64", "mixed"):
            U, P, info = ctx.flow3d(inp["Visc"], inp["Dens"], inp["phys"], rtol_Pat=0.8e-6, rtol_U_final=5e-7,
                                    precision=form)
            rU, fU, rP = api.residual(ctx, U, P)
            rel = max(rU / fU, rP / info["norm_rP0"])
            log(form, {k: info[k] for k in ("nitP", "nitU", "nitZ", "t_total")}, "rel_res", rel)
            assert rel <= 1e-6
            out[form] = (U, P)
        dU = np.linalg.norm(out["mixed"][0] - out["fp64"][0]) / np.linalg.norm(out["fp64"][0])
        log("mixed vs fp64", dU)
        assert dU < 1e-4
    finally:
        ctx.close()
