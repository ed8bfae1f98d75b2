"""The C++ harness m3bench drives the C ABI without Python: an exported c0
case solved by m3bench gives the solution the Python binding gets for the
same inputs, and meets the residual target."""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from tests.conftest import relerr


def test_m3bench_matches_the_python_binding(tmp_path, c0, c0_inputs):
    from m3tet_sph_amd import api
    exe = os.path.join(os.path.dirname(api.LIB_PATH), "m3bench")
    case = str(tmp_path / "case")
    api.export_case(case, c0, c0_inputs)
    out = str(tmp_path)
    r = subprocess.run([exe, case, "--steps", "2", "--warmup", "1", "--precision", "fp64", "--out", out],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["harness"] == "m3bench" and line["velocity_dofs"] == 3 * c0.nnod
    assert line["rel_res_U"] <= 1e-6 and line["rel_res_P"] <= 1e-6
    U = np.load(os.path.join(out, "U.npy"))
    P = np.load(os.path.join(out, "P.npy"))
    i = c0_inputs
    ctx = api.StokesContext(c0)
    try:
        ctx.set_bc(i["ifix"], i["vfix"])
        Up, Pp, info = ctx.flow3d(i["Visc"], i["Dens"], i["phys"], rtol_Pat=0.8e-6, rtol_U_final=0.5e-6)
    finally:
        ctx.close()
    assert info["nitZ"] == line["nitZ"] and info["nitP"] == line["nitP"]
    assert relerr(U, Up) < 1e-10 and relerr(P, Pp) < 1e-10
