"""GPU tests of the time-loop steps around the Stokes solve: the material law
and the time step on the device against oracle/material.py, assembly from the
device-resident fields, and a short time loop (material law -> solve -> dt ->
semi-Lagrangian step) that must match the same loop driven from host arrays."""
import os

import numpy as np
import pytest

from conftest import relerr
from oracle import material as OM

pytestmark = pytest.mark.gpu

ROCK = dict(Visc0=2.0, Dens0=3300.0, T_ref=0.5, alpha=3e-5, Ea=2.5e5, Va=3e-6, g=-9.81, L0=1e6,
            a_UM2LM=30.0, minVisc=1e-3, maxVisc=1e3)


@pytest.fixture(scope="module")
def ctx(c0, c0_inputs):
    from m3tet_sph_amd import api
    print("[gpu] StokesContext on c0 (time loop)", flush=True)
    c = api.StokesContext(c0)
    c.set_bc(c0_inputs["ifix"], c0_inputs["vfix"])
    yield c
    c.close()


def test_rock_prop_matches_oracle(c0, ctx):
    from m3tet_sph_amd import api
    T = np.random.default_rng(1).random(c0.nnod) * 900.0 + 300.0
    v, d = ctx.rock_prop(T, api.rock_params(**ROCK))
    vr, dr = OM.rock_prop(T, c0.GCOORD, **ROCK)
    assert relerr(v, vr) < 1e-13 and relerr(d, dr) < 1e-15
    assert (v >= 1e-3).all() and (v <= 1e3).all()
    with pytest.raises(api.M3Error):  # kJ/mol instead of J/mol (rock_prop.m:75-77)
        ctx.rock_prop(T, api.rock_params(Ea=250.0))


def test_calculate_dt_matches_oracle(c0, ctx):
    from m3tet_sph_amd.fields import rotation_frames
    rng = np.random.default_rng(2)
    U = rng.standard_normal(3 * c0.nnod)
    R = rotation_frames(c0.GCOORD)
    Uxyz = np.einsum("nij,nj->ni", R, U.reshape(-1, 3))
    dt = ctx.calculate_dt(U, 0.5, 1e9)
    assert dt == pytest.approx(OM.calculate_dt(Uxyz, c0.GCOORD, 0.5, 1e9), rel=1e-14)
    assert ctx.calculate_dt(np.zeros(3 * c0.nnod), 0.5, 3.0) == 3.0


def test_device_fields_assemble_the_same_operator(c0, c0_inputs, ctx):
    from m3tet_sph_amd import api
    T = c0_inputs["T"]
    rp = api.rock_params(Visc0=1.0, Dens0=1.0, T_ref=0.5, alpha=1.0)
    v, d = ctx.rock_prop(T, rp)
    ctx.assemble(v, d, c0_inputs["phys"])
    K_host = api.get_operator(ctx, 0).toarray()
    ctx.rock_prop(T, rp, download=False)
    ctx.assemble(None, None, c0_inputs["phys"])
    K_dev = api.get_operator(ctx, 0).toarray()
    # same inputs; only the order of the FP64 atomic scatter differs
    assert relerr(K_dev, K_host) < 1e-14


def test_short_time_loop_on_device_fields(c0, c0_inputs, ctx):
    from m3tet_sph_amd import api, learned as L
    nV = c0.nVnod
    _, _, parent, _ = L.vertex_graph(c0)
    rp = api.rock_params(Visc0=1.0, Dens0=1.0, T_ref=0.5, alpha=1.0, Ea=3e4, minVisc=1e-3)
    phys = c0_inputs["phys"]
    runs = []
    for device_fields in (False, True):
        Tv = c0_inputs["T"][:nV].copy()
        U = P = None
        hist = []
        for step in range(3):
            Tn = L.nodal_temperature(c0, Tv, parent)
            if device_fields:
                ctx.rock_prop(Tn, rp, download=False)
                U, P, info = ctx.flow3d(None, None, phys, P0=P, U0=U, rtol_Pat=1e-8, rtol_U_final=1e-9)
            else:
                v, d = OM.rock_prop(Tn, c0.GCOORD, **{k: getattr(rp, k) for k in
                                                      ("Visc0", "Dens0", "T_ref", "alpha", "Ea")}, minVisc=1e-3)
                U, P, info = ctx.flow3d(v, d, phys, P0=P, U0=U, rtol_Pat=1e-8, rtol_U_final=1e-9)
            dt = ctx.calculate_dt(U, 0.5, 1e9) * 0.05
            Tv = ctx.sl_advect(Tv, U, dt, 1.0, 0.0)
            hist.append((U.copy(), P.copy(), Tv.copy(), info["nitP"]))
        runs.append(hist)
    for (U1, P1, T1, n1), (U2, P2, T2, n2) in zip(*runs):
        assert relerr(U2, U1) < 1e-9 and relerr(P2, P1) < 1e-9 and relerr(T2, T1) < 1e-12


def test_timeloop_driver_snapshots_and_restart(tmp_path):
    """The time-loop driver on c0 with diffusion, a second solve of every
    step from the extrapolated start (formed on the device), snapshot /
    restart files in the reference's MAT layout, and a restart from the last
    snapshot."""
    import json
    from m3tet_sph_amd import matio, timeloop
    out = str(tmp_path)
    timeloop.main(["--config", "c0", "--steps", "3", "--kappa", "1e2", "--snapshots", "2", "--extrap", "2",
                   "--out", out])
    summ = json.load(open(os.path.join(out, "timeloop_c0.json")))["summary"]
    assert summ["max_rel_res"] <= 1e-6 and summ["thermal"]["nit"] > 0
    assert summ["extrap"]["max_rel_res"] <= 1e-6
    files = sorted(os.listdir(out))
    assert "c0_Data_0001.mat" in files and "c0_Data_0002.mat" in files
    assert "c0_RestartFile1.mat" in files and "c0_RestartFile2.mat" in files
    VAR, t, istep, iplot = matio.load_restart_file(os.path.join(out, "c0_RestartFile2.mat"))
    assert istep == 2 and iplot == 2 and t > 0
    assert set(VAR) >= {"T", "Ux", "Uy", "Uz", "Uth", "Uph", "Ur", "P", "Visc", "Dens"}
    out2 = str(tmp_path / "r")
    timeloop.main(["--config", "c0", "--steps", "1", "--restart", os.path.join(out, "c0_RestartFile2.mat"),
                   "--out", out2])
    assert os.path.exists(os.path.join(out2, "timeloop_c0.json"))
