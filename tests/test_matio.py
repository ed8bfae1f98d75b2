"""MAT-file (Level 5) reader and writer of the snapshot / restart files.

Pinned against a file written by MATLAB itself (tests/golden/
EarthShell_n37k.mat: the reference's mesh, SETUP_TEST/MESHES, v7 compressed,
a struct of uint32 / double / int32 / char fields, doubles stored in
narrower integer types) and against SciPy's independent implementation in
both directions, compressed and uncompressed.
"""
import os

import numpy as np
import pytest
import scipy.io as sio

from m3tet_sph_amd import matio as MI

GOLD = os.path.join(os.path.dirname(__file__), "golden", "EarthShell_n37k.mat")


def _same(a, b):
    if isinstance(a, dict):
        return set(a) == set(b) and all(_same(a[k], b[k]) for k in a)
    if isinstance(a, str):
        return a == b
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and np.array_equal(a, b)


def test_reads_a_matlab_written_file():
    ours = MI.loadmat(GOLD)
    ref = sio.loadmat(GOLD, mat_dtype=True)["MESH"][0, 0]
    M = ours["MESH"]
    assert list(M) == list(ref.dtype.names)
    for f in ref.dtype.names:
        if f == "name":
            assert M[f] == str(ref[f][0]) == "EarthShell_n37092"
            continue
        assert M[f].shape == ref[f].shape, f
        assert np.array_equal(M[f], ref[f]), f
    assert M["EL2NOD"].dtype == np.uint32 and M["EL2NOD"].shape == (10, 26360)
    assert M["GCOORD"].shape == (3, 37092) and M["PointID"].dtype == np.int32
    assert M["r_cmb"].dtype == np.float64   # class double, stored as uint16


def _sample(rng):
    return {
        "T": rng.random(17),
        "M2": rng.standard_normal((4, 3)),
        "I": rng.integers(-5, 5, (3, 5)).astype(np.int32),
        "U8": rng.integers(0, 255, 9).astype(np.uint8),
        "B": rng.random((2, 3)) > 0.5,
        "s": "snapshot",
        "x": 2.5,
        "E": np.zeros((0, 0)),
        "VAR": {"T": rng.random(7), "P": rng.random(3), "nested": {"k": np.int64(3), "name": "ab"}},
    }


@pytest.mark.parametrize("compress", [True, False])
def test_scipy_reads_what_we_write(tmp_path, compress):
    d = _sample(np.random.default_rng(0))
    p = str(tmp_path / "a.mat")
    MI.savemat(p, d, compress=compress)
    r = sio.loadmat(p, mat_dtype=True, squeeze_me=False)
    assert np.array_equal(r["T"], d["T"].reshape(-1, 1))           # 1-D -> column
    assert np.array_equal(r["M2"], d["M2"])
    assert r["I"].dtype == np.int32 and np.array_equal(r["I"], d["I"])
    assert r["U8"].dtype == np.uint8 and np.array_equal(r["U8"], d["U8"].reshape(-1, 1))
    assert r["B"].dtype == bool and np.array_equal(r["B"], d["B"])
    assert str(r["s"][0]) == "snapshot"
    assert r["x"].shape == (1, 1) and r["x"][0, 0] == 2.5
    assert r["E"].shape == (0, 0)
    V = r["VAR"][0, 0]
    assert np.array_equal(V["T"], d["VAR"]["T"].reshape(-1, 1))
    assert str(V["nested"][0, 0]["name"][0]) == "ab"
    assert V["nested"][0, 0]["k"][0, 0] == 3


@pytest.mark.parametrize("compress", [True, False])
def test_we_read_what_scipy_writes(tmp_path, compress):
    rng = np.random.default_rng(1)
    d = {"A": rng.random((5, 4)), "v": rng.random(6), "i": np.arange(7, dtype=np.int32), "s": "hello",
         "S": {"a": rng.random((2, 2)), "b": np.uint16(7)}, "L": np.array([[True, False]])}
    p = str(tmp_path / "b.mat")
    sio.savemat(p, d, do_compression=compress, oned_as="column")
    r = MI.loadmat(p)
    assert np.array_equal(r["A"], d["A"])
    assert np.array_equal(r["v"], d["v"].reshape(-1, 1))
    assert np.array_equal(r["i"], d["i"].reshape(-1, 1)) and r["i"].dtype == np.int32
    assert r["s"] == "hello"
    assert np.array_equal(r["S"]["a"], d["S"]["a"]) and r["S"]["b"][0, 0] == 7
    assert r["L"].dtype == bool and np.array_equal(r["L"], d["L"])


def test_round_trip(tmp_path):
    d = _sample(np.random.default_rng(2))
    p = str(tmp_path / "c.mat")
    MI.savemat(p, d)
    r = MI.loadmat(p)
    assert _same(r["M2"], d["M2"]) and _same(r["I"], d["I"]) and _same(r["B"], d["B"])
    assert np.array_equal(r["T"].ravel(), d["T"]) and r["s"] == d["s"]
    assert np.array_equal(r["VAR"]["nested"]["k"], np.array([[3]]))


def test_output_and_restart_files(tmp_path):
    rng = np.random.default_rng(3)
    VAR = {"T": rng.random(11), "Ux": rng.random(11), "P": rng.random(4)}
    S, C = {"outdir": str(tmp_path)}, {"prefix": "Plume"}
    p = MI.write_output_file(S, VAR, C, 0.25, 12, 1e-3, 7)
    assert os.path.basename(p) == "Plume_Data_0007.mat"
    r = sio.loadmat(p)
    assert set(k for k in r if not k.startswith("__")) == {"istep", "time", "dt", "VAR"}
    V, t, istep, iplot = MI.load_restart_file(p)
    assert t == 0.25 and istep == 12 and iplot is None
    assert all(np.array_equal(V[k], VAR[k]) for k in VAR)
    p = MI.write_restart_file(S, VAR, C, 12, 8, 0.25, 1e-3, 2)
    assert os.path.basename(p) == "Plume_RestartFile2.mat"
    V, t, istep, iplot = MI.load_restart_file(p)
    assert (t, istep, iplot) == (0.25, 12, 8)


def test_rejects_bad_files(tmp_path):
    p = str(tmp_path / "bad.mat")
    with open(p, "wb") as f:
        f.write(b"not a mat file")
    with pytest.raises(MI.MatError):
        MI.loadmat(p)
    good = str(tmp_path / "g.mat")
    MI.savemat(good, {"A": np.arange(100.0)}, compress=False)
    with open(good, "rb") as f:
        buf = f.read()
    with open(p, "wb") as f:
        f.write(buf[:-40])
    with pytest.raises(MI.MatError):
        MI.loadmat(p)
    with pytest.raises(MI.MatError):
        MI.savemat(good, {"bad name": 1.0})
