"""Mesh set-up on the GPU: find_unique_bars (m3_unique_bars) against the
host numbering, and whole hierarchies built with it against the host build."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


def _host(pairs):
    from m3tet_sph_amd import mesh as M
    return M.unique_first_order(pairs)


@pytest.mark.parametrize("n,nodes", [(1, 2), (1000, 30), (200_000, 5_000), (1_000_000, 400_000)])
def test_unique_bars_match_the_host(n, nodes):
    from m3tet_sph_amd import api
    rng = np.random.default_rng(n)
    pairs = rng.integers(0, nodes, (n, 2)).astype(np.int32)
    first, bar_of = api.unique_bars(pairs, 0)
    ub, inv = _host(pairs.astype(np.int64))
    assert np.array_equal(pairs[first], ub)
    assert np.array_equal(bar_of, inv)
    assert np.all(np.diff(first) > 0)           # first-appearance order


def test_unique_bars_edge_cases():
    from m3tet_sph_amd import api
    first, bar_of = api.unique_bars(np.zeros((0, 2), np.int32), 0)
    assert len(first) == 0 and len(bar_of) == 0
    p = np.array([[3, 7], [7, 3], [3, 7], [1, 1]], np.int32)   # both orientations, a loop
    first, bar_of = api.unique_bars(p, 0)
    assert list(first) == [0, 3] and list(bar_of) == [0, 0, 0, 1]
    with pytest.raises(api.M3Error):
        api.unique_bars(np.array([[-1, 2]], np.int32), 0)


@pytest.mark.parametrize("name", ["c0", "c1", "c4"])
def test_hierarchy_built_on_the_gpu_equals_the_host_build(name):
    from m3tet_sph_amd import mesh as M
    t = time.perf_counter()
    mg = M.build_config(name, device=0)
    tg = time.perf_counter() - t
    t = time.perf_counter()
    mh = M.build_config(name)
    th = time.perf_counter() - t
    log(f"{name}: hierarchy set-up {th:.2f} s host, {tg:.2f} s with GPU bar numbering")
    assert mg.nmg == mh.nmg
    for a, b in zip(mg.levels, mh.levels):
        assert np.array_equal(a.EL2NOD, b.EL2NOD) and a.nnod == b.nnod and a.nVnod == b.nVnod
    assert np.array_equal(mg.GCOORD, mh.GCOORD) and np.array_equal(mg.PointID, mh.PointID)
    assert all(np.array_equal(mg.classes[k], mh.classes[k]) for k in mh.classes)
    assert np.array_equal(mg.X20_iso, mh.X20_iso)
