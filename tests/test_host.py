"""CPU tests of the product's host side: the C-ABI library loads and exports
every entry point include/lbm_amd.h declares, and the host-only geometry
helpers build the reference's node sets (checked against the reference's
own lattice fixture and against brute force)."""
import os
import re

import numpy as np
import pytest

import lbm_amd as L
import workloads as W
from conftest import ROOT

GOLDEN = os.path.join(ROOT, "tests", "golden")


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "lbm_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(lbm_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(L.EXPORTED)
    assert "gfx950" in L.version()


def test_library_has_gfx950_code_object():
    blob = open(L.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def read_lattice_file(path):
    """Reader for the reference's lattice text format
    (LatticeReader.hpp:367-522): dims, extents, six counts, then 'i j type'
    for every non-fluid node."""
    with open(path) as f:
        lines = [ln.split() for ln in f if ln.strip()]
    assert int(lines[0][0]) == 2
    nr, nc = int(lines[1][0]), int(lines[1][1])
    counts = [int(x) for x in lines[2]]
    cls = np.zeros((nc, nr), dtype=np.uint8)
    for i, j, t in lines[3:3 + sum(counts) - counts[0]]:
        cls[int(j), int(i)] = int(t)
    return cls, counts


def test_cavity_geometry_matches_reference_fixture():
    """test/assets/lid-lattice.txt is the reference's 129x129 lid-driven
    cavity; the cavity helper must produce the same node classes."""
    ref, counts = read_lattice_file(os.path.join(GOLDEN, "lid-lattice.txt"))
    assert counts == [16129, 0, 385, 127, 0, 0]
    got = L.geometry_cavity(129, 129)
    assert np.array_equal(got, ref)


def test_channel_geometry():
    cls = L.geometry_channel(7, 11)
    assert np.all(cls[:, 0] == L.BOUNDARY) and np.all(cls[:, -1] == L.BOUNDARY)
    assert np.all(cls[0, 1:-1] == L.INLET) and np.all(cls[-1, 1:-1] == L.OUTLET)
    assert np.all(cls[1:-1, 1:-1] == L.FLUID)


def test_disc_rasterisation_and_bounds():
    cls = L.geometry_channel(40, 50)
    added = L.add_disc(cls, 20, 25, 7)
    ii, jj = np.meshgrid(np.arange(40), np.arange(50))
    inside = (ii - 20) ** 2 + (jj - 25) ** 2 <= 49
    assert added == inside.sum()
    assert np.all(cls[inside] == L.OBSTACLE)
    assert not np.any(cls[~inside] == L.OBSTACLE)
    with pytest.raises(L.LbmError):
        L.add_disc(cls, 7, 25, 7)  # would touch row 0 (ConstructionInfo.hpp:388-402)


def test_porous_generator():
    a = L.geometry_channel(300, 260)
    b = L.geometry_channel(300, 260)
    fa, na = L.geometry_porous(a, 7, 4, 16, 0.3)
    fb, nb = L.geometry_porous(b, 7, 4, 16, 0.3)
    assert np.array_equal(a, b) and na == nb and fa == fb
    assert 0.3 <= fa < 0.3 + (np.pi * 16 ** 2 + 1) / (300 * 260)
    # discs never touch the perimeter
    solid = a == L.OBSTACLE
    assert not solid[:, 0].any() and not solid[:, -1].any() and not solid[0].any() and not solid[-1].any()
    c = L.geometry_channel(300, 260)
    L.geometry_porous(c, 8, 4, 16, 0.3)
    assert not np.array_equal(a, c)


def test_workload_builders():
    wl = W.channel(64, 256, radius=6, centre=(33, 51))
    assert wl.collision == L.TRT and wl.outlet_policy == L.ZOUHE_PRESSURE
    # tau- from the magic parameter (TRTCollision.hpp:130)
    assert np.isclose((wl.tau_even - 0.5) * (wl.tau_odd - 0.5), 3 / 16)
    ux = wl.inlet["ux"]
    assert len(ux) == 62 and np.isclose(ux.max(), 0.04, rtol=1e-3) and ux.min() > 0
    cav = W.cavity()
    assert np.isclose(cav.tau, 3 * 0.1 * 127 / 100 + 0.5)
    assert cav.wet_count() == 127 * 127 + 127


def test_copy_probe_rejects_bad_sizes_before_touching_a_device():
    """lbm_copy_bandwidth validates its arguments first: these calls make no
    HIP call, so they run without a GPU."""
    for nbytes, reps in ((1000, 5), (1 << 20, 0), (1 << 20, 5000)):
        with pytest.raises(L.LbmError) as e:
            L.copy_bandwidth(nbytes, reps)
        assert e.value.code == 1


def test_pattern_probe_rejects_bad_arguments_before_touching_a_device():
    """lbm_pattern_bandwidth validates its arguments first (no HIP call)."""
    bad = ((9, 4, 64, 64, 5), (L.PATTERN_AB, 3, 64, 64, 5), (L.PATTERN_AA, 4, 1, 64, 5),
           (L.PATTERN_AA, 4, 63, 3, 5), (L.PATTERN_BLOCKED, 8, 64, 64, 1), (L.PATTERN_AB, 8, 1 << 20, 1 << 20, 5))
    for pattern, size, nr, nc, reps in bad:
        with pytest.raises(L.LbmError) as e:
            L.pattern_bandwidth(pattern, size, nr, nc, reps)
        assert e.value.code == 1


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU: --gpus N > 1 exits with a message before
    any device is touched."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr + r.stdout
