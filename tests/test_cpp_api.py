"""The LLALBM-compatible C++ host layer (HIPPolicy, Lattice, ConstructionInfo,
FlowAnalyzer).  tests/cpp/lid_cavity.cpp is written like the reference's
examples/1_LidDrivenCavity129.cpp with the HIP policies; the GPU test runs it
and checks its populations and lift/drag history against the oracle driven
the same way (per-step inlet function, flow analysis every 100 steps)."""
import math
import os
import subprocess

import numpy as np
import pytest

import lbm_amd as L
import oracle as O
from conftest import ROOT
from helpers import max_abs_wet

CPP_DIR = os.path.join(ROOT, "tests", "cpp")
PKG_INC = os.path.join(ROOT, "latticeboltzmannmethod-fonnesu-grassi-guffanti_amd", "include")


def test_headers_compile_standalone(tmp_path):
    """Every public header compiles on its own (C++20, g++)."""
    hdrs = ["llalbm_hip.hpp"] + [f"llalbm_hip/{h}" for h in sorted(os.listdir(os.path.join(PKG_INC, "llalbm_hip")))]
    for h in hdrs:
        src = tmp_path / "t.cpp"
        src.write_text(f"#include <{h}>\nint main() {{ return 0; }}\n")
        r = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-Wall", "-I", PKG_INC, str(src)],
                           capture_output=True, text=True)
        assert r.returncode == 0, (h, r.stderr[-2000:])


def _cavity_with_disc():
    nr, nc = 65, 97
    cls = np.full((nc, nr), L.FLUID, dtype=np.uint8)
    cls[0, :] = cls[-1, :] = L.BOUNDARY
    cls[:, 0] = cls[:, -1] = L.BOUNDARY
    cls[1:nc - 1, 0] = L.INLET                 # row 0, columns 1..95
    ii, jj = np.meshgrid(np.arange(nr), np.arange(nc))
    cls[(ii - 30) ** 2 + (jj - 40) ** 2 <= 36] = L.OBSTACLE
    return cls


@pytest.mark.gpu
def test_cpp_driver_matches_oracle(tmp_path):
    exe = os.path.join(CPP_DIR, "lid_cavity")
    assert os.path.exists(exe), "tests/cpp/lid_cavity is built by __graft_entry__.build()"
    steps = 500
    out = tmp_path / "pop.bin"
    r = subprocess.run([exe, str(out), str(steps)], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [ln.split() for ln in r.stdout.splitlines() if ln.startswith("lift_drag")]
    cpp_ld = np.array([[float(a), float(b)] for _, a, b in lines])
    cls = _cavity_with_disc()
    nc, nr = cls.shape
    pop = np.fromfile(out, dtype=np.float64).reshape(9, nc, nr)

    te = 0.58
    to = 0.25 * (1.0 / (te - 0.5)) + 0.5      # enforce_magic_parameter() default 1/4
    orc = O.OracleLattice(cls, collision=O.TRT, tau_even=te, tau_odd=to, inlet_policy=O.ZOUHE_VELOCITY,
                          outlet_policy=O.NOOP)
    orc.init()
    ni = orc.count(O.LIST_INLET)
    ld = []
    for i in range(steps):
        # math.exp is the C library's exp, the one std::exp calls in the driver
        orc.set_open_values(0, ux=np.full(ni, 0.1 * (1.0 - math.exp(-i * 0.01))), uy=np.zeros(ni))
        orc.step(1)
        if i > 0 and i % 100 == 0:
            ld.append(orc.lift_drag())
    err = max_abs_wet(pop, orc.populations(), cls)
    print(f"C++ driver vs oracle: max|dP| = {err:.3e}", flush=True)
    assert err <= 1e-12
    assert cpp_ld.shape == (4, 2)
    ld_err = float(np.max(np.abs(cpp_ld - np.array(ld))))
    print(f"C++ driver vs oracle: max|d(lift, drag)| = {ld_err:.3e}", flush=True)
    assert ld_err <= 1e-12


def test_readers_and_writer_match_the_reference_formats(tmp_path):
    """The C++ lattice/obstacle readers and the result writer (host only):
    lid-lattice.txt (the reference's 129 x 129 fixture) reads as the cavity
    plane of lbm_geometry_cavity; wing1.txt marks the same nodes the Python
    reader of the parity tests marks; the writer emits one text row per
    lattice row of u_x / u_y."""
    exe = os.path.join(CPP_DIR, "io_check")
    if not os.path.exists(exe):
        subprocess.run(["make", "-s", "-C", CPP_DIR, "io_check"], check=True)
    golden = os.path.join(ROOT, "tests", "golden")
    out_cls, out_obst, results = tmp_path / "cls.bin", tmp_path / "obst.bin", tmp_path / "results"
    r = subprocess.run([exe, os.path.join(golden, "lid-lattice.txt"), str(out_cls), os.path.join(golden, "wing1.txt"),
                        "100", "400", str(out_obst), str(results)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "dims 129 129" in r.stdout
    cls = np.fromfile(out_cls, dtype=np.uint8).reshape(129, 129)      # (NC, NR) = ColMajor plane
    assert np.array_equal(cls, L.geometry_cavity(129, 129))
    pts = np.loadtxt(os.path.join(golden, "wing1.txt"), skiprows=1, dtype=np.int64)
    expect = np.zeros((400, 100), dtype=np.uint8)
    expect[pts[:, 1], pts[:, 0]] = L.OBSTACLE
    assert np.array_equal(np.fromfile(out_obst, dtype=np.uint8).reshape(400, 100), expect)
    assert f"obstacles {len(pts)}" in r.stdout
    ux = (results / "output-7u_x.txt").read_text().splitlines()
    uy = (results / "output-7u_y.txt").read_text().splitlines()
    assert len(ux) == 3 and len(uy) == 3
    assert [float(x) for x in ux[1].split()] == [0.25 * (4 + j) for j in range(4)]
    assert [float(x) for x in uy[2].split()] == [-(2.0 + j) for j in range(4)]
