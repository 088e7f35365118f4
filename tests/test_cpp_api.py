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
@pytest.mark.parametrize("tune", ["", "tune"])
def test_cpp_driver_matches_oracle(tmp_path, tune):
    """tune: Lattice::set_autotune before perform_lbm (the shape and
    occupancy change no bit)."""
    exe = os.path.join(CPP_DIR, "lid_cavity")
    assert os.path.exists(exe), "tests/cpp/lid_cavity is built by __graft_entry__.build()"
    steps = 500
    out = tmp_path / "pop.bin"
    r = subprocess.run([exe, str(out), str(steps)] + ([tune] if tune else []), capture_output=True, text=True,
                       cwd=tmp_path, timeout=120)
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


# ------------------------------------------------ file-driven lattices, output

GOLDEN = os.path.join(ROOT, "tests", "golden")


def _run(exe, args, cwd, timeout=300):
    path = os.path.join(CPP_DIR, exe)
    assert os.path.exists(path), f"tests/cpp/{exe} is built by __graft_entry__.build()"
    r = subprocess.run([path] + [str(a) for a in args], capture_output=True, text=True, cwd=cwd, timeout=timeout)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    return r.stdout


def _fmt(v):
    """A double as std::ostream prints it by default (6 significant digits)."""
    return f"{v:g}"


def _check_velocity_files(results, k, u, nr, nc):
    """results/output-<k>u_x.txt / u_y.txt (LatticeWriter.hpp:61-93): one row
    per lattice row, every value followed by a space.  u: (2, NC, NR)."""
    for comp, name in ((0, "u_x"), (1, "u_y")):
        lines = (results / f"output-{k}{name}.txt").read_text().splitlines()
        assert len(lines) == nr
        for i in (0, 1, nr // 2, nr - 1):
            want = "".join(_fmt(u[comp, j, i]) + " " for j in range(nc))
            assert lines[i] == want, (k, name, i)


def _lid_oracle(steps_cb=None):
    cls = L.geometry_cavity(129, 129)
    orc = O.OracleLattice(cls, collision=O.BGK, tau=0.881, inlet_policy=O.ZOUHE_VELOCITY, outlet_policy=O.ZOUHE_VELOCITY)
    return cls, orc


def _lid_speed(t):
    return 0.1 * (1.0 - math.exp(-t * 0.02))


@pytest.mark.gpu
def test_lattice_file_run_saves_and_getters(tmp_path):
    """Lattice(path) from the reference's lid-lattice.txt fixture
    (Lattice.hpp:207-239, LatticeReader.hpp:367-522) with a ramped lid: the
    getters and update_nodes before the run (InletUpdates.cpp:33-44, checked
    inside the driver); perform_lbm(300, 1, 50) writes output-<k>u_x/u_y
    from asynchronous downloads, each equal to the oracle's u of step 50 k
    as the reference's writer prints it; after the run get_global_rho/u
    are the fields of the last step and the populations match the oracle
    to 1e-12 (FP64)."""
    steps, every = 300, 50
    out = _run("lattice_file", [os.path.join(GOLDEN, "lid-lattice.txt"), steps, every, 0, "f64_"], tmp_path)
    assert "update_nodes checked" in out
    cls, orc = _lid_oracle()
    ni = orc.count(O.LIST_INLET)
    orc.init()
    results = tmp_path / "results"
    for i in range(steps):
        orc.set_open_values(0, ux=np.full(ni, _lid_speed(i)), uy=np.zeros(ni))
        orc.step(1)
        if i % every == 0:
            _, u = orc.macro()
            _check_velocity_files(results, i // every, u, 129, 129)
    assert not (results / f"output-{steps // every}u_x.txt").exists()
    nc, nr = cls.shape
    pop = np.fromfile(tmp_path / "f64_pop.bin").reshape(9, nc, nr)
    err = max_abs_wet(pop, orc.populations(), cls)
    rho_o, u_o = orc.macro()
    rho = np.fromfile(tmp_path / "f64_rho.bin").reshape(nc, nr)
    u = np.fromfile(tmp_path / "f64_u.bin").reshape(2, nc, nr)
    fl = cls == L.FLUID
    print(f"lattice file run vs oracle: max|dP| = {err:.3e}, max|d rho| = {np.max(np.abs(rho - rho_o)[fl]):.3e}",
          flush=True)
    assert err <= 1e-12
    assert np.max(np.abs(rho - rho_o)[fl]) <= 1e-12 and np.max(np.abs(u - u_o)[:, fl]) <= 1e-12
    # the lid keeps rho = 1 and the prescribed velocity of the last step (Q1)
    lid = cls == L.INLET
    assert np.all(rho[lid] == 1.0) and np.all(u[0][lid] == _lid_speed(steps - 1))


@pytest.mark.gpu
@pytest.mark.parametrize("storage,bound", [(1, 2e-5), (2, 3e-3)])
def test_lattice_file_shifted_storage(tmp_path, storage, bound):
    """The same file-built cavity stored as FP32 / FP16 shifted populations
    through Lattice::set_storage, 2000 steps: L2 relative velocity error of
    the fluid nodes against the FP64 run within the stated bound.  The lid
    ramps up from rest, so the FP16 scale taken from the first step's
    values is the capped 2^15, and the run stays inside the FP16 range
    (the driver's perform_lbm reports LBM_ERR_RANGE otherwise when built
    with LLALBM_DEBUG; here the range is checked by the bound)."""
    steps = 2000
    lid = os.path.join(GOLDEN, "lid-lattice.txt")
    _run("lattice_file", [lid, steps, 0, 0, "a_"], tmp_path)
    out = _run("lattice_file", [lid, steps, 0, storage, "b_"], tmp_path)
    scale = float(out.split("scale ")[1].split()[0])
    assert scale == (32768.0 if storage == 2 else 1.0)
    cls = L.geometry_cavity(129, 129)
    fl = cls == L.FLUID
    ua = np.fromfile(tmp_path / "a_u.bin").reshape(2, 129, 129)[:, fl]
    ub = np.fromfile(tmp_path / "b_u.bin").reshape(2, 129, 129)[:, fl]
    e = float(np.linalg.norm(ub - ua) / np.linalg.norm(ua))
    print(f"storage {storage}: L2 rel velocity error vs FP64 after {steps} steps = {e:.3e} (bound {bound})", flush=True)
    assert e <= bound


def _foil_classes():
    nr, nc = 100, 400
    cls = np.zeros((nc, nr), dtype=np.uint8)
    cls[0, :] = cls[-1, :] = L.BOUNDARY
    cls[:, 0] = cls[:, -1] = L.BOUNDARY
    cls[0, 1:99] = L.INLET
    cls[nc - 1, 1:99] = L.OUTLET
    pts = np.loadtxt(os.path.join(GOLDEN, "NACA-foil.txt"), skiprows=1, dtype=np.int64)
    cls[pts[:, 1], pts[:, 0]] = L.OBSTACLE
    return cls


@pytest.mark.gpu
def test_foil_channel_flow_analysis_and_output(tmp_path):
    """examples/8_1_NACAProfileFlight.cpp's set-up through the compat layer:
    the aerofoil read by ConstructionInfo::read_obstacle_from_file
    (ConstructionInfo.hpp:493-503, LatticeReader.hpp:532-575), TRT, ramped
    Zou-He inlet, no-op outlet, FlowAnalyzer on every obstacle node saving
    its CSVs (FlowAnalyzer.hpp:259-268,348-451) from asynchronous device
    reductions.  Every contribution_<k>.csv row and total_lift_drag.csv
    row equals the oracle's per-point / total lift and drag at step
    100 (k + 1) as printed; the velocity files equal the oracle's u; the
    populations match to 1e-12."""
    steps, save, every = 400, 100, 100
    out = _run("foil_channel", [os.path.join(GOLDEN, "NACA-foil.txt"), steps, save, every, tmp_path / "pop.bin"],
               tmp_path)
    cls = _foil_classes()
    nc, nr = cls.shape
    orc = O.OracleLattice(cls, collision=O.TRT, tau_even=0.9, tau_odd=1.2, inlet_policy=O.ZOUHE_VELOCITY,
                          outlet_policy=O.NOOP)
    ni, no = orc.count(O.LIST_INLET), orc.count(O.LIST_OUTLET)
    assert f"inlets {ni} outlets {no}" in out
    obst_ij, obst_mask = orc.node_list(O.LIST_OBSTACLE)
    assert f"obstacles {len(obst_ij)} " in out
    orc.set_open_values(1, ux=np.zeros(no), uy=np.zeros(no))
    orc.init()
    totals = []
    fa = tmp_path / "flow_analysis"
    for i in range(steps):
        orc.set_open_values(0, ux=np.full(ni, 0.15 * (1 - math.exp(-i * 0.001))), uy=np.zeros(ni))
        orc.step(1)
        if i % save == 0:
            _, u = orc.macro()
            _check_velocity_files(tmp_path / "results", i // save, u, nr, nc)
        if i > 0 and i % every == 0:
            lift, drag, per = orc.lift_drag(per_point=True)
            totals.append((lift, drag))
            rows = (fa / f"contribution_{len(totals) - 1}.csv").read_text().splitlines()
            assert rows[0] == "i,j,lift,drag" and len(rows) == len(obst_ij) + 1
            for q, row in enumerate(rows[1:]):
                assert row == f"{obst_ij[q][0]},{obst_ij[q][1]},{_fmt(per[q][0])},{_fmt(per[q][1])}", (i, q)
    rows = (fa / "total_lift_drag.csv").read_text().splitlines()
    assert rows[0] == "iteration,total_lift,total_drag" and len(rows) == len(totals) + 1
    for k, (lift, drag) in enumerate(totals):
        assert rows[k + 1] == f"{k},{_fmt(lift)},{_fmt(drag)}"
    pop = np.fromfile(tmp_path / "pop.bin").reshape(9, nc, nr)
    err = max_abs_wet(pop, orc.populations(), cls)
    print(f"foil channel vs oracle: max|dP| = {err:.3e}, {len(totals)} analyses", flush=True)
    assert err <= 1e-12


@pytest.mark.gpu
def test_policy_statics_reference_loop(tmp_path):
    """The reference's perform_lbm loop body written against the policy
    statics (OMPPolicy.hpp:50-163 order, Lattice.hpp:322-410) with
    HIPPolicy bound to the lattice's context: the populations after 300
    steps match the oracle to 1e-12, and update_nodes left the host
    velocity tensor holding the last step's inlet value."""
    steps = 300
    out = _run("policy_loop", [steps, tmp_path / "pop.bin"], tmp_path)
    nr, nc = 48, 120
    cls = np.zeros((nc, nr), dtype=np.uint8)
    cls[0, :] = cls[-1, :] = L.BOUNDARY
    cls[:, 0] = cls[:, -1] = L.BOUNDARY
    cls[0, 1:47] = L.INLET
    cls[nc - 1, 1:47] = L.OUTLET
    L.add_disc(cls, 22, 30, 6)
    te = 0.6
    to = (3.0 / 16.0) * (1.0 / (te - 0.5)) + 0.5
    orc = O.OracleLattice(cls, collision=O.TRT, tau_even=te, tau_odd=to, inlet_policy=O.ZOUHE_VELOCITY,
                          outlet_policy=O.NOOP)
    ij, _ = orc.node_list(O.LIST_INLET)
    y = ij[:, 0] - 0.5
    prof = 4.0 * 0.05 * y * (46.0 - y) / (46.0 * 46.0)
    orc.set_open_values(1, ux=np.zeros(orc.count(O.LIST_OUTLET)), uy=np.zeros(orc.count(O.LIST_OUTLET)))
    orc.init()
    for i in range(steps):
        orc.set_open_values(0, ux=prof * (1.0 - math.exp(-i * 0.01)), uy=np.zeros(len(ij)))
        orc.step(1)
    pop = np.fromfile(tmp_path / "pop.bin").reshape(9, nc, nr)
    err = max_abs_wet(pop, orc.populations(), cls)
    print(f"policy statics loop vs oracle: max|dP| = {err:.3e}", flush=True)
    assert err <= 1e-12
    y10 = 10 - 0.5
    want = 4.0 * 0.05 * y10 * (46.0 - y10) / (46.0 * 46.0) * (1.0 - math.exp(-(steps - 1) * 0.01))
    assert f"inlet u(10, 0) {want:.17g}" in out


@pytest.mark.gpu
def test_ps_obstacle_driver_matches_oracle(tmp_path):
    """test/unit/LidWithPSObstacle.cpp through the compat layer
    (HIPPSBounceBackPolicy as the obstacle policy, its tau from
    allowed_tau(0.02, 10), compute_obstacle_weight leaving b = 1 as the
    reference does) against the oracle's PSBounceBackPolicy over 400
    steps, to 1e-12."""
    steps = 400
    out = _run("ps_obstacle", [steps, tmp_path / "pop.bin"], tmp_path)
    cls = L.geometry_cavity(100, 100)
    L.add_disc(cls, 50, 50, 20)
    te = 0.9
    to = 0.25 * (1.0 / (te - 0.5)) + 0.5
    tau_ps = 0.5 + 0.02 / ((1.0 / 3.0) * 10)
    orc = O.OracleLattice(cls, collision=O.TRT, tau_even=te, tau_odd=to, inlet_policy=O.ZOUHE_VELOCITY,
                          outlet_policy=O.ZOUHE_VELOCITY)
    tok = out.split()
    assert int(tok[tok.index("obstacles") + 1]) == orc.count(O.LIST_OBSTACLE)
    assert float(tok[tok.index("tau_ps") + 1]) == tau_ps
    assert float(tok[tok.index("b0") + 1]) == 1.0          # the list keeps its weights
    assert float(tok[tok.index("min_block_weight") + 1]) < 1.0   # though partial blocks were computed
    orc.set_partial_bounce_back(tau_ps)
    ni = orc.count(O.LIST_INLET)
    orc.init()
    for i in range(steps):
        orc.set_open_values(0, ux=np.full(ni, 0.2 * (1.0 - math.exp(-((500 * 500 * i) / (2 * 1000 * 1000))))),
                            uy=np.zeros(ni))
        orc.step(1)
    pop = np.fromfile(tmp_path / "pop.bin").reshape(9, 100, 100)
    err = max_abs_wet(pop, orc.populations(), cls)
    print(f"PS obstacle driver vs oracle: max|dP| = {err:.3e}", flush=True)
    assert err <= 1e-12
