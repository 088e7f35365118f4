"""GPU parity of the FP64 drop-in path against the CPU oracle (the
reference's Serial policy semantics), through the C ABI.

Bar: max |P_gpu - P_oracle| <= 1e-12 over wet-node population slots after
1000 steps (the engine is built to match bit for bit; the tests report
whether it does), plus rho/u of the macro fields, lift/drag and the FP64
reductions.  Wall/obstacle slots are excluded: the reference keeps
order-dependent scratch there (never read by a wet node)."""
import os
import sys

import numpy as np
import pytest

import lbm_amd as L
import oracle as O
import workloads as W
from conftest import ROOT
from helpers import max_abs_wet, oracle_from, wet_mask

pytestmark = pytest.mark.gpu
TOL = 1e-12
THREADS = min(16, os.cpu_count() or 1)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def compare_state(gpu, orc, cls, what):
    pg, po = gpu.populations(), orc.populations()
    assert np.all(np.isfinite(po[:, wet_mask(cls)])), f"{what}: the oracle run itself is not finite"
    err = max_abs_wet(pg, po, cls)
    exact = err == 0.0
    say(f"  {what}: max|dP| = {err:.3e} ({'bit-exact' if exact else 'not bit-exact'})")
    assert err <= TOL, f"{what}: {err}"
    return err


def run_pair(wl, steps, lift_every=0, values=None):
    """Steps the GPU engine and the oracle side by side; values(i) may
    return per-step inlet/outlet values (dict per list) like the reference's
    update functions."""
    say(f"case {wl.name}: {wl.nr}x{wl.nc}, {steps} steps")
    O.set_threads(THREADS)
    orc = oracle_from(wl)
    gpu = wl.make_lattice(L.F64)
    lifts = []
    for i in range(steps):
        if values is not None:
            v = values(i)
            for which, kw in v.items():
                orc.set_open_values(which, **kw)
                gpu.set_open_values(which, **kw)
        flags = L.STEP_MACRO if i == steps - 1 else 0
        if values is None and lift_every == 0:
            gpu.step(steps, L.STEP_MACRO)
            orc.step(steps)
            break
        gpu.step(1, flags)
        orc.step(1)
        if lift_every and i > 0 and i % lift_every == 0:
            lg = gpu.lift_drag(per_point=True)
            lo = orc.lift_drag(per_point=True)
            assert np.max(np.abs(lg[2] - lo[2])) <= TOL
            assert abs(lg[0] - lo[0]) <= 1e-12 * max(1.0, abs(lo[0])) and abs(lg[1] - lo[1]) <= 1e-12 * max(1.0, abs(lo[1]))
            lifts.append((lo[0], lo[1]))
    assert gpu.time == steps == orc.time
    return gpu, orc, lifts


def check_macro(gpu, orc, cls):
    rg, ug = gpu.macro()
    ro, uo = orc.macro()
    m = (cls == L.FLUID) | (cls == L.INLET) | (cls == L.OUTLET)
    assert np.max(np.abs(rg[m] - ro[m])) <= TOL
    assert np.max(np.abs(ug[:, m] - uo[:, m])) <= TOL


def test_cavity129_bgk_1000_steps():
    wl = W.cavity(129)
    gpu, orc, _ = run_pair(wl, 1000)
    compare_state(gpu, orc, wl.node_class, "cavity129 BGK")
    check_macro(gpu, orc, wl.node_class)
    red_g, red_o = gpu.reduce(), orc.reduce()
    assert np.allclose(red_g, red_o, rtol=1e-13, atol=1e-15)


def test_channel_trt_cylinder_lift_drag():
    # Re 20 keeps tau+ = 0.548 at this resolution (Re 200 would need the
    # full 512 x 2048 lattice to stay stable).
    wl = W.channel(64, 256, radius=6, centre=(33, 51), re=20.0)
    gpu, orc, lifts = run_pair(wl, 1000, lift_every=100)
    assert len(lifts) == 9
    compare_state(gpu, orc, wl.node_class, "channel TRT + pressure outlet")
    check_macro(gpu, orc, wl.node_class)
    ag, ao = gpu.after_collision(), orc.after_collision()
    assert max_abs_wet(ag, ao, wl.node_class) <= TOL


def test_wing_time_varying_inlet_noop_outlet():
    """examples/7_FlatBottomFoil.cpp: 100 x 400 channel, wing1.txt obstacle
    file, TRT (0.6, 1.67), inlet u = 0.15 (1 - exp(-0.001 t)), no-op outlet,
    lift/drag every 100 steps."""
    nr, nc = 100, 400
    cls = np.full((nc, nr), L.FLUID, dtype=np.uint8)
    cls[:, 0] = cls[:, -1] = L.BOUNDARY
    cls[0, :] = cls[-1, :] = L.BOUNDARY
    cls[0, 1:nr - 1] = L.INLET
    cls[-1, 1:nr - 1] = L.OUTLET
    pts = np.loadtxt(os.path.join(GOLDEN, "wing1.txt"), skiprows=1, dtype=np.int64)
    cls[pts[:, 1], pts[:, 0]] = L.OBSTACLE
    wl = W.Workload(name="wing1", node_class=cls, collision=L.TRT, tau_even=0.6, tau_odd=1.67,
                    inlet_policy=L.ZOUHE_VELOCITY, outlet_policy=L.OPEN_NOOP)
    ni = int(np.count_nonzero(cls == L.INLET))

    def values(t):
        return {0: dict(ux=np.full(ni, 0.15 * (1 - np.exp(-t * 0.001))), uy=np.zeros(ni)),
                1: dict(ux=0.0, uy=0.0)}

    gpu, orc, lifts = run_pair(wl, 1000, lift_every=100, values=values)
    compare_state(gpu, orc, cls, "wing1 time-varying inlet")
    check_macro(gpu, orc, cls)


def test_porous_pressure_bgk():
    wl = W.porous(128, seed=11, rmin=3, rmax=8, nc=160)
    gpu, orc, _ = run_pair(wl, 1000)
    compare_state(gpu, orc, wl.node_class, "porous 128x160 pressure BGK")
    check_macro(gpu, orc, wl.node_class)
    red_g, red_o = gpu.reduce(), orc.reduce()
    assert np.allclose(red_g, red_o, rtol=1e-12, atol=1e-14)


def test_channel_full_size_101_steps():
    wl = W.channel()  # 512 x 2048, cylinder r = 26 at (260, 410)
    gpu, orc, _ = run_pair(wl, 101, lift_every=100)
    compare_state(gpu, orc, wl.node_class, "channel 512x2048")


def test_upload_download_roundtrip_and_state_machine():
    wl = W.channel(40, 60, radius=5, centre=(20, 25))
    gpu = L.Lattice(wl.node_class, **wl.lattice_kwargs())
    with pytest.raises(L.LbmError):
        gpu.step(1)                     # no populations yet
    rng = np.random.default_rng(5)
    p = rng.uniform(0.01, 0.2, size=(9, wl.nc, wl.nr))
    gpu.upload_populations(p)
    q = gpu.populations()
    m = wet_mask(wl.node_class)
    assert np.array_equal(q[:, m], p[:, m])
    assert np.all(q[:, ~m] == 0.0)
    with pytest.raises(L.LbmError):
        gpu.after_collision()           # no step yet
    # an uploaded state steps like the oracle from the same state
    orc = O.OracleLattice(wl.node_class, **wl.lattice_kwargs())
    orc.set_populations(p)
    gpu.step(20)
    orc.step(20)
    assert max_abs_wet(gpu.populations(), orc.populations(), wl.node_class) <= TOL


def test_nan_is_reported():
    wl = W.cavity(33)
    gpu = wl.make_lattice()
    p = gpu.populations()
    p[4, 10, 12] = np.nan
    gpu.upload_populations(p)
    gpu.step(3)
    with pytest.raises(L.LbmError) as e:
        gpu.step(1, L.STEP_CHECK)
    assert e.value.code == 3


def test_invalid_geometries_are_rejected():
    cls = L.geometry_channel(16, 16)
    bad = cls.copy()
    bad[5, 5] = L.INLET                 # inlet inside the domain
    with pytest.raises(L.LbmError) as e:
        L.Lattice(bad)
    assert e.value.code == 1
    bad = cls.copy()
    bad[5, 5] = L.SOLID                 # SOLID next to fluid
    with pytest.raises(L.LbmError) as e:
        L.Lattice(bad)
    assert e.value.code == 5


def test_graph_replay_equals_eager_steps():
    """lbm_step replays captured step loops for runs of >= 32 steps; with
    new prescribed values and downloads between calls, every call must end
    bit for bit where eager launches end."""
    wl = W.channel(64, 256, radius=6, centre=(32, 60), re=20)
    say("case graph replay vs eager, channel 64x256, 5 calls of 70 steps")
    g = wl.make_lattice(L.F64)
    e = wl.make_lattice(L.F64)
    for call in range(5):
        scale = 1.0 - 0.5 ** (call + 1)
        for lat in (g, e):
            lat.set_open_values(0, ux=wl.inlet["ux"] * scale, uy=wl.inlet["uy"])
        g.step(70)
        e.step(70, L.STEP_NO_GRAPH)
        assert np.array_equal(g.populations(), e.populations()), f"call {call}"
        assert g.lift_drag() == e.lift_drag()
    orc = oracle_from(wl, threads=THREADS)
    for call in range(5):
        scale = 1.0 - 0.5 ** (call + 1)
        orc.set_open_values(0, ux=wl.inlet["ux"] * scale, uy=wl.inlet["uy"])
        orc.step(70)
    compare_state(g, orc, wl.node_class, "graph-replayed channel")
    for lat in (g, e):
        lat.close()
    orc.close()


@pytest.mark.parametrize("storage", [L.F32_SHIFTED, L.F16_SHIFTED])
def test_occupancy_changes_between_graph_replays(storage):
    """Captured step loops are rebuilt when the occupancy (or the shape)
    changes: replays at one occupancy, then another, end bit for bit where
    eager launches end, with downloads and reductions in between."""
    wl = W.porous(300, seed=7, nc=260)
    say(f"case occupancy and shape changes between graph replays, porous 300x260, storage {storage}")
    g = wl.make_lattice(storage)
    e = wl.make_lattice(storage)
    for occ, shape in ((6, 1), (3, L.LAUNCH_WET_RUNS), (0, 2), (7, L.LAUNCH_RUN_TABLE)):
        g.set_occupancy(occ)
        g.set_launch(shape)
        g.step(70)
        e.step(70, L.STEP_NO_GRAPH)
        assert np.array_equal(g.populations(), e.populations()), (occ, shape)
        assert np.array_equal(g.reduce(), e.reduce()), (occ, shape)
    for lat in (g, e):
        lat.close()


def test_autotune_keeps_state_and_results():
    """The autotuner (launch shape and occupancy) must not change the state,
    and every tile shape at every occupancy must give the same bits."""
    wl = W.porous(256, seed=5, nc=320)
    say("case autotune, tile shapes (row segments, wet runs) and occupancies, porous 256x320 FP64/FP32/FP16")
    for storage in (L.F64, L.F32_SHIFTED, L.F16_SHIFTED):
        ref = wl.make_lattice(storage)
        ref.step(80)
        for seg in (1, 2, 4, 8, L.LAUNCH_WET_RUNS, L.LAUNCH_RUN_TABLE):
            for occ in (0, 7, 5, 3, 2):
                lat = wl.make_lattice(storage)
                lat.step(30)
                before = lat.populations()
                if seg == 1 and occ == 0:
                    chosen, ms = lat.autotune(steps=2)
                    assert chosen in (1, 2, 4, 8, L.LAUNCH_WET_RUNS, L.LAUNCH_RUN_TABLE) and ms > 0
                    assert lat.occupancy() in (0, 7, 6, 5, 4, 3)
                assert np.array_equal(before, lat.populations())
                lat.set_launch(seg)
                lat.set_occupancy(occ)
                assert lat.occupancy() == occ
                lat.step(50)
                assert np.array_equal(lat.populations(), ref.populations()), (storage, seg, occ)
                lat.close()
        ref.close()


def test_occupancy_defaults_and_range():
    """Until set or tuned the occupancy is the storage format's default (7
    work-groups per CU for FP32 and FP16, 5 for FP64); 0 (no limit) and 2..7
    are accepted, anything else is rejected and leaves it unchanged."""
    wl = W.porous(128, seed=5, nc=128)
    for storage, default in ((L.F64, 5), (L.F32_SHIFTED, 7), (L.F16_SHIFTED, 7)):
        lat = wl.make_lattice(storage)
        try:
            assert lat.occupancy() == default
            for bad in (-1, 1, 8, 64):
                with pytest.raises(L.LbmError) as e:
                    lat.set_occupancy(bad)
                assert e.value.code == 1
            assert lat.occupancy() == default
            for ok in (0, 2, 7):
                lat.set_occupancy(ok)
                assert lat.occupancy() == ok
        finally:
            lat.close()


def test_new_geometry_under_the_tabled_shape():
    """lbm_set_geometry on a context whose launch shape reads run tables
    (LBM_LAUNCH_RUN_TABLE) rebuilds the tables for the new geometry: the run
    on the second geometry is bit for bit the run of a fresh lattice with
    that shape, and matches the oracle."""
    a = W.porous(256, seed=5, nc=320)
    b = W.porous(256, seed=9, nc=320)
    assert not np.array_equal(a.node_class, b.node_class)
    say("case new geometry on a tabled-run context, porous 256x320 (two obstacle sets), 200 steps")
    lat = a.make_lattice(L.F64)
    lat.set_launch(L.LAUNCH_RUN_TABLE)
    lat.step(20)
    lat.set_geometry(b.node_class)
    assert lat.launch() == L.LAUNCH_RUN_TABLE
    lat.set_open_values(0, **b.inlet)
    lat.set_open_values(1, **b.outlet)
    lat.init_equilibrium()
    lat.step(200)
    fresh = b.make_lattice(L.F64)
    fresh.set_launch(L.LAUNCH_RUN_TABLE)
    fresh.step(200)
    assert np.array_equal(lat.populations(), fresh.populations())
    orc = oracle_from(b, threads=THREADS)
    orc.step(200)
    compare_state(lat, orc, b.node_class, "second geometry, tabled runs")
    for x in (lat, fresh, orc):
        x.close()


def test_pattern_probe_measures_plausible_rates():
    """The access-pattern ceilings the benchmark reports: every pattern and
    element size gives a finite rate between a floor no working HBM stream
    misses and the part's 8 TB/s nominal peak (4096^2 lattice: two buffers
    of 0.6-2.4 GB together, beyond the last-level cache)."""
    say("case access-pattern probe, 4096x4096, every pattern and element size")
    for size in (2, 4, 8):
        for p in sorted(L.PATTERN_NAMES):
            gbs = L.pattern_bandwidth(p, size, 4096, 4096, reps=10)
            say(f"  {L.PATTERN_NAMES[p]} {size} B: {gbs:.0f} GB/s")
            assert np.isfinite(gbs) and 200.0 < gbs < 8000.0, (p, size, gbs)


def test_copy_probe_measures_a_plausible_rate():
    """The copy-kernel ceiling the benchmark reports: positive, below the
    part's 8 TB/s nominal peak, and above a floor no working HBM stream
    misses."""
    say("case copy probe, 1 GiB (4x the last-level cache)")
    gbs = L.copy_bandwidth(1 << 30, reps=10)
    say(f"  copy kernel: {gbs:.0f} GB/s")
    assert 1000.0 < gbs < 8000.0
