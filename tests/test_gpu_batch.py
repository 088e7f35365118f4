"""Ensembles (lbm_batch): many independent lattices advanced by one
persistent kernel per step over a tile table, with HIP-graph replay of the
step loop and an autotuned tile shape.

Bar: every sample of an ensemble ends bit for bit where the same lattice
stepped alone ends (every storage format, samples in either population
layout), graph replays equal eager launches bit for bit across calls with
other work in between, the FP64 ensemble matches the CPU oracle to 1e-12,
and the per-sample FP64 reductions (and the Darcy permeability computed
from them) agree with the single-lattice reduction to 1e-12 relative."""
import sys

import numpy as np
import pytest

import lbm_amd as L
import workloads as W
from helpers import max_abs_wet, oracle_from

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def small_ensemble(n=6, seed=11, sizes=(128, 192, 256)):
    return W.ensemble(n_samples=n, seed=seed, sizes=sizes)


def run_alone(wl, storage, schedule):
    """schedule: list of step counts (one lbm_step call each)."""
    lat = wl.make_lattice(storage)
    for k in schedule:
        lat.step(k)
    p = lat.populations()
    lat.close()
    return p


@pytest.mark.parametrize("layouts", ["blocked", "mixed"])
@pytest.mark.parametrize("launch", ["tiles", "wet-runs"])
@pytest.mark.parametrize("storage", [L.F64, L.F32_SHIFTED, L.F16_SHIFTED])
def test_ensemble_equals_lattices_alone(storage, launch, layouts):
    """mixed: every other sample stores its populations in planes."""
    wls = small_ensemble()
    say(f"case ensemble of {len(wls)} ({', '.join(w.name for w in wls)}), storage {storage}, {launch}, {layouts}")
    lats = [w.make_lattice(storage, options=L.OPT_PLANAR_LAYOUT if layouts == "mixed" and k % 2 else 0)
            for k, w in enumerate(wls)]
    schedule = [1, 40, 7, 64]          # first step, graph + eager remainder, eager, graph replays
    with L.Batch(lats) as b:
        if launch == "wet-runs":
            b.set_launch(L.LAUNCH_WET_RUNS, 0)
            assert b.launch() == (L.LAUNCH_WET_RUNS, 0)
        for k in schedule:
            b.step(k, L.STEP_CHECK)
        t = [lat.time for lat in lats]
    assert t == [sum(schedule)] * len(lats)
    worst = 0.0
    for w, lat in zip(wls, lats):
        ref = run_alone(w, storage, [sum(schedule)])
        err = max_abs_wet(lat.populations(), ref, w.node_class)
        worst = max(worst, err)
        lat.close()
    say(f"  max |P_ensemble - P_alone| = {worst:.3e}")
    assert worst == 0.0


@pytest.mark.parametrize("launch", ["tiles", "wet-runs"])
def test_large_ensemble_equals_lattices_alone(launch):
    """Enough tiles (and wet runs) that the work-groups are dealt to the XCDs
    in chunks of 1024 (the remapping is the identity below 8 x 1024 groups):
    every sample still ends where it ends alone, bit for bit."""
    wls = small_ensemble(n=5, seed=23, sizes=(960, 1024))
    tiles = sum((w.nr + 63) // 64 * w.nc for w in wls)
    say(f"case large ensemble of {len(wls)} ({', '.join(w.name for w in wls)}), {tiles} tiles, FP32, {launch}")
    assert tiles // 4 >= 2 * 8 * 1024
    lats = [w.make_lattice(L.F32_SHIFTED) for w in wls]
    with L.Batch(lats) as b:
        if launch == "wet-runs":
            b.set_launch(L.LAUNCH_WET_RUNS, 0)
        else:
            b.set_launch(1, 0)
        b.step(1)
        b.step(40, L.STEP_CHECK)
    for w, lat in zip(wls, lats):
        ref = run_alone(w, L.F32_SHIFTED, [41])
        assert max_abs_wet(lat.populations(), ref, w.node_class) == 0.0, w.name
        lat.close()


@pytest.mark.parametrize("segments", [1, 2, 4, 8])
def test_every_tile_shape_gives_the_same_result(segments):
    wls = small_ensemble(n=4, seed=17)
    say(f"case tile shape R={segments}")
    lats = [w.make_lattice(L.F32_SHIFTED) for w in wls]
    with L.Batch(lats) as b:
        b.set_launch(segments, 64)     # a small persistent grid: many tiles per wave
        b.step(70)
    for w, lat in zip(wls, lats):
        ref = run_alone(w, L.F32_SHIFTED, [70])
        assert max_abs_wet(lat.populations(), ref, w.node_class) == 0.0
        lat.close()


@pytest.mark.parametrize("launch", ["tiles", "wet-runs"])
def test_graph_replays_equal_eager_launches(launch):
    wls = small_ensemble(n=5, seed=23)
    say(f"case graph vs eager ({launch}), 4 calls of 64 steps with reductions and downloads in between")
    a = [w.make_lattice(L.F32_SHIFTED) for w in wls]
    e = [w.make_lattice(L.F32_SHIFTED) for w in wls]
    with L.Batch(a) as ba, L.Batch(e) as be:
        if launch == "wet-runs":
            ba.set_launch(L.LAUNCH_WET_RUNS, 0)
        for call in range(4):
            ba.step(64)
            be.step(64, L.STEP_NO_GRAPH)
            ra, re = ba.reduce(), be.reduce()
            assert np.array_equal(ra, re), f"call {call}: reductions differ"
            for w, x, y in zip(wls, a, e):
                assert max_abs_wet(x.populations(), y.populations(), w.node_class) == 0.0, f"call {call}"
    for lat in a + e:
        lat.close()


@pytest.mark.parametrize("launch", ["tiles", "wet-runs"])
def test_fp64_ensemble_matches_oracle(launch):
    wls = W.ensemble(n_samples=3, seed=29, sizes=(128, 192))
    steps = 300
    say(f"case FP64 ensemble vs oracle (populations and permeability), {steps} steps, {launch}")
    lats = [w.make_lattice(L.F64) for w in wls]
    with L.Batch(lats) as b:
        if launch == "wet-runs":
            b.set_launch(L.LAUNCH_WET_RUNS, 0)
        b.step(steps)
        red = b.reduce()
    nu = (0.8 - 0.5) / 3.0
    for s, (w, lat) in enumerate(zip(wls, lats)):
        orc = oracle_from(w, threads=8)
        orc.step(steps)
        err = max_abs_wet(lat.populations(), orc.populations(), w.node_class)
        # per-sample Darcy permeability: ensemble FP64 reduction on the GPU
        # against the oracle's own FP64 reduction of the same sample
        k_gpu = W.darcy_permeability(red[s], nu, w.nc - 1, 0.001, w.nr * w.nc)
        k_orc = W.darcy_permeability(orc.reduce(), nu, w.nc - 1, 0.001, w.nr * w.nc)
        say(f"  {w.name}: max|dP| = {err:.3e}, permeability {k_gpu:.10e} (oracle {k_orc:.10e})")
        assert err <= 1e-12
        assert abs(k_gpu - k_orc) <= 1e-10 * abs(k_orc)
        orc.close()
        lat.close()


def test_per_sample_reductions_and_permeability():
    wls = small_ensemble(n=6, seed=31)
    steps = 500
    lats = [w.make_lattice(L.F64) for w in wls]
    with L.Batch(lats) as b:
        b.step(steps)
        red = b.reduce()
    nu = (0.8 - 0.5) / 3.0
    for s, (w, lat) in enumerate(zip(wls, lats)):
        single = lat.reduce()
        rel = np.abs(red[s] - single) / np.maximum(np.abs(single), 1e-300)
        assert np.all(rel[[0, 3, 4]] <= 1e-12), (s, red[s], single)
        assert np.all(np.abs(red[s][1:3] - single[1:3]) <= 1e-12 * np.abs(single[0])), (s, red[s], single)
        k_b = W.darcy_permeability(red[s], nu, w.nc - 1, 0.001, w.nr * w.nc)
        k_1 = W.darcy_permeability(single, nu, w.nc - 1, 0.001, w.nr * w.nc)
        say(f"  {w.name}: permeability {k_b:.6e} (alone {k_1:.6e})")
        assert abs(k_b - k_1) <= 1e-10 * abs(k_1) + 1e-18
        lat.close()


def test_autotune_keeps_the_state():
    wls = small_ensemble(n=4, seed=37)
    lats = [w.make_lattice(L.F32_SHIFTED) for w in wls]
    with L.Batch(lats) as b:
        b.step(10)
        before = [lat.populations() for lat in lats]
        seg, blocks, ms = b.autotune(steps=2)
        say(f"  autotune: R={seg}, {blocks} work-groups, {ms:.4f} ms/step")
        assert seg in (1, 2, 4, 8, L.LAUNCH_WET_RUNS) and blocks % 8 == 0 and ms > 0
        assert seg != L.LAUNCH_WET_RUNS or blocks == 0
        for x, lat in zip(before, lats):
            assert np.array_equal(x, lat.populations())
        b.step(30)
    for w, lat in zip(wls, lats):
        ref = run_alone(w, L.F32_SHIFTED, [40])
        assert max_abs_wet(lat.populations(), ref, w.node_class) == 0.0
        lat.close()


@pytest.mark.parametrize("launch", ["tiles", "wet-runs"])
def test_time_varying_inlets_in_an_ensemble(launch):
    """Channels with a velocity inlet that changes between calls, mixed with
    a porous sample: each ends where it ends alone with the same schedule."""
    wls = [W.channel(96, 256, radius=8, centre=(48, 60), re=20), W.channel(128, 192, radius=6, centre=(60, 50), re=20),
           small_ensemble(n=1, seed=41)[0]]
    lats = [w.make_lattice(L.F64) for w in wls]
    alone = [w.make_lattice(L.F64) for w in wls]

    def set_values(lat, w, k):
        if w.inlet_policy == L.ZOUHE_VELOCITY:
            lat.set_open_values(0, ux=w.inlet["ux"] * (1.0 - np.exp(-0.05 * (k + 1))), uy=w.inlet["uy"])

    with L.Batch(lats) as b:
        if launch == "wet-runs":
            b.set_launch(L.LAUNCH_WET_RUNS, 0)
        for k in range(5):
            for w, lat, lat1 in zip(wls, lats, alone):
                set_values(lat, w, k)
                set_values(lat1, w, k)
                lat1.step(37)
            b.step(37)
    for w, lat, lat1 in zip(wls, lats, alone):
        assert max_abs_wet(lat.populations(), lat1.populations(), w.node_class) == 0.0, w.name
        lat.close()
        lat1.close()


def test_membership_rules():
    wls = small_ensemble(n=2, seed=43)
    lats = [w.make_lattice(L.F64) for w in wls]
    b = L.Batch(lats)
    with pytest.raises(L.LbmError) as e:
        lats[0].step(1)
    assert e.value.code == 4
    with pytest.raises(L.LbmError) as e:
        lats[0].autotune(steps=1)       # the ensemble tunes itself (lbm_batch_autotune)
    assert e.value.code == 4
    with pytest.raises(L.LbmError):
        lats[0].close()
    with pytest.raises(L.LbmError):
        L.Batch([lats[0]])              # already in an ensemble
    b.step(3)
    lats[1].init_equilibrium()          # states now differ
    with pytest.raises(L.LbmError) as e:
        b.step(1)
    assert e.value.code == 4
    with pytest.raises(L.LbmError):
        b.step(1, L.STEP_MACRO)
    with pytest.raises(L.LbmError):
        b.set_launch(L.LAUNCH_WET_RUNS, 64)   # wet runs: one wave per run only
    b.close()
    lats[0].step(2)                     # free again
    for lat in lats:
        lat.close()
