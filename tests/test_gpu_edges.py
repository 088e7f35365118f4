"""Edge cases of the FP64 drop-in path against the CPU oracle: tiny and
ragged lattice sizes (rows not a multiple of 64, single interior node),
inlets and outlets on every side of the domain (Zou-He velocity and
pressure closures for TOP, BOTTOM, LEFT, RIGHT in one lattice), fluid nodes
on the perimeter, obstacles touching the walls, and the shifted storage
formats on the smallest lattices.  Bar: bit-exact populations (checked at
<= 1e-12), identical lift/drag."""
import os
import sys

import numpy as np
import pytest

import lbm_amd as L
import workloads as W
from helpers import max_abs_wet, oracle_from, wet_mask

pytestmark = pytest.mark.gpu
TOL = 1e-12
THREADS = min(16, os.cpu_count() or 1)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def boxed(nr, nc):
    cls = np.full((nc, nr), L.FLUID, dtype=np.uint8)
    cls[0, :] = cls[-1, :] = L.BOUNDARY
    cls[:, 0] = cls[:, -1] = L.BOUNDARY
    return cls


def open_values(cls, which, rng, policy):
    n = len(W.open_rows(cls, which)[0])
    if policy == L.ZOUHE_PRESSURE:
        return dict(rho=1.0 + 1e-3 * rng.uniform(-1, 1, n))
    if policy == L.ZOUHE_VELOCITY:
        return dict(ux=0.03 * rng.uniform(0, 1, n), uy=0.01 * rng.uniform(-1, 1, n))
    return {}


def run_case(name, cls, collision, inlet_policy, outlet_policy, steps, seed=0, tau=0.7):
    rng = np.random.default_rng(seed)
    te, to = W.trt_taus((tau - 0.5) / 3.0)
    wl = W.Workload(name=name, node_class=cls, collision=collision, tau=tau, tau_even=te, tau_odd=to,
                    inlet_policy=inlet_policy, outlet_policy=outlet_policy,
                    inlet=open_values(cls, 0, rng, inlet_policy), outlet=open_values(cls, 1, rng, outlet_policy))
    say(f"case {name}: {cls.shape[1]}x{cls.shape[0]}, {steps} steps, collision {collision}, "
        f"inlet {inlet_policy}, outlet {outlet_policy}")
    # the same run with row-segment waves (default shape), wet-run waves and
    # tabled wet-run waves
    shapes = ("row segments", "wet runs", "tabled wet runs")
    gpus = [wl.make_lattice(L.F64) for _ in shapes]
    gpus[1].set_launch(L.LAUNCH_WET_RUNS)
    gpus[2].set_launch(L.LAUNCH_RUN_TABLE)
    orc = oracle_from(wl, threads=THREADS)
    # the prescribed values change half way (per-step update functions)
    half = steps // 2
    for g in gpus:
        g.step(half)
    orc.step(half)
    for which, pol in ((0, inlet_policy), (1, outlet_policy)):
        vals = open_values(cls, which, rng, pol)
        if vals:
            for g in gpus:
                g.set_open_values(which, **vals)
            orc.set_open_values(which, **vals)
    for g in gpus:
        g.step(steps - half, L.STEP_MACRO)
    orc.step(steps - half)
    po = orc.populations()
    assert np.all(np.isfinite(po[:, wet_mask(cls)])), f"{name}: the oracle run itself is not finite"
    for shape, gpu in zip(shapes, gpus):
        err = max_abs_wet(gpu.populations(), po, cls)
        say(f"  {shape}: max|dP| = {err:.3e}")
        assert err <= TOL, f"{name} ({shape}): {err}"
        if np.any(cls == L.OBSTACLE):
            lg, lo = gpu.lift_drag(), orc.lift_drag()
            assert abs(lg[0] - lo[0]) <= TOL and abs(lg[1] - lo[1]) <= TOL, (shape, lg, lo)
        rg, ug = gpu.macro()
        ro, uo = orc.macro()
        m = cls == L.FLUID
        assert np.max(np.abs(rg[m] - ro[m])) <= TOL and np.max(np.abs(ug[:, m] - uo[:, m])) <= TOL, shape
        gpu.close()
    orc.close()


@pytest.mark.parametrize("nr,nc", [(3, 3), (4, 5), (63, 65), (64, 64), (65, 63), (127, 130), (200, 70)])
def test_ragged_sizes_channel_with_obstacles(nr, nc):
    """Inlet column 0 (velocity), outlet column NC-1 (pressure), random
    obstacle nodes, some touching the walls."""
    cls = boxed(nr, nc)
    cls[0, 1:nr - 1] = L.INLET
    cls[nc - 1, 1:nr - 1] = L.OUTLET
    rng = np.random.default_rng(nr * 1000 + nc)
    if nr > 4 and nc > 4:
        inner = rng.random((nc - 4, nr - 2)) < 0.15
        cls[2:nc - 2, 1:nr - 1][inner] = L.OBSTACLE
    run_case(f"ragged{nr}x{nc}", cls, L.TRT, L.ZOUHE_VELOCITY, L.ZOUHE_PRESSURE, 300, seed=nr + nc)


@pytest.mark.parametrize("inlet_policy,outlet_policy", [(L.ZOUHE_VELOCITY, L.ZOUHE_PRESSURE),
                                                        (L.ZOUHE_PRESSURE, L.ZOUHE_VELOCITY),
                                                        (L.ZOUHE_VELOCITY, L.OPEN_NOOP)])
def test_open_boundaries_on_every_side(inlet_policy, outlet_policy):
    """Inlets on the top row and the left column, outlets on the bottom row
    and the right column: every Zou-He side closure in one lattice."""
    nr, nc = 90, 140
    cls = boxed(nr, nc)
    cls[1:nc - 1, 0] = L.INLET          # top row
    cls[0, 1:nr - 1] = L.INLET          # left column
    cls[1:nc - 1, nr - 1] = L.OUTLET    # bottom row
    cls[nc - 1, 1:nr - 1] = L.OUTLET    # right column
    L.add_disc(cls, 45, 60, 9)
    run_case(f"allsides{inlet_policy}{outlet_policy}", cls, L.BGK, inlet_policy, outlet_policy, 400, seed=7)


@pytest.mark.parametrize("nr", [40, 64, 65, 128])
def test_wet_top_and_bottom_rows(nr):
    """Wet nodes on row 0 and row NR-1 of every interior column: with
    NR <= 64 one 64-row segment holds both (the perimeter kernel takes it
    once); with NR = 65 the last segment holds row NR-1 alone; with NR = 128
    row NR-1 ends a full segment."""
    nc = 50
    cls = boxed(nr, nc)
    cls[1:nc - 1, 0] = L.INLET          # top row
    cls[1:nc - 1, nr - 1] = L.OUTLET    # bottom row
    L.add_disc(cls, nr // 2, 25, min(5, nr // 2 - 3))
    run_case(f"top_bottom{nr}", cls, L.BGK, L.ZOUHE_VELOCITY, L.ZOUHE_PRESSURE, 300, seed=11)


def test_fluid_on_the_perimeter():
    """Fluid nodes on the domain edge (no wall): the out-of-domain slots keep
    their initial values, the same definition in oracle and kernel."""
    nr, nc = 70, 66
    cls = np.full((nc, nr), L.FLUID, dtype=np.uint8)
    cls[0, :] = L.BOUNDARY
    cls[nc - 1, 10:20] = L.BOUNDARY
    cls[1:nc - 1, 0] = L.INLET
    L.add_disc(cls, 35, 30, 6)
    run_case("perimeter_fluid", cls, L.TRT, L.ZOUHE_VELOCITY, L.OPEN_NOOP, 300, seed=3)


@pytest.mark.parametrize("storage", [L.F32_SHIFTED, L.F16_SHIFTED])
@pytest.mark.parametrize("nr,nc", [(3, 3), (5, 4), (65, 66)])
def test_shifted_storage_on_tiny_lattices(storage, nr, nc):
    """The shifted formats on the smallest lattices: finite, and close to
    FP64 (within the format's resolution)."""
    cls = boxed(nr, nc)
    cls[0, 1:nr - 1] = L.INLET
    cls[nc - 1, 1:nr - 1] = L.OUTLET
    rng = np.random.default_rng(5)
    wl = W.Workload(name="tiny", node_class=cls, collision=L.BGK, tau=0.8, inlet_policy=L.ZOUHE_PRESSURE,
                    outlet_policy=L.ZOUHE_PRESSURE, inlet=open_values(cls, 0, rng, L.ZOUHE_PRESSURE),
                    outlet=open_values(cls, 1, rng, L.ZOUHE_PRESSURE))
    ref = wl.make_lattice(L.F64)
    lo = wl.make_lattice(storage)
    ref.step(200)
    lo.step(200, L.STEP_CHECK)
    err = max_abs_wet(lo.populations(), ref.populations(), cls)
    say(f"case tiny {nr}x{nc} storage {storage}: max|dP| vs FP64 = {err:.3e}")
    assert err <= (1e-6 if storage == L.F32_SHIFTED else 1e-3)
    ref.close()
    lo.close()


def compact_slots(cls):
    """Compact slot count of a node-class plane: each column starts at the
    first index congruent to the row of its first wet node modulo 64 and
    holds its wet nodes contiguously (the layout lbm_set_geometry builds)."""
    acc = 0
    for col in wet_mask(cls):
        rows = np.nonzero(col)[0]
        if len(rows):
            acc += (int(rows[0]) - acc) % 64
            acc += len(rows)
    return acc


def test_dry_lanes_past_the_last_wet_node():
    """The last wet node of the lattice ends a 64-slot group and its wave has
    dry lanes after it: those lanes read the slot one past the last one used
    (their own, for bounce-back), which every plane keeps spare."""
    nr, nc = 66, 10
    cls = boxed(nr, nc)
    cls[nc - 2, 30] = L.OBSTACLE          # shifts the last column's slots off the row numbering
    cls[0, 1:nr - 1] = L.INLET
    slots = compact_slots(cls)
    last_row = int(np.nonzero(wet_mask(cls)[nc - 2])[0][-1])
    assert slots % 64 == 0 and last_row % 64 != 63, (slots, last_row)
    run_case("dry_tail", cls, L.BGK, L.ZOUHE_VELOCITY, L.OPEN_NOOP, 200, seed=13)
    # the shifted formats read the same addresses
    for storage in (L.F32_SHIFTED, L.F16_SHIFTED):
        wl = W.Workload(name="dry_tail", node_class=cls, collision=L.BGK, tau=0.8, inlet_policy=L.ZOUHE_VELOCITY,
                        outlet_policy=L.OPEN_NOOP, inlet=dict(ux=np.full(nr - 2, 0.02), uy=np.zeros(nr - 2)),
                        outlet={})
        ref, lo = wl.make_lattice(L.F64), wl.make_lattice(storage)
        ref.step(200)
        lo.step(200, L.STEP_CHECK)
        err = max_abs_wet(lo.populations(), ref.populations(), cls)
        say(f"  storage {storage}: max|dP| vs FP64 = {err:.3e}")
        assert err <= (1e-6 if storage == L.F32_SHIFTED else 1e-3)
        ref.close()
        lo.close()
