"""Partially saturated bounce-back obstacles (PSBounceBackPolicy<2>,
PSBounceBackPolicy.hpp:177-297) on the GPU against the oracle, through the
C ABI: the reference's own test set-up (test/unit/LidWithPSObstacle.cpp:
100 x 100 cavity with a disc of radius 20, TRT 0.9 with magic 1/4, policy
tau from allowed_tau(0.02, 10), weights left at 1) and per-obstacle
weights, BGK and TRT.  Bar: bit-exact populations (checked at <= 1e-12),
post-collision populations and lift/drag; the argument and state rules of
lbm_set_partial_bounce_back."""
import math
import os
import sys

import numpy as np
import pytest

import lbm_amd as L
import oracle as O
from helpers import max_abs_wet, wet_mask

pytestmark = pytest.mark.gpu
TOL = 1e-12
THREADS = min(16, os.cpu_count() or 1)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def lid_with_disc():
    """LidWithPSObstacle.cpp:59-69: 100 x 100, perimeter walls, inlet on row 0
    columns 1..98, disc of radius 20 at (50, 50)."""
    cls = L.geometry_cavity(100, 100)
    L.add_disc(cls, 50, 50, 20)
    return cls


def lid_speed(t):
    # LidWithPSObstacle.cpp:46-53 (x component on row 0)
    return 0.2 * (1.0 - math.exp(-((500 * 500 * t) / (2 * 1000 * 1000))))


def run_case(cls, kwargs, tau_ps, b, steps, values, shapes=(0, L.LAUNCH_WET_RUNS)):
    O.set_threads(THREADS)
    orc = O.OracleLattice(cls, **kwargs)
    gpus = []
    for shape in shapes:
        g = L.Lattice(cls, storage=L.F64, **kwargs)
        if shape:
            g.set_launch(shape)
        gpus.append(g)
    ni = len(gpus[0].open_nodes(0))
    for i in range(steps):
        ux, uy = values(i, ni)
        if i == 0:
            orc.set_open_values(0, ux=ux, uy=uy)
            orc.set_partial_bounce_back(tau_ps, b)
            orc.init()
            for g in gpus:
                g.set_open_values(0, ux=ux, uy=uy)
                g.init_equilibrium()
                g.set_partial_bounce_back(tau_ps, b)
        else:
            orc.set_open_values(0, ux=ux, uy=uy)
            for g in gpus:
                g.set_open_values(0, ux=ux, uy=uy)
        orc.step(1)
        for g in gpus:
            g.step(1)
    return orc, gpus


def check(orc, gpus, cls, name):
    po = orc.populations()
    assert np.all(np.isfinite(po[:, wet_mask(cls)])), f"{name}: oracle run not finite"
    ao = orc.after_collision()
    lo = orc.lift_drag()
    for k, g in enumerate(gpus):
        err = max_abs_wet(g.populations(), po, cls)
        erra = max_abs_wet(g.after_collision(), ao, cls)
        lg = g.lift_drag()
        say(f"  {name} shape {k}: max|dP| = {err:.3e}, max|dF*| = {erra:.3e}, lift/drag {lg[0]:.6e}/{lg[1]:.6e}")
        assert err <= TOL and erra <= TOL
        assert abs(lg[0] - lo[0]) <= TOL and abs(lg[1] - lo[1]) <= TOL


def test_reference_set_up_lid_with_ps_obstacle():
    """LidWithPSObstacle.cpp's run (TRT 0.9 / magic 1/4, PSBB tau from
    allowed_tau(0.02, 10), weights 1 -- compute_obstacle_weight's values
    never reach the list, Lattice.hpp:251-276), 600 steps."""
    cls = lid_with_disc()
    te = 0.9
    to = 0.25 * (1.0 / (te - 0.5)) + 0.5
    tau_ps = 0.5 + 0.02 / ((1.0 / 3.0) * 10)
    kw = dict(collision=L.TRT, tau_even=te, tau_odd=to, inlet_policy=L.ZOUHE_VELOCITY, outlet_policy=L.ZOUHE_VELOCITY)
    say(f"case lid with PS obstacle 100x100: tau+ {te} tau- {to}, PSBB tau {tau_ps}, 600 steps")
    orc, gpus = run_case(cls, kw, tau_ps, None, 600, lambda t, n: (np.full(n, lid_speed(t)), np.zeros(n)))
    check(orc, gpus, cls, "lid PSBB b=1")


@pytest.mark.parametrize("collision", [L.BGK, L.TRT])
def test_per_obstacle_weights(collision):
    """Weights in [0, 1] per wet obstacle (list order), a disc and a block,
    changing inlet speed, 400 steps."""
    cls = L.geometry_channel(60, 140)
    L.add_disc(cls, 30, 40, 8)
    cls[95:101, 20:26] = L.OBSTACLE
    kw = dict(collision=collision, tau=0.75, tau_even=0.7, tau_odd=1.1, inlet_policy=L.ZOUHE_VELOCITY,
              outlet_policy=L.OPEN_NOOP)
    nob = int(O.OracleLattice(cls, **kw).count(O.LIST_OBSTACLE))
    b = np.random.default_rng(4).uniform(0, 1, nob)
    say(f"case channel 60x140 PSBB with {nob} weighted obstacles, collision {collision}")
    orc, gpus = run_case(cls, kw, 0.83, b, 400,
                         lambda t, n: (np.full(n, 0.04 * (1 - math.exp(-t * 0.02))), np.zeros(n)))
    check(orc, gpus, cls, f"weighted PSBB collision {collision}")


def test_rules():
    cls = lid_with_disc()
    kw = dict(collision=L.BGK, tau=0.8, inlet_policy=L.ZOUHE_VELOCITY)
    lat = L.Lattice(cls, storage=L.F32_SHIFTED, **kw)
    with pytest.raises(L.LbmError) as e:
        lat.set_partial_bounce_back(0.9)
    assert e.value.code == L.ERR_UNSUPPORTED
    lat.set_partial_bounce_back(0.0)                  # off: allowed in every storage
    lat = L.Lattice(cls, storage=L.F64, **kw)
    lat.init_equilibrium()
    lat.step(2)
    with pytest.raises(L.LbmError) as e:
        lat.set_partial_bounce_back(0.9)              # after the run started
    assert e.value.code == L.ERR_STATE
    lat2 = L.Lattice(cls, storage=L.F64, **kw)
    lat2.set_partial_bounce_back(0.9)
    lat2.init_equilibrium()
    with pytest.raises(L.LbmError) as e:
        L.Batch([lat2])
    assert e.value.code == L.ERR_UNSUPPORTED
    lat2.step(3)
    before = lat2.populations()
    with pytest.raises(L.LbmError) as e:
        lat2.autotune(steps=2)                        # timing launches would advance the carried values
    assert e.value.code == L.ERR_STATE
    lat2.set_occupancy(4)
    assert np.array_equal(lat2.populations(), before)
    bad = np.ones(int(lat2.node_counts()[6]))
    bad[0] = np.nan
    with pytest.raises(L.LbmError):
        lat2.set_partial_bounce_back(0.9, bad)
