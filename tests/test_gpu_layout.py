"""The two population layouts: blocked (the 9 directions of 64 consecutive
wet nodes together, the default wherever a buffer stays below 4 GiB) and
9 planes (LBM_OPT_PLANAR_LAYOUT, and every larger lattice).

Bar: the layout changes addresses only, so after the same steps both give
the same populations, rho/u and FP64 reductions bit for bit, in every
storage format and launch shape.  The FP64 default (blocked) is held to
the CPU oracle by the parity tests, so this carries that bar over to the
planes."""
import sys

import numpy as np
import pytest

import lbm_amd as L
import workloads as W

pytestmark = pytest.mark.gpu

SHAPES = (1, 2, 8, L.LAUNCH_WET_RUNS, L.LAUNCH_RUN_TABLE)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def cases():
    return [W.porous(700, nc=530), W.cavity_trt(300), W.channel(128, 512, radius=10, centre=(64, 100))]


@pytest.mark.parametrize("storage,options", [(L.F64, 0), (L.F32_SHIFTED, 0), (L.F16_SHIFTED, 0),
                                             (L.F16_SHIFTED, L.OPT_F16_ROUND_TO_NEAREST)])
def test_planar_and_blocked_layouts_agree_bit_for_bit(storage, options):
    for wl in cases():
        say(f"case {wl.name} {wl.nr}x{wl.nc}, storage {storage}, options {options}, shapes {SHAPES}")
        blocked = wl.make_lattice(storage, options=options)
        planar = wl.make_lattice(storage, options=options | L.OPT_PLANAR_LAYOUT)
        try:
            assert blocked.layout() == L.LAYOUT_BLOCKED and planar.layout() == L.LAYOUT_PLANES
            for shape in SHAPES:
                for lat in (blocked, planar):
                    lat.init_equilibrium()
                    lat.set_launch(shape)
                    lat.step(1)                     # the collide-only first step
                    lat.step(299, L.STEP_MACRO | L.STEP_CHECK)
                pb, pp = blocked.populations(), planar.populations()
                assert np.isfinite(pb).all()
                assert np.array_equal(pb, pp), f"shape {shape}: populations differ"
                rb, ub = blocked.macro()
                rp, up = planar.macro()
                assert np.array_equal(rb, rp) and np.array_equal(ub, up), f"shape {shape}: rho/u differ"
                assert np.array_equal(blocked.reduce(), planar.reduce()), f"shape {shape}: reductions differ"
            say("  identical in every shape")
        finally:
            blocked.close()
            planar.close()


def test_layouts_round_trip_the_same_state():
    wl = W.porous(300, nc=200)
    rng = np.random.default_rng(4)
    say(f"case upload/download round trip {wl.nr}x{wl.nc}, FP64, both layouts")
    w = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
    P = (1.0 + 0.01 * rng.standard_normal((9, wl.nc, wl.nr))) * w[:, None, None]
    out = []
    for opt in (0, L.OPT_PLANAR_LAYOUT):
        lat = wl.make_lattice(L.F64, options=opt)
        try:
            lat.upload_populations(P)
            out.append(lat.populations())
            lat.step(5, L.STEP_CHECK)
            out.append(lat.populations())
        finally:
            lat.close()
    wet = (wl.node_class == L.FLUID) | (wl.node_class == L.INLET) | (wl.node_class == L.OUTLET)
    assert np.array_equal(out[0][:, wet], P[:, wet])
    assert np.array_equal(out[0], out[2]) and np.array_equal(out[1], out[3])
