"""The largest lattices the 32-bit plane addressing admits.  The kernels
address a population as a plane base plus a 32-bit byte offset spanning up
to three planes, so a context accepts a lattice only while
3 * plane * element size < 2^32 (plane = compact slots rounded up to 64,
with one spare slot).  Here, just below and just above that bound: the
largest FP64 lattice keeps a closed box at rest exactly (mass and momentum
from the FP64 reductions), in both launch shapes; one column more is
rejected with LBM_ERR_UNSUPPORTED; FP32 storage accepts that size.

A closed box of NR - 2 = 13312 = 208 * 64 fluid rows fills whole 64-slot
groups, so its slot count is 1 + (NC - 2) * 13312."""
import sys

import numpy as np
import pytest

import lbm_amd as L

pytestmark = pytest.mark.gpu

NR = 13314
NC_FIT = 13445      # 3 * plane * 8 = 4,294,878,720 bytes < 2^32
NC_OVER = 13446     # 3 * plane * 8 = 4,295,198,208 bytes > 2^32


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def closed_box(nr, nc):
    cls = np.full((nc, nr), L.FLUID, dtype=np.uint8)
    cls[0, :] = cls[-1, :] = L.BOUNDARY
    cls[:, 0] = cls[:, -1] = L.BOUNDARY
    return cls


def plane_bytes(nr, nc, esize):
    slots = 1 + (nc - 2) * (nr - 2)
    plane = (slots + 1 + 63) // 64 * 64
    return 3 * plane * esize


def test_largest_fp64_lattice_keeps_the_rest_state():
    assert plane_bytes(NR, NC_FIT, 8) < 2 ** 32 <= plane_bytes(NR, NC_OVER, 8)
    cls = closed_box(NR, NC_FIT)
    fluid = (NR - 2) * (NC_FIT - 2)
    say(f"case FP64 {NR}x{NC_FIT} closed box at rest, {fluid} fluid nodes, "
        f"3 planes = {plane_bytes(NR, NC_FIT, 8)} bytes")
    lat = L.Lattice(cls, collision=L.BGK, tau=0.8, storage=L.F64)
    try:
        lat.init_equilibrium()
        m0 = lat.reduce()
        for shape in (1, L.LAUNCH_WET_RUNS):
            lat.set_launch(shape)
            lat.step(3, L.STEP_CHECK)
            red = lat.reduce()
            say(f"  shape {shape}: mass {red[0]:.6f} (start {m0[0]:.6f}), momentum ({red[1]:.3e}, {red[2]:.3e})")
            assert red[4] == fluid
            assert abs(red[0] - m0[0]) <= 1e-12 * m0[0]
            assert abs(red[0] - fluid) <= 1e-9 * fluid
            assert red[1] == 0.0 and red[2] == 0.0
    finally:
        lat.close()


def test_one_column_more_is_rejected_in_fp64_and_accepted_in_fp32():
    cls = closed_box(NR, NC_OVER)
    say(f"case {NR}x{NC_OVER}: FP64 rejected, FP32 accepted")
    with pytest.raises(L.LbmError) as e:
        L.Lattice(cls, storage=L.F64)
    assert e.value.code == 5
    lat = L.Lattice(cls, collision=L.BGK, tau=0.8, storage=L.F32_SHIFTED)
    try:
        lat.init_equilibrium()
        lat.step(3, L.STEP_CHECK)
        red = lat.reduce()
        fluid = (NR - 2) * (NC_OVER - 2)
        say(f"  FP32: mass {red[0]:.6f} for {fluid} fluid nodes")
        assert red[4] == fluid and abs(red[0] - fluid) <= 1e-9 * fluid
        assert red[1] == 0.0 and red[2] == 0.0
    finally:
        lat.close()
