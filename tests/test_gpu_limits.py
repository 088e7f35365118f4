"""The largest lattices the 32-bit population addressing admits.  The
kernels address a population as a direction base plus a 32-bit byte offset
spanning up to three planes, so a context accepts a lattice only while
3 * plane * element size < 2^32 (plane = compact slots rounded up to 64,
with one spare slot); lattices whose whole buffer, 9 * plane * element
size, stays below 2^32 are stored blocked, larger ones in planes.  Here,
just below and just above those bounds: the largest FP64 lattice keeps a
closed box at rest exactly (mass and momentum from the FP64 reductions), in
both launch shapes; one column more is rejected with LBM_ERR_UNSUPPORTED;
FP32 storage accepts that size.  The momentum a lid puts into the largest
lattices of each layout (FP64 and FP32 in planes, the largest blocked FP32
and FP16 ones) after a few steps is the linear function of the width that
two narrow cavities fix: every column under the lid away from the side
walls computes the same values, so a population read from a wrong slot
anywhere along the lid, up to the highest offsets, would show.

A closed box of NR - 2 = 13312 = 208 * 64 fluid rows fills whole 64-slot
groups, so its slot count is 1 + (NC - 2) * 13312; so does a lid cavity of
NR_LID = NR - 1 rows (the lid row and NR_LID - 2 fluid rows per column)."""
import sys

import numpy as np
import pytest

import lbm_amd as L

pytestmark = pytest.mark.gpu

NR = 13314
NC_FIT = 13445      # 3 * plane * 8 = 4,294,878,720 bytes < 2^32
NC_OVER = 13446     # 3 * plane * 8 = 4,295,198,208 bytes > 2^32
NC_F32_BLOCKED = 8964   # 9 * plane * 4 = 4,294,879,488 bytes < 2^32 (8965: 4,295,358,720)
NR_LID = NR - 1


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def closed_box(nr, nc):
    cls = np.full((nc, nr), L.FLUID, dtype=np.uint8)
    cls[0, :] = cls[-1, :] = L.BOUNDARY
    cls[:, 0] = cls[:, -1] = L.BOUNDARY
    return cls


def plane_bytes(nr, nc, esize, planes=3):
    slots = 1 + (nc - 2) * (nr - 2)
    plane = (slots + 1 + 63) // 64 * 64
    return planes * plane * esize


def test_largest_fp64_lattice_keeps_the_rest_state():
    assert plane_bytes(NR, NC_FIT, 8) < 2 ** 32 <= plane_bytes(NR, NC_OVER, 8)
    cls = closed_box(NR, NC_FIT)
    fluid = (NR - 2) * (NC_FIT - 2)
    say(f"case FP64 {NR}x{NC_FIT} closed box at rest, {fluid} fluid nodes, "
        f"3 planes = {plane_bytes(NR, NC_FIT, 8)} bytes")
    lat = L.Lattice(cls, collision=L.BGK, tau=0.8, storage=L.F64)
    try:
        assert lat.layout() == L.LAYOUT_PLANES
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
        assert lat.layout() == L.LAYOUT_PLANES
        lat.init_equilibrium()
        lat.step(3, L.STEP_CHECK)
        red = lat.reduce()
        fluid = (NR - 2) * (NC_OVER - 2)
        say(f"  FP32: mass {red[0]:.6f} for {fluid} fluid nodes")
        assert red[4] == fluid and abs(red[0] - fluid) <= 1e-9 * fluid
        assert red[1] == 0.0 and red[2] == 0.0
    finally:
        lat.close()


def lid_momentum(nr, nc, storage, options, steps, layout=None):
    """x and y momentum of the fluid nodes of an nr x nc lid-driven cavity
    (BGK, tau 0.8, lid u = (0.1, 0)) after `steps` steps; checks the layout
    the lattice got when one is given."""
    cls = L.geometry_cavity(nr, nc)
    lat = L.Lattice(cls, collision=L.BGK, tau=0.8, storage=storage, inlet_policy=L.ZOUHE_VELOCITY,
                    outlet_policy=L.OPEN_NOOP, options=options)
    try:
        if layout is not None:
            assert lat.layout() == {"blocked": L.LAYOUT_BLOCKED, "planes": L.LAYOUT_PLANES}[layout]
        lat.set_open_values(0, ux=0.1, uy=0.0)
        lat.init_equilibrium()
        lat.step(steps, L.STEP_CHECK)
        red = lat.reduce()
        return red[1], red[2]
    finally:
        lat.close()


@pytest.mark.parametrize("storage,options,nc,layout", [
    (L.F64, 0, NC_FIT, "planes"),
    (L.F32_SHIFTED, 0, NC_FIT, "planes"),
    (L.F32_SHIFTED, 0, NC_F32_BLOCKED, "blocked"),
    (L.F16_SHIFTED, L.OPT_F16_ROUND_TO_NEAREST, NC_FIT, "blocked"),
])
def test_lid_momentum_of_the_largest_lattices_is_linear_in_the_width(storage, options, nc, layout):
    esize = {L.F64: 8, L.F32_SHIFTED: 4, L.F16_SHIFTED: 2}[storage]
    assert (plane_bytes(NR, nc, esize, 9) < 2 ** 32) == (layout == "blocked")
    assert plane_bytes(NR, nc, esize) < 2 ** 32
    steps = 4
    say(f"case lid cavity {NR_LID}x{nc}, storage {storage}, options {options} ({layout}), {steps} steps")
    n1, n2 = 64, 96
    m1 = lid_momentum(64, n1, storage, options, steps)
    m2 = lid_momentum(64, n2, storage, options, steps)
    m = lid_momentum(NR_LID, nc, storage, options, steps, layout)
    for k, name in ((0, "x"), (1, "y")):
        expect = m1[k] + (m2[k] - m1[k]) * (nc - n1) / (n2 - n1)
        say(f"  momentum {name}: {m[k]:.15e}, from the narrow cavities {expect:.15e}")
        assert abs(m[k] - expect) <= 1e-10 * abs(m1[0] + (m2[0] - m1[0]) * (nc - n1) / (n2 - n1))
    assert abs(m[0]) > 1e-3 * nc
