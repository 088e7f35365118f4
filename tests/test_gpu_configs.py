"""The lid-driven TRT cavity (the 4096 x 4096, Re 1000 configuration and
scaled-down copies), the full-size channel over 1000 steps, corner-typed
open nodes, the FP16 range guard, asynchronous downloads and argument
checks of the C ABI.

Bars: FP64 bit-exact against the oracle (checked at <= 1e-12); the shifted
storage formats against the FP64 GPU run within the L2 relative bounds
written in each test; exact rest states and mass where the arithmetic
makes them exact."""
import os
import sys

import numpy as np
import pytest

import lbm_amd as L
import oracle as O
import workloads as W
from helpers import max_abs_wet, oracle_from, wet_mask

pytestmark = pytest.mark.gpu
TOL = 1e-12
THREADS = min(16, os.cpu_count() or 1)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def l2rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# ----------------------------------------------------------- TRT cavity

def test_cavity_trt_fp64_matches_oracle():
    """TRT lid cavity, magic 3/16, u_lid 0.1, Re 1000 scaled to 256 x 256
    (tau+ 0.576), 1000 steps: populations, rho/u and the FP64 reductions
    against the oracle."""
    wl = W.cavity_trt(256)
    say(f"case {wl.name}: TRT tau+ {wl.tau_even:.4f} tau- {wl.tau_odd:.4f}, 1000 steps")
    O.set_threads(THREADS)
    orc = oracle_from(wl)
    gpu = wl.make_lattice(L.F64)
    gpu.step(1000, L.STEP_MACRO | L.STEP_CHECK)
    orc.step(1000)
    cls = wl.node_class
    err = max_abs_wet(gpu.populations(), orc.populations(), cls)
    say(f"  max|dP| = {err:.3e}")
    assert err <= TOL
    rg, ug = gpu.macro()
    ro, uo = orc.macro()
    m = wet_mask(cls)
    assert np.max(np.abs(rg[m] - ro[m])) <= TOL and np.max(np.abs(ug[:, m] - uo[:, m])) <= TOL
    assert np.allclose(gpu.reduce(), orc.reduce(), rtol=1e-13, atol=1e-15)


# stated bounds: L2 relative error of the fluid velocity and of rho - 1
# against the FP64 GPU run (256 x 256 TRT cavity, Re 1000, 4000 steps)
SHIFTED_BOUNDS = [
    ("fp32", L.F32_SHIFTED, 0, 1e-5, 1e-5),
    ("fp16-sr", L.F16_SHIFTED, 0, 2e-3, 2e-2),
    ("fp16-rn", L.F16_SHIFTED, L.OPT_F16_ROUND_TO_NEAREST, 5e-2, 2e-1),
]


@pytest.mark.parametrize("name,storage,opts,bu,br", SHIFTED_BOUNDS, ids=[b[0] for b in SHIFTED_BOUNDS])
def test_cavity_trt_shifted_storage_errors(name, storage, opts, bu, br):
    wl = W.cavity_trt(256)
    steps = 4000
    ref = wl.make_lattice(L.F64)
    lo = wl.make_lattice(storage, options=opts)
    ref.step(steps, L.STEP_MACRO)
    lo.step(steps, L.STEP_MACRO | L.STEP_CHECK)
    fl = wl.node_class == L.FLUID
    r0, u0 = ref.macro()
    r1, u1 = lo.macro()
    eu = l2rel(u1[:, fl], u0[:, fl])
    er = l2rel(r1[fl] - 1.0, r0[fl] - 1.0)
    dm = abs(lo.reduce()[0] - ref.reduce()[0]) / ref.reduce()[0]
    say(f"  {name}: velocity {eu:.3e} (bound {bu}), density {er:.3e} (bound {br}), mass {dm:.2e}, scale {lo.scale}")
    assert eu <= bu and er <= br and dm <= 1e-5


@pytest.mark.slow
def test_cavity_trt_4096_full_size():
    """The 4096 x 4096 configuration itself in all three storage formats.
    (a) Lid at rest: shifted storage holds exact zeros, so after 100 steps
    its momentum is exactly zero and its mass bit for bit the mass after
    the initialisation; FP64 (whose nine weights sum to 1 + 2^-52, not 1,
    so the rest state moves by roundings) keeps the mass to 1e-12 and the
    momentum below 1e-9.  (b) Lid at 0.1: after 300
    steps the FP32 / FP16 fields are within the stated L2 bounds of the
    FP64 run (velocity: FP32 1e-5, FP16 stochastic 5e-3) and the FP64
    reduced mass agrees to 1e-6 relative."""
    wl = W.cavity_trt(4096)
    fl = wl.node_class == L.FLUID
    nw = wl.wet_count()
    say(f"case {wl.name}: {nw} wet nodes")
    for storage in (L.F64, L.F32_SHIFTED, L.F16_SHIFTED):
        lat = L.Lattice(wl.node_class, storage=storage, **wl.lattice_kwargs())
        ni = len(lat.open_nodes(0))
        lat.set_open_values(0, ux=np.zeros(ni), uy=np.zeros(ni))
        lat.init_equilibrium()
        m0 = lat.reduce()[0]
        lat.step(100, L.STEP_CHECK)
        red = lat.reduce()
        say(f"  storage {storage} at rest: mass {red[0]!r} (after init {m0!r}, wet {nw}), "
            f"momentum {red[1]!r}, {red[2]!r}")
        assert abs(m0 - nw) <= 1e-12 * nw
        if storage == L.F64:
            assert abs(red[0] - m0) <= 1e-12 * nw
            assert abs(red[1]) <= 1e-9 and abs(red[2]) <= 1e-9
        else:
            assert red[0] == m0
            assert red[1] == 0.0 and red[2] == 0.0
        lat.close()
    runs = {}
    for storage in (L.F64, L.F32_SHIFTED, L.F16_SHIFTED):
        lat = wl.make_lattice(storage)
        lat.step(300, L.STEP_MACRO | L.STEP_CHECK)
        rho, u = lat.macro()
        runs[storage] = (u[:, fl].copy(), lat.reduce()[0])
        lat.close()
    u64, m64 = runs[L.F64]
    for storage, bound in ((L.F32_SHIFTED, 1e-5), (L.F16_SHIFTED, 5e-3)):
        u, m = runs[storage]
        e = l2rel(u, u64)
        say(f"  storage {storage} lid 0.1, 300 steps: velocity {e:.3e} (bound {bound}), mass {abs(m - m64) / m64:.2e}")
        assert e <= bound and abs(m - m64) <= 1e-6 * m64


# ------------------------------------------------------ full-size channel

@pytest.mark.slow
@pytest.mark.timeout(900)
def test_channel_full_size_1000_steps():
    """The 512 x 2048 channel (Re 200, cylinder r 26, parabolic Zou-He
    inlet, Zou-He pressure outlet) against the oracle over 1000 steps, with
    the per-point and total lift/drag compared every 100 steps."""
    wl = W.channel()
    say(f"case {wl.name}: 1000 steps, lift/drag every 100")
    O.set_threads(THREADS)
    orc = oracle_from(wl)
    gpu = wl.make_lattice(L.F64)
    for k in range(10):
        gpu.step(100 if k else 101, L.STEP_MACRO if k == 9 else 0)
        orc.step(100 if k else 101)
        if k < 9:
            lg = gpu.lift_drag(per_point=True)
            lo = orc.lift_drag(per_point=True)
            assert np.max(np.abs(lg[2] - lo[2])) <= TOL
            assert abs(lg[0] - lo[0]) <= TOL and abs(lg[1] - lo[1]) <= TOL
            say(f"  step {gpu.time - 1}: lift {lo[0]:.6e} drag {lo[1]:.6e}")
    gpu.step(0)
    assert gpu.time == orc.time == 1001
    err = max_abs_wet(gpu.populations(), orc.populations(), wl.node_class)
    say(f"  max|dP| after {gpu.time} steps = {err:.3e}")
    assert err <= TOL
    rg, ug = gpu.macro()
    ro, uo = orc.macro()
    m = wet_mask(wl.node_class)
    assert np.max(np.abs(rg[m] - ro[m])) <= TOL and np.max(np.abs(ug[:, m] - uo[:, m])) <= TOL


# ------------------------------------------------ corner-typed open nodes

@pytest.mark.parametrize("inlet_policy,outlet_policy", [(L.ZOUHE_VELOCITY, L.ZOUHE_PRESSURE),
                                                        (L.ZOUHE_VELOCITY, L.ZOUHE_VELOCITY),
                                                        (L.ZOUHE_PRESSURE, L.OPEN_NOOP)])
@pytest.mark.parametrize("collision", [L.BGK, L.TRT])
def test_corner_typed_open_nodes(collision, inlet_policy, outlet_policy):
    """Inlets and outlets that include the domain corners (ZouHePolicy.hpp:
    119-120 leaves corner-typed nodes alone; their out-of-domain slots stay
    frozen): inlet = column 0 down to row NR-2, with the top-left corner,
    outlet = the whole bottom row with both bottom corners, walls
    elsewhere, an obstacle.
    Both launch shapes against the oracle, values changed half way."""
    nr, nc = 70, 90
    cls = np.full((nc, nr), L.FLUID, dtype=np.uint8)
    cls[0, :] = cls[-1, :] = L.BOUNDARY
    cls[:, 0] = cls[:, -1] = L.BOUNDARY
    cls[0, :nr - 1] = L.INLET              # column 0, rows 0..NR-2: the top-left corner (0, 0)
    cls[:, nr - 1] = L.OUTLET              # row NR-1, columns 0..NC-1: both bottom corners
    L.add_disc(cls, 30, 40, 7)
    rng = np.random.default_rng(17)
    te, to = W.trt_taus(0.06)

    def vals(which, pol):
        n = int(np.count_nonzero(cls == (L.INLET if which == 0 else L.OUTLET)))
        if pol == L.ZOUHE_PRESSURE:
            return dict(rho=1.0 + 1e-3 * rng.uniform(-1, 1, n))
        if pol == L.ZOUHE_VELOCITY:
            return dict(ux=0.03 * rng.uniform(0, 1, n), uy=0.01 * rng.uniform(-1, 1, n))
        return {}

    wl = W.Workload(name="corners", node_class=cls, collision=collision, tau=0.68, tau_even=te, tau_odd=to,
                    inlet_policy=inlet_policy, outlet_policy=outlet_policy, inlet=vals(0, inlet_policy),
                    outlet=vals(1, outlet_policy))
    say(f"case corners {nr}x{nc}: collision {collision}, inlet {inlet_policy}, outlet {outlet_policy}")
    gpus = [wl.make_lattice(L.F64), wl.make_lattice(L.F64)]
    gpus[1].set_launch(L.LAUNCH_WET_RUNS)
    orc = oracle_from(wl, threads=THREADS)
    # btypes (aliases.hpp:28-38): 5..8 are the corner types
    assert np.count_nonzero(orc.node_list(O.LIST_INLET)[1] >= 5) == 1
    assert np.count_nonzero(orc.node_list(O.LIST_OUTLET)[1] >= 5) == 2
    for g in gpus:
        g.step(250)
    orc.step(250)
    for which, pol in ((0, inlet_policy), (1, outlet_policy)):
        v = vals(which, pol)
        if v:
            for g in gpus:
                g.set_open_values(which, **v)
            orc.set_open_values(which, **v)
    for g in gpus:
        g.step(250, L.STEP_MACRO)
    orc.step(250)
    po = orc.populations()
    assert np.all(np.isfinite(po[:, wet_mask(cls)]))
    for shape, g in zip(("row segments", "wet runs"), gpus):
        err = max_abs_wet(g.populations(), po, cls)
        say(f"  {shape}: max|dP| = {err:.3e}")
        assert err <= TOL
        lg, lo = g.lift_drag(), orc.lift_drag()
        assert abs(lg[0] - lo[0]) <= TOL and abs(lg[1] - lo[1]) <= TOL


# ----------------------------------------------------------- FP16 range

def test_fp16_range_guard():
    """A flow faster than the FP16 storage scale allows is reported: a
    cavity whose scale is set to 2^21 (room for |f - w| < 0.031) with a
    lid at 0.15 (|f - w| about 0.06 at the lid) returns LBM_ERR_RANGE under
    LBM_STEP_CHECK, in both rounding modes; the default scale for the same
    flow does not."""
    wl = W.cavity(65, u_lid=0.15)
    for opts in (0, L.OPT_F16_ROUND_TO_NEAREST):
        lat = wl.make_lattice(L.F16_SHIFTED, f16_scale=float(2 ** 21), options=opts)
        assert lat.scale == 2.0 ** 21
        with pytest.raises(L.LbmError) as e:
            lat.step(50, L.STEP_CHECK)
        assert e.value.code == L.ERR_RANGE, e.value
        lat.close()
    lat = wl.make_lattice(L.F16_SHIFTED)
    say(f"  default scale for lid 0.15: {lat.scale}")
    lat.step(500, L.STEP_CHECK)
    lat.close()


def test_fp16_scale_follows_values_set_after_init():
    """Values set after lbm_init_equilibrium and before the first step still
    choose the FP16 scale (the state is exactly zero then); values set
    later do not change it."""
    wl = W.cavity(33, u_lid=0.3)
    lat = L.Lattice(wl.node_class, storage=L.F16_SHIFTED, **wl.lattice_kwargs())
    lat.init_equilibrium()
    s_rest = lat.scale
    lat.set_open_values(0, **wl.inlet)
    s_lid = lat.scale
    assert s_rest == 32768.0 and s_lid == 2048.0, (s_rest, s_lid)
    ref = wl.make_lattice(L.F16_SHIFTED)      # values set before init
    assert ref.scale == s_lid
    lat.step(20)
    ref.step(20)
    assert np.array_equal(lat.populations(), ref.populations())
    lat.set_open_values(0, ux=0.0, uy=0.0)
    assert lat.scale == s_lid


def test_f16_scale_argument_is_validated():
    cls = L.geometry_cavity(16, 16)
    for bad in (3.0, 0.5, float("inf"), float("nan"), -4.0, 2.0 ** 25):
        with pytest.raises(L.LbmError) as e:
            L.Lattice(cls, storage=L.F16_SHIFTED, f16_scale=bad)
        assert e.value.code == L.ERR_INVALID
    assert L.Lattice(cls, storage=L.F16_SHIFTED, f16_scale=1024.0).scale == 1.0   # applied at init


# ------------------------------------------------- argument checks, async

def test_failed_open_values_change_nothing():
    """lbm_set_open_values checks every value before taking any: a rejected
    call leaves the prescribed values and the results unchanged."""
    wl = W.channel(40, 96, radius=5, centre=(20, 30))
    a = wl.make_lattice(L.F64)
    b = wl.make_lattice(L.F64)
    n = len(a.open_nodes(1))
    bad = np.ones(n)
    bad[n // 2] = -1.0
    with pytest.raises(L.LbmError) as e:
        a.set_open_values(1, rho=bad * 1.01)
    assert e.value.code == L.ERR_INVALID
    bad_u = np.zeros(len(a.open_nodes(0)))
    bad_u[3] = np.nan
    with pytest.raises(L.LbmError):
        a.set_open_values(0, ux=bad_u)
    a.step(60)
    b.step(60)
    assert np.array_equal(a.populations(), b.populations())


def test_lift_drag_rejects_masks_naming_dry_nodes():
    """A direction bit must name a wet neighbour (NodeIdentification.hpp:
    137-211); a bit towards a wall or obstacle node is LBM_ERR_INVALID."""
    wl = W.channel(40, 96, radius=5, centre=(20, 30))
    lat = wl.make_lattice(L.F64)
    lat.step(5)
    orc_like = O.OracleLattice(wl.node_class, **wl.lattice_kwargs())
    ij, masks = orc_like.node_list(O.LIST_OBSTACLE)
    lat.lift_drag(ij, masks)                         # the reference's own masks pass
    q = 0
    i, j = ij[q]
    # direction d whose neighbour is another obstacle node (inside the disc)
    for d in range(1, 9):
        ii, jj = i + [0, 0, -1, 0, 1, -1, -1, 1, 1][d], j + [0, 1, 0, -1, 0, 1, -1, -1, 1][d]
        if wl.node_class[jj, ii] == L.OBSTACLE:
            break
    else:
        pytest.fail("no dry neighbour found")
    bad = masks.copy()
    bad[q] |= 1 << d
    with pytest.raises(L.LbmError) as e:
        lat.lift_drag(ij, bad)
    assert e.value.code == L.ERR_INVALID


def test_async_macro_downloads_match_sync():
    """lbm_download_macro_async + lbm_wait deliver the fields of the
    LBM_STEP_MACRO step they follow, while later MACRO steps write the
    other device copy; three downloads in flight, checked against
    synchronous downloads of a twin lattice."""
    wl = W.channel(64, 200, radius=6, centre=(32, 50), re=20)
    a = wl.make_lattice(L.F64)
    b = wl.make_lattice(L.F64)
    bufs, tickets, want = [], [], []
    for k in range(3):
        a.step(37, L.STEP_MACRO)
        b.step(37, L.STEP_MACRO)
        rho = np.full((a.nc, a.nr), -1.0)
        u = np.full((2, a.nc, a.nr), -1.0)
        tickets.append(a.macro_async(rho, u))
        bufs.append((rho, u))
        want.append(b.macro())
    wall = wl.node_class == L.BOUNDARY
    for t, (rho, u), (rw, uw) in zip(tickets, bufs, want):
        a.wait(t)
        assert np.array_equal(rho, rw) and np.array_equal(u, uw)
        # nodes the steps never write hold rho = 1, u = 0 (Lattice.hpp:295-296)
        assert np.all(rho[wall] == 1.0) and np.all(u[:, wall] == 0.0)
    with pytest.raises(L.LbmError):
        a.wait(tickets[-1] + 5)
