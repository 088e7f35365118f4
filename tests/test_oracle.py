"""Pins the CPU oracle.  The reference ships no golden vectors or
known-answer tests (its only assertion is test/unit/InletUpdates.cpp:44), so
each oracle function is checked against what the mathematics fixes: moment
identities, conservation laws, closed forms, brute force on tiny inputs,
the reference's own geometry fixtures, and two textbook flows (Poiseuille,
Ghia et al. 1982).  Each check is chosen so that a plausible slip -- a wrong
weight, sign, direction index or opposite -- fails it."""
import numpy as np
import pytest

import oracle as O
from helpers import CX, CY, DCOL, DROW, OPP, W, oracle_from, wet_mask

RNG = np.random.default_rng(1234)


# ---------------------------------------------------------------- equilibrium

@pytest.mark.parametrize("rho,ux,uy", [(1.0, 0.0, 0.0), (1.02, 0.05, -0.03), (0.97, -0.1, 0.08), (1.3, 0.2, 0.2)])
def test_feq_moments(rho, ux, uy):
    """Sum f = rho, sum c f = rho u, sum c c f = rho (u u + I/3): D2Q9 second-order equilibrium."""
    e = O.feq(rho, ux, uy)
    assert np.isclose(e.sum(), rho, rtol=0, atol=1e-15)
    assert np.isclose((CX * e).sum(), rho * ux, atol=1e-15)
    assert np.isclose((CY * e).sum(), rho * uy, atol=1e-15)
    assert np.isclose((CX * CX * e).sum(), rho * (ux * ux + 1 / 3), atol=1e-15)
    assert np.isclose((CY * CY * e).sum(), rho * (uy * uy + 1 / 3), atol=1e-15)
    assert np.isclose((CX * CY * e).sum(), rho * ux * uy, atol=1e-15)


def test_feq_textbook_form():
    """f_d = w_d rho (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u^2)."""
    for _ in range(20):
        rho, ux, uy = 1 + 0.1 * RNG.standard_normal(), 0.1 * RNG.standard_normal(), 0.1 * RNG.standard_normal()
        cu = CX * ux + CY * uy
        ref = W * rho * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (ux * ux + uy * uy))
        assert np.allclose(O.feq(rho, ux, uy), ref, rtol=0, atol=2e-16)


def test_moments_inverse_of_feq():
    for _ in range(20):
        rho, ux, uy = 1 + 0.1 * RNG.standard_normal(), 0.1 * RNG.standard_normal(), 0.1 * RNG.standard_normal()
        r, a, b = O.moments(O.feq(rho, ux, uy))
        assert np.isclose(r, rho, atol=1e-15) and np.isclose(a, ux, atol=1e-15) and np.isclose(b, uy, atol=1e-15)


# ------------------------------------------------------------------ collision

def _rand_state():
    rho, ux, uy = 1 + 0.05 * RNG.standard_normal(), 0.05 * RNG.standard_normal(), 0.05 * RNG.standard_normal()
    p = O.feq(rho, ux, uy) * (1 + 0.05 * RNG.standard_normal(9))
    r, a, b = O.moments(p)
    return p, O.feq(r, a, b)


@pytest.mark.parametrize("tau", [0.51, 0.8, 1.0, 1.7])
def test_bgk_conserves_and_relaxes(tau):
    p, e = _rand_state()
    o = O.collide_bgk(tau, p, e)
    assert np.isclose(o.sum(), p.sum(), atol=1e-15)
    assert np.isclose((CX * o).sum(), (CX * p).sum(), atol=1e-15)
    assert np.isclose((CY * o).sum(), (CY * p).sum(), atol=1e-15)
    # non-equilibrium part scaled by exactly (1 - 1/tau)
    assert np.allclose(o - e, (1 - 1 / tau) * (p - e), atol=1e-16)


@pytest.mark.parametrize("te,to", [(0.8, 0.8), (0.52, 9.5), (1.7, 0.6)])
def test_trt_closed_form(te, to):
    """Symmetric part relaxes with 1/tau+, antisymmetric part with 1/tau-."""
    p, e = _rand_state()
    o = O.collide_trt(te, to, p, e)
    sym = lambda f: 0.5 * (f + f[OPP])
    asym = lambda f: 0.5 * (f - f[OPP])
    assert np.allclose(sym(o) - sym(e), (1 - 1 / te) * (sym(p) - sym(e)), atol=1e-16)
    assert np.allclose(asym(o) - asym(e), (1 - 1 / to) * (asym(p) - asym(e)), atol=1e-16)
    assert np.isclose(o.sum(), p.sum(), atol=1e-15)
    assert np.isclose((CX * o).sum(), (CX * p).sum(), atol=1e-15)


def test_trt_equal_rates_is_bgk():
    for tau in (0.6, 1.0, 1.4):
        p, e = _rand_state()
        assert np.allclose(O.collide_trt(tau, tau, p, e), O.collide_bgk(tau, p, e), atol=1e-16)


# ----------------------------------------------------------------- Zou-He

SIDES = {O.BT_TOP: (0, -1), O.BT_BOTTOM: (0, 1), O.BT_LEFT: (1, 0), O.BT_RIGHT: (-1, 0)}  # inward normal (x, y up)


def _unknowns(side):
    nx, ny = SIDES[side]
    return [d for d in range(1, 9) if CX[d] * nx + CY[d] * ny > 0]


@pytest.mark.parametrize("side", list(SIDES))
def test_zouhe_velocity_moments(side):
    """After the closure the node carries exactly rho and rho*u with the
    rho of the Zou-He mass balance, and only the three unknowns change."""
    ux, uy = 0.06, -0.02
    p = O.feq(1.01, 0.03, 0.01) * (1 + 0.02 * RNG.standard_normal(9))
    q, rho = O.zouhe_velocity(side, p, ux, uy)
    changed = [d for d in range(9) if q[d] != p[d]]
    assert set(changed) <= set(_unknowns(side)) and len(_unknowns(side)) == 3
    assert np.isclose(q.sum(), rho, atol=1e-15)
    assert np.isclose((CX * q).sum(), rho * ux, atol=1e-15)
    assert np.isclose((CY * q).sum(), rho * uy, atol=1e-15)


def test_zouhe_corner_is_noop():
    p = O.feq(1.0, 0.01, 0.0)
    for bt in (5, 6, 7, 8, O.BT_NONE):
        q, _ = O.zouhe_velocity(bt, p, 0.1, 0.0)
        assert np.array_equal(q, p)


@pytest.mark.parametrize("side", list(SIDES))
def test_zouhe_pressure_moments(side):
    rho0 = 1.003
    p = O.feq(1.0, 0.02, -0.01) * (1 + 0.02 * RNG.standard_normal(9))
    q, un = O.zouhe_pressure(side, p, rho0)
    assert np.isclose(q.sum(), rho0, atol=1e-15)
    nx, ny = SIDES[side]
    # tangential momentum zero, normal momentum rho0 * un
    assert np.isclose((CX * q).sum(), rho0 * un * abs(nx), atol=1e-15)
    assert np.isclose((CY * q).sum(), rho0 * un * abs(ny), atol=1e-15)
    assert set(d for d in range(9) if q[d] != p[d]) <= set(_unknowns(side))


# ---------------------------------------------------------- lattice structure

def _box(nr, nc):
    cls = np.full((nc, nr), O.FLUID, dtype=np.uint8)
    cls[0, :] = cls[-1, :] = O.BOUNDARY
    cls[:, 0] = cls[:, -1] = O.BOUNDARY
    return cls


def test_node_lists_and_types():
    """Row-major list order, btypes by position, masks and wet-obstacle
    elimination (NodeIdentification.hpp:45-242) against brute force."""
    nr, nc = 9, 12
    cls = _box(nr, nc)
    cls[0, 1:-1] = O.INLET         # left column (j = 0), rows 1..nr-2
    cls[-1, 1:-1] = O.OUTLET       # right column
    cls[0, 0] = O.INLET            # (row 0, col 0): a corner
    cls[4:8, 3:6] = O.OBSTACLE     # columns 4..7, rows 3..5 -> has an inner (dry) node row 4
    lat = O.OracleLattice(cls)
    ij, tag = lat.node_list(O.LIST_INLET)
    assert [tuple(x) for x in ij] == sorted(tuple(x) for x in ij)
    assert (0, 0) in [tuple(x) for x in ij]
    for (i, j), t in zip(ij, tag):
        if (i, j) == (0, 0):
            assert t == 8  # LEFT_TOP_C
        elif j == 0:
            assert t == O.BT_LEFT
    oij, mask = lat.node_list(O.LIST_OBSTACLE)
    wet = wet_mask(cls)
    expect = {}
    for j in range(nc):
        for i in range(nr):
            if cls[j, i] != O.OBSTACLE:
                continue
            m = 0
            for d in range(1, 9):
                ii, jj = i + DROW[d], j + DCOL[d]
                if 0 <= ii < nr and 0 <= jj < nc and wet[jj, ii]:
                    m |= 1 << d
            if m:
                expect[(i, j)] = m
    got = {tuple(x): int(t) for x, t in zip(oij, mask)}
    assert got == expect
    assert (4, 5) not in got  # interior obstacle node is not wet
    assert lat.count(O.LIST_FLUID) == int(np.count_nonzero(cls == O.FLUID))


# ------------------------------------------------------------ streaming / BB

def _pure_stream_lattice(cls):
    # BGK with tau = 1e300: 1/tau = 1e-300 and 1 - 1/tau == 1.0 exactly, so
    # the collision returns the populations unchanged and a step is pure
    # streaming plus bounce-back.
    return O.OracleLattice(cls, collision=O.BGK, tau=1e300, inlet_policy=O.NOOP, outlet_policy=O.NOOP)


def test_streaming_moves_along_velocity_and_bounces_back():
    """A population of direction d moves by (-c_y, +c_x) in (row, column) per
    step; into a wall or obstacle it comes back into the opposite slot of
    the same node; slots fed from outside keep their value."""
    nr, nc = 10, 11
    cls = _box(nr, nc)
    cls[5, 4] = O.OBSTACLE   # column 5, row 4
    lat = _pure_stream_lattice(cls)
    p0 = RNG.uniform(1.0, 2.0, size=(9, nc, nr))
    lat.set_populations(p0)
    lat.step(1)
    p1 = lat.populations()
    wet = wet_mask(cls)
    for j in range(nc):
        for i in range(nr):
            if not wet[j, i]:
                continue
            for d in range(9):
                si, sj = i - DROW[d], j - DCOL[d]
                if wet[sj, si]:
                    want = p0[d, sj, si]
                else:
                    want = p0[OPP[d], j, i]
                assert p1[d, j, i] == want, (i, j, d)


def test_closed_box_streaming_is_a_permutation():
    nr, nc = 16, 13
    cls = _box(nr, nc)
    cls[6:9, 5:8] = O.OBSTACLE
    lat = _pure_stream_lattice(cls)
    p0 = RNG.uniform(0.0, 1.0, size=(9, nc, nr))
    lat.set_populations(p0)
    wet = wet_mask(cls)
    lat.step(7)
    p7 = lat.populations()
    assert np.array_equal(np.sort(p0[:, wet].ravel()), np.sort(p7[:, wet].ravel()))


def test_closed_box_conserves_mass_and_rest_state():
    nr, nc = 20, 24
    cls = _box(nr, nc)
    cls[8:12, 6:10] = O.OBSTACLE
    for coll in (O.BGK, O.TRT):
        lat = O.OracleLattice(cls, collision=coll, tau=0.7, tau_even=0.7, tau_odd=1.1)
        lat.init()
        lat.step(50)
        rho, u = lat.macro()
        fl = cls == O.FLUID
        assert np.max(np.abs(u[:, fl])) < 1e-15       # a resting fluid stays at rest
        lat2 = O.OracleLattice(cls, collision=coll, tau=0.7, tau_even=0.7, tau_odd=1.1)
        lat2.init()
        p = lat2.populations()
        wet = wet_mask(cls)
        p[:, wet] *= 1 + 0.01 * RNG.standard_normal(p[:, wet].shape)
        lat2.set_populations(p)
        m0 = lat2.reduce()[0]
        lat2.step(200)
        assert abs(lat2.reduce()[0] - m0) < 1e-12 * m0


def test_noop_outlet_slots_stay_frozen():
    """Q2: with the no-op outlet, slots whose source is outside the domain are
    never written and keep their initial equilibrium value w_d."""
    nr, nc = 12, 30
    cls = _box(nr, nc)
    cls[-1, 1:-1] = O.OUTLET       # right column
    cls[0, 1:-1] = O.INLET         # left column
    lat = O.OracleLattice(cls, collision=O.TRT, tau_even=0.6, tau_odd=1.67, inlet_policy=O.ZOUHE_VELOCITY,
                          outlet_policy=O.NOOP)
    lat.set_open_values(0, ux=0.05, uy=0.0)
    lat.init()
    w0 = lat.populations()[:, -1, 5].copy()
    lat.step(100)
    p = lat.populations()
    for d in (3, 6, 7):            # sources at column NC (outside)
        assert np.all(p[d, -1, 1:-1] == w0[d])


def test_inlet_update_values_reach_the_lattice():
    """The reference's one assertion (test/unit/InletUpdates.cpp:44): after
    update_nodes the inlet velocity equals the callback's value."""
    cls = _box(10, 10)
    cls[0, 1:-1] = O.INLET         # left column
    lat = O.OracleLattice(cls, collision=O.BGK, tau=0.8)
    lat.set_open_values(0, ux=0.0375, uy=-0.01)
    lat.init()
    lat.step(1)
    _, u = lat.macro()
    assert np.all(u[0, 0, 1:-1] == 0.0375) and np.all(u[1, 0, 1:-1] == -0.01)


# ---------------------------------------------------------------- lift / drag

def test_lift_drag_is_momentum_of_incoming_populations():
    """FlowAnalyzer.hpp:386-430 = sum over the obstacle's wet directions d of
    the post-collision population opp(d) of the neighbour at +c_d, times
    the unit vector of opp(d); checked against that definition on random
    post-collision fields of a small obstacle."""
    nr, nc = 12, 14
    cls = _box(nr, nc)
    cls[5:8, 4:7] = O.OBSTACLE
    cls[6, 8] = O.OBSTACLE
    lat = O.OracleLattice(cls, collision=O.BGK, tau=0.9, inlet_policy=O.NOOP, outlet_policy=O.NOOP)
    lat.set_populations(RNG.uniform(0.01, 0.2, size=(9, nc, nr)))
    lat.step(1)
    A = lat.after_collision()
    lift, drag, pp = lat.lift_drag(per_point=True)
    ij, masks = lat.node_list(O.LIST_OBSTACLE)
    tot = np.zeros(2)
    for k, ((i, j), m) in enumerate(zip(ij, masks)):
        f = np.zeros(2)
        for d in range(1, 9):
            if m & (1 << d):
                o = OPP[d]
                v = A[o, j + DCOL[d], i + DROW[d]]
                f += v * np.array([CX[o], CY[o]]) / np.hypot(CX[o], CY[o])
        assert np.allclose(pp[k], [f[1], f[0]], atol=1e-15)
        tot += f
    assert np.isclose(lift, tot[1], atol=1e-14) and np.isclose(drag, tot[0], atol=1e-14)


# ------------------------------------------------------------- flows (slow)

@pytest.mark.slow
def test_poiseuille_profile():
    """Pressure-driven channel, TRT with magic 3/16: the steady profile is the
    parabola with zeros on the half-way walls (rows 0.5 and NR-1.5)."""
    nr, nc = 22, 40
    cls = _box(nr, nc)
    cls[0, 1:-1] = O.INLET
    cls[-1, 1:-1] = O.OUTLET
    nu = 0.1
    te = 3 * nu + 0.5
    to = (3 / 16) / (te - 0.5) + 0.5
    lat = O.OracleLattice(cls, collision=O.TRT, tau_even=te, tau_odd=to, inlet_policy=O.ZOUHE_PRESSURE,
                          outlet_policy=O.ZOUHE_PRESSURE)
    lat.set_open_values(0, rho=1.0005)
    lat.set_open_values(1, rho=0.9995)
    lat.init()
    lat.step(6000)
    _, u = lat.macro()
    ux = u[0, nc // 2, 1:-1]
    y = np.arange(1, nr - 1) - 0.5
    h = nr - 2
    shape = y * (h - y)
    a = float(np.dot(ux, shape) / np.dot(shape, shape))
    assert a > 0
    assert np.max(np.abs(ux - a * shape)) < 2e-3 * np.max(ux)
    assert np.max(np.abs(u[1, nc // 2, 1:-1])) < 1e-3 * np.max(ux)


# ------------------------------------------- velocity-prescribed open nodes

def _feq_textbook(rho, ux, uy):
    """w_d rho (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u.u), per node (arrays over nodes)."""
    ux = np.asarray(ux, dtype=np.float64)[None]
    uy = np.asarray(uy, dtype=np.float64)[None]
    cu = CX[:, None] * ux + CY[:, None] * uy
    return W[:, None] * rho * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (ux * ux + uy * uy))


def _open_box(nr, nc):
    """Velocity inlets on the top row and the left column (corner (0, 0)
    included: a corner-typed inlet node), Zou-He pressure outlet on the
    right column, walls elsewhere, one obstacle."""
    cls = _box(nr, nc)
    cls[1:-1, 0] = O.INLET     # top row (i = 0), columns 1..nc-2
    cls[0, 1:-1] = O.INLET     # left column (j = 0), rows 1..nr-2
    cls[0, 0] = O.INLET        # corner
    cls[-1, 1:-1] = O.OUTLET   # right column
    cls[6:9, 5:8] = O.OBSTACLE
    return cls


def _open_values(rng, n):
    return 0.02 + 0.03 * rng.uniform(size=n), 0.01 * rng.standard_normal(n)


@pytest.mark.parametrize("coll", [O.BGK, O.TRT])
def test_velocity_inlet_equilibrium_uses_unit_density(coll):
    """Q1 (Lattice.hpp:295,357; ZouHePolicy.hpp:58,71): a velocity-prescribed
    node keeps global_rho = 1 -- update_macro touches the fluid list only and
    the Zou-He density goes to a by-value copy -- so its collision relaxes
    towards feq(1, U) with the prescribed U of the step, not towards the
    equilibrium of its own moments.  Checked against an independent numpy
    collision on every inlet node, over several steps with new values."""
    rng = np.random.default_rng(7)
    nr, nc = 14, 18
    cls = _open_box(nr, nc)
    tau, te, to = 0.83, 0.74, 1.31
    lat = O.OracleLattice(cls, collision=coll, tau=tau, tau_even=te, tau_odd=to,
                          inlet_policy=O.ZOUHE_VELOCITY, outlet_policy=O.ZOUHE_PRESSURE)
    ij, _ = lat.node_list(O.LIST_INLET)
    n = len(ij)
    lat.set_open_values(1, rho=0.999)
    lat.set_open_values(0, *_open_values(rng, n))
    lat.init()
    lat.step(5)
    ii, jj = ij[:, 0], ij[:, 1]
    for _ in range(4):
        ux, uy = _open_values(rng, n)
        lat.set_open_values(0, ux=ux, uy=uy)
        p = lat.populations()[:, jj, ii]
        lat.step(1)
        a = lat.after_collision()[:, jj, ii]
        e = _feq_textbook(1.0, ux, uy)
        if coll == O.BGK:
            want = (1 - 1 / tau) * p + (1 / tau) * e
        else:
            sym = 0.5 * ((p + p[OPP]) - (e + e[OPP]))
            asym = 0.5 * ((p - p[OPP]) - (e - e[OPP]))
            want = p - sym / te - asym / to
        assert np.max(np.abs(a - want)) < 2e-16 * 8
        # the moments of these nodes are not (1, U): the check above would
        # fail with the equilibrium of the moments
        r_mom = p.sum(axis=0)
        assert np.min(np.abs(r_mom - 1.0)) > 1e-9
        rho, u = lat.macro()
        assert np.all(rho[jj, ii] == 1.0)
        assert np.all(u[0][jj, ii] == ux) and np.all(u[1][jj, ii] == uy)


def test_zouhe_closes_each_step_with_that_steps_values():
    """Q3 (Lattice.hpp:351,405-410; InletOutletInitializer.hpp:90-129): the
    values set before step i (update_nodes at the top of the loop body)
    close the Zou-He boundary at the end of step i.  After the step every
    TOP and LEFT inlet node holds density rho_zh -- the Zou-He mass balance
    of its known populations -- and momentum rho_zh U(i); it does not match
    the values of step i-1 or i+1.  The corner-typed inlet (0, 0) is left
    alone (ZouHePolicy.hpp:119-120): its out-of-domain slots keep w_d."""
    rng = np.random.default_rng(11)
    nr, nc = 14, 18
    cls = _open_box(nr, nc)
    lat = O.OracleLattice(cls, collision=O.TRT, tau_even=0.7, tau_odd=1.4,
                          inlet_policy=O.ZOUHE_VELOCITY, outlet_policy=O.ZOUHE_PRESSURE)
    ij, tag = lat.node_list(O.LIST_INLET)
    n = len(ij)
    ii, jj = ij[:, 0], ij[:, 1]
    top, left = tag == O.BT_TOP, tag == O.BT_LEFT
    corner = (ii == 0) & (jj == 0)
    assert top.sum() == nc - 2 and left.sum() == nr - 2 and corner.sum() == 1
    lat.set_open_values(1, rho=1.0)
    U = [_open_values(rng, n) for _ in range(8)]
    lat.set_open_values(0, ux=U[0][0], uy=U[0][1])
    lat.init()
    w = O.feq(1.0, 0.0, 0.0)
    for s in range(7):
        lat.set_open_values(0, ux=U[s][0], uy=U[s][1])
        lat.step(1)
        p = lat.populations()[:, jj, ii]
        rho = p.sum(axis=0)
        jx, jy = (CX[:, None] * p).sum(axis=0), (CY[:, None] * p).sum(axis=0)
        sel = top | left
        # closed-form mass balance from the known populations
        known_top = p[0] + p[1] + p[3] + 2 * (p[2] + p[5] + p[6])
        known_left = p[0] + p[2] + p[4] + 2 * (p[3] + p[7] + p[6])
        rho_zh = np.where(top, known_top / (1 + U[s][1]), known_left / (1 - U[s][0]))
        assert np.max(np.abs(rho[sel] - rho_zh[sel])) < 1e-14
        assert np.max(np.abs(jx[sel] - rho[sel] * U[s][0][sel])) < 1e-14
        assert np.max(np.abs(jy[sel] - rho[sel] * U[s][1][sel])) < 1e-14
        for other in (s - 1, s + 1):
            if other >= 0:
                assert np.min(np.abs(jx[sel] - rho[sel] * U[other][0][sel])) > 1e-8
        # slots of (0, 0) whose source (0, 0) - c_d lies outside the domain:
        # push offsets with +1 row (d = 4, 7, 8) or +1 column (d = 1, 5, 8)
        for d in (1, 4, 5, 7, 8):
            assert p[d][corner][0] == w[d]


def test_reduce_matches_population_sums():
    """orc_reduce: out[0] = mass of the wet nodes, out[1], out[2] = momentum
    of the fluid nodes, out[3] = fluid mass, out[4] = fluid count, against
    numpy sums over populations() on a lattice with inlet, outlet and
    obstacle nodes (so fluid != wet)."""
    rng = np.random.default_rng(3)
    nr, nc = 14, 18
    cls = _open_box(nr, nc)
    lat = O.OracleLattice(cls, collision=O.BGK, tau=0.9, inlet_policy=O.ZOUHE_VELOCITY,
                          outlet_policy=O.ZOUHE_PRESSURE)
    ij, _ = lat.node_list(O.LIST_INLET)
    lat.set_open_values(0, *_open_values(rng, len(ij)))
    lat.set_open_values(1, rho=0.998)
    lat.init()
    lat.step(30)
    p = lat.populations()
    fl = cls == O.FLUID
    wet = wet_mask(cls)
    out = lat.reduce()
    pf = p[:, fl]
    assert np.isclose(out[0], p[:, wet].sum(), rtol=1e-14, atol=0)
    assert np.isclose(out[1], (CX[:, None] * pf).sum(), rtol=0, atol=1e-14)
    assert np.isclose(out[2], (CY[:, None] * pf).sum(), rtol=0, atol=1e-14)
    assert np.isclose(out[3], pf.sum(), rtol=1e-14, atol=0)
    assert out[4] == fl.sum()
    assert abs(out[1]) > 1e-3  # the flow carries momentum: the check has teeth


# ------------------------------------------- partially saturated bounce-back

def _ps_lattice(b=None, tau_ps=None, p=None):
    """Lid cavity 22 x 26 with two obstacles (a disc and a block), BGK;
    PSBounceBackPolicy on the obstacles when tau_ps is given."""
    nr, nc = 22, 26
    cls = _box(nr, nc)
    cls[1:-1, 0] = O.INLET
    ii, jj = np.meshgrid(np.arange(nr), np.arange(nc))
    cls[(ii - 9) ** 2 + (jj - 8) ** 2 <= 9] = O.OBSTACLE
    cls[16:19, 13:17] = O.OBSTACLE
    lat = O.OracleLattice(cls, collision=O.BGK, tau=0.8, inlet_policy=O.ZOUHE_VELOCITY, outlet_policy=O.NOOP)
    ni = lat.count(O.LIST_INLET)
    lat.set_open_values(0, ux=np.full(ni, 0.06), uy=np.zeros(ni))
    if tau_ps is not None:
        lat.set_partial_bounce_back(tau_ps, b)
    lat.init()
    if p is not None:
        lat.set_populations(p)
    return cls, lat


def _ps_slots(lat):
    """(row, column, direction, obstacle index) of every slot a wet obstacle
    writes: direction d of the neighbour o + c_d for each bit d."""
    ij, masks = lat.node_list(O.LIST_OBSTACLE)
    out = []
    for q, ((i, j), m) in enumerate(zip(ij, masks)):
        for d in range(1, 9):
            if m & (1 << d):
                out.append((i + DROW[d], j + DCOL[d], d, q))
    return out


def _state_after(steps):
    cls, lat = _ps_lattice(tau_ps=0.9)
    lat.step(steps)
    return cls, lat.populations()


def _moments_feq(p, i, j):
    f = p[:, j, i]
    rho = f.sum()
    ux, uy = (CX * f).sum() / rho, (CY * f).sum() / rho
    return _feq_textbook(rho, np.array([ux]), np.array([uy]))[:, 0]


def test_psbb_weight_zero_with_infinite_tau_is_bounce_back():
    """b = 0 and tau -> infinity leave p_q - (p_q - e_q) / tau = p_q: the
    policy reduces to half-way bounce-back, bit for bit, over 60 steps."""
    cls, plain = _ps_lattice()
    _, ps = _ps_lattice(b=np.zeros(plain.count(O.LIST_OBSTACLE)), tau_ps=1e300)
    plain.step(60)
    ps.step(60)
    wet = wet_mask(cls)
    assert np.array_equal(plain.populations()[:, wet], ps.populations()[:, wet])


def test_psbb_closed_forms_of_one_step():
    """From the same state, one step with weight b = 1 gives, on every slot an
    obstacle writes, P_d(n) - e_d(n) + w_d (the old slot value relaxed by
    the equilibrium of n, the obstacle's rho = 1); with b = 0 and tau = 1 it
    gives e_opp(d)(n); with b = 0.3 the weighted mix of both.  Equilibria
    from the textbook form at the moments of the state."""
    cls, p0 = _state_after(40)
    _, lat = _ps_lattice(tau_ps=0.9)
    n_obst = lat.count(O.LIST_OBSTACLE)
    slots = _ps_slots(lat)
    assert len(slots) > 20
    res = {}
    for name, b, tau in (("b1", np.ones(n_obst), 0.9), ("b0", np.zeros(n_obst), 1.0),
                         ("mix", np.full(n_obst, 0.3), 1.0)):
        _, l2 = _ps_lattice(b=b, tau_ps=tau, p=p0)
        l2.step(1)
        res[name] = l2.populations()
    for (i, j, d, q) in slots:
        e = _moments_feq(p0, i, j)
        want1 = p0[d, j, i] - e[d] + W[d]
        assert abs(res["b1"][d, j, i] - want1) <= 1e-15, (i, j, d)
        assert abs(res["b0"][d, j, i] - e[OPP[d]]) <= 1e-15, (i, j, d)
        mix = 0.7 * res["b0"][d, j, i] + 0.3 * res["b1"][d, j, i]
        assert abs(res["mix"][d, j, i] - mix) <= 1e-15, (i, j, d)
    # the other wet slots do not depend on the weights
    wet = wet_mask(cls)
    other = np.ones_like(res["b1"], dtype=bool)
    for (i, j, d, q) in slots:
        other[d, j, i] = False
    other &= wet[None]
    assert np.array_equal(res["b0"][other], res["mix"][other])


def test_psbb_uses_each_obstacles_own_weight():
    """Weights differ per obstacle (list order): every written slot is the
    mix, with the weight of the obstacle it comes from, of the b = 0 and
    b = 1 results."""
    cls, p0 = _state_after(25)
    _, lat = _ps_lattice(tau_ps=1.3)
    n_obst = lat.count(O.LIST_OBSTACLE)
    rng = np.random.default_rng(2)
    bw = rng.uniform(0, 1, n_obst)
    out = {}
    for name, b in (("0", np.zeros(n_obst)), ("1", np.ones(n_obst)), ("w", bw)):
        _, l2 = _ps_lattice(b=b, tau_ps=1.3, p=p0)
        l2.step(1)
        out[name] = l2.populations()
    for (i, j, d, q) in _ps_slots(lat):
        mix = (1 - bw[q]) * out["0"][d, j, i] + bw[q] * out["1"][d, j, i]
        assert abs(out["w"][d, j, i] - mix) <= 2e-15, (i, j, d, q)
