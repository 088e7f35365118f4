"""GPU tests beyond bit parity: the FP64 path against published physics
(Ghia, Ghia & Shin 1982), size-independent properties at the full
configuration sizes, and the stated error of the shifted FP32 / FP16
storage paths against the FP64 path."""
import sys

import numpy as np
import pytest

import lbm_amd as L
import workloads as W
from helpers import wet_mask

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


# Ghia, Ghia & Shin (1982), J. Comput. Phys. 48:387, Table I (Re = 100):
# u / U_lid on the vertical centre line versus y (0 = bottom, 1 = lid).
GHIA_Y = np.array([0.0547, 0.0625, 0.0703, 0.1016, 0.1719, 0.2813, 0.4531, 0.5000, 0.6172, 0.7344, 0.8516,
                   0.9531, 0.9609, 0.9688, 0.9766])
GHIA_U = np.array([-0.03717, -0.04192, -0.04775, -0.06434, -0.10150, -0.15662, -0.21090, -0.20581, -0.13641,
                   0.00332, 0.23151, 0.68717, 0.73722, 0.78871, 0.84123])
# Table II (Re = 100): v / U_lid on the horizontal centre line versus x.
GHIA_X = np.array([0.0625, 0.0703, 0.0781, 0.0938, 0.1563, 0.2266, 0.2344, 0.5000, 0.8047, 0.8594, 0.9063,
                   0.9453, 0.9531, 0.9609, 0.9688])
GHIA_V = np.array([0.09233, 0.10091, 0.10890, 0.12317, 0.16077, 0.17507, 0.17527, 0.05454, -0.24533, -0.22445,
                   -0.16914, -0.10313, -0.08864, -0.07391, -0.05906])


def test_cavity129_matches_ghia_re100():
    wl = W.cavity(129)
    say(f"case {wl.name}: 30000 steps, Ghia et al. Re=100")
    lat = wl.make_lattice(L.F64)
    lat.step(30000, L.STEP_MACRO | L.STEP_CHECK)
    _, u = lat.macro()
    n = 129
    # Lid (Zou-He) on row 0, half-way walls at row n-1.5 and columns 0.5, n-1.5.
    height = n - 1.5
    width = n - 2
    ycoord = (height - np.arange(n)) / height          # row i -> y
    jc = (n - 1) // 2                                  # column at x = 0.5
    ucol = u[0, jc, :] / 0.1
    uprof = np.interp(GHIA_Y, ycoord[::-1], ucol[::-1])
    xcoord = (np.arange(n) - 0.5) / width
    ic = int(round(height - 0.5 * height))             # row at y = 0.5
    vrow = u[1, :, ic] / 0.1
    vprof = np.interp(GHIA_X, xcoord, vrow)
    say(f"  max |u - Ghia| = {np.max(np.abs(uprof - GHIA_U)):.4f}, max |v - Ghia| = {np.max(np.abs(vprof - GHIA_V)):.4f}")
    assert np.max(np.abs(uprof - GHIA_U)) < 0.03
    assert np.max(np.abs(vprof - GHIA_V)) < 0.03


def test_porous_full_size_rest_state_and_mass():
    """8192 x 8192 porous domain at full size: with equal inlet and outlet
    density the fluid must stay at rest (|u| ~ round-off) and the wet mass
    must not drift."""
    wl = W.porous(8192)
    no = int(np.count_nonzero(wl.node_class == L.OUTLET))
    ni = int(np.count_nonzero(wl.node_class == L.INLET))
    say(f"case {wl.name}: rest state, {wl.wet_count()} wet nodes, solid {wl.info['solid']:.4f}")
    lat = L.Lattice(wl.node_class, **wl.lattice_kwargs())
    lat.set_open_values(0, rho=np.ones(ni))
    lat.set_open_values(1, rho=np.ones(no))
    lat.init_equilibrium()
    m0 = lat.reduce()[0]
    lat.step(100, L.STEP_MACRO | L.STEP_CHECK)
    m1 = lat.reduce()
    _, u = lat.macro()
    umax = float(np.max(np.abs(u)))
    say(f"  max|u| = {umax:.3e}, mass drift = {(m1[0] - m0) / m0:.3e}")
    assert umax < 1e-13
    assert abs(m1[0] - m0) < 1e-12 * m0


def _l2rel(a, b, m):
    return float(np.sqrt(np.sum((a[..., m] - b[..., m]) ** 2) / np.sum(b[..., m] ** 2)))


# Velocity L2 error bounds after the stated runs; measured values are 5-10x
# below them (FP32: cavity 1.2e-7, porous 1.6e-7; FP16 rounded to nearest:
# 5.2e-3 / 2.4e-3).  FP16 with its default stochastic rounding is held to
# the same bound here; over long runs it is far more accurate (see the
# 4096^2 cavity errors stated by bench.py).
SHIFTED = [(L.F32_SHIFTED, 0, 1e-6), (L.F16_SHIFTED, L.OPT_F16_ROUND_TO_NEAREST, 2.5e-2),
           (L.F16_SHIFTED, 0, 2.5e-2)]


@pytest.mark.parametrize("storage,options,tol_u", SHIFTED)
def test_shifted_storage_error_cavity(storage, options, tol_u):
    wl = W.cavity(129)
    say(f"case {wl.name}: storage {storage} options {options} vs FP64, 5000 steps")
    ref = wl.make_lattice(L.F64)
    lo = wl.make_lattice(storage, options=options)
    ref.step(5000, L.STEP_MACRO)
    lo.step(5000, L.STEP_MACRO)
    r0, u0 = ref.macro()
    r1, u1 = lo.macro()
    m = wl.node_class == L.FLUID
    eu = _l2rel(u1, u0, m)
    er = float(np.sqrt(np.sum((r1[m] - r0[m]) ** 2) / np.sum((r0[m] - 1) ** 2)))
    say(f"  L2 rel velocity error {eu:.3e}, density error {er:.3e}, scale {lo.scale}")
    assert eu < tol_u


@pytest.mark.parametrize("storage,options,tol_u", SHIFTED)
def test_shifted_storage_error_porous(storage, options, tol_u):
    wl = W.porous(512, seed=5)
    say(f"case {wl.name}: storage {storage} options {options} vs FP64, 2000 steps")
    ref = wl.make_lattice(L.F64)
    lo = wl.make_lattice(storage, options=options)
    ref.step(2000, L.STEP_MACRO)
    lo.step(2000, L.STEP_MACRO)
    r0, u0 = ref.macro()
    r1, u1 = lo.macro()
    m = wl.node_class == L.FLUID
    eu = _l2rel(u1, u0, m)
    er = float(np.sqrt(np.sum((r1[m] - r0[m]) ** 2) / np.sum((r0[m] - 1) ** 2)))
    red0, red1 = ref.reduce(), lo.reduce()
    say(f"  L2 rel velocity error {eu:.3e}, density error {er:.3e}, mass rel diff {(red1[0]-red0[0])/red0[0]:.3e}")
    assert eu < tol_u
