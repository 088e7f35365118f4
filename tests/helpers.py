"""Helpers shared by the tests: physical velocity set, oracle set-up from a
workload, and comparisons restricted to wet-node slots."""
import numpy as np

import lbm_amd as L
import oracle as O

# Physical velocity of each D2Q9 direction, read off the reference's moment
# sums (DefaultEquilibrium.hpp:148-149): ux = f1+f5+f8-f3-f6-f7,
# uy = f2+f5+f6-f4-f7-f8, with x along +column and y along -row.
CX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
CY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
# Displacement of a population of direction d per step in (row, column).
DROW = -CY
DCOL = CX
OPP = np.array([int(np.nonzero((CX == -CX[d]) & (CY == -CY[d]))[0][0]) for d in range(9)])


def wet_mask(node_class):
    c = node_class
    return (c == L.FLUID) | (c == L.INLET) | (c == L.OUTLET)


def oracle_from(wl, threads=None):
    if threads:
        O.set_threads(threads)
    lat = O.OracleLattice(wl.node_class, **wl.lattice_kwargs())
    if wl.inlet:
        lat.set_open_values(0, **wl.inlet)
    if wl.outlet:
        lat.set_open_values(1, **wl.outlet)
    lat.init()
    return lat


def max_abs_wet(a, b, node_class):
    m = wet_mask(node_class)
    return float(np.max(np.abs(a[:, m] - b[:, m]))) if m.any() else 0.0
