"""ctypes front end of the CPU oracle (test infrastructure only).

Only the tests, the smoke check and the CPU-baseline leg of the benchmark use
this module, and only as the checker.  The library it loads restates the
LLALBM Serial policy (see lbm_oracle.c for the reference file:line map).

Arrays follow the reference's Eigen ColMajor layout: populations are
``(9, NC, NR)`` in C order (element (i, j, d) at i + NR*(j + NC*d)), the
density is ``(NC, NR)`` and the velocity ``(2, NC, NR)``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblbm_oracle.so")

# reference InputNodeType (aliases.hpp:41-48)
FLUID, SOLID, BOUNDARY, INLET, OUTLET, OBSTACLE = range(6)
# reference btypes (aliases.hpp:28-38)
BT_NONE, BT_TOP, BT_RIGHT, BT_BOTTOM, BT_LEFT = range(5)
BGK, TRT = 0, 1
ZOUHE_VELOCITY, ZOUHE_PRESSURE, NOOP = 0, 1, 2
LIST_FLUID, LIST_WALL, LIST_INLET, LIST_OUTLET, LIST_OBSTACLE = range(5)

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        dp = C.POINTER(C.c_double)
        i64p = C.POINTER(C.c_int64)
        i32p = C.POINTER(C.c_int32)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int64, C.c_int64, C.c_void_p, C.c_int, C.c_double, C.c_double,
                                 C.c_double, C.c_int, C.c_int, C.POINTER(C.c_int)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_count.restype = C.c_int64
        L.orc_count.argtypes = [C.c_void_p, C.c_int]
        L.orc_get_list.argtypes = [C.c_void_p, C.c_int, i64p, i32p]
        L.orc_set_open_values.argtypes = [C.c_void_p, C.c_int, dp, dp, dp]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_feq.argtypes = [C.c_double, C.c_double, C.c_double, dp]
        L.orc_moments.argtypes = [dp, dp, dp, dp]
        L.orc_collide_bgk_node.argtypes = [C.c_double, dp, dp, dp]
        L.orc_collide_trt_node.argtypes = [C.c_double, C.c_double, dp, dp, dp]
        L.orc_zouhe_velocity_node.restype = C.c_double
        L.orc_zouhe_velocity_node.argtypes = [C.c_int, dp, C.c_double, C.c_double]
        L.orc_zouhe_pressure_node.restype = C.c_double
        L.orc_zouhe_pressure_node.argtypes = [C.c_int, dp, C.c_double]
        L.orc_init.argtypes = [C.c_void_p]
        L.orc_step.argtypes = [C.c_void_p, C.c_int64]
        L.orc_get_step.restype = C.c_int64
        L.orc_get_step.argtypes = [C.c_void_p]
        for f in (L.orc_get_populations, L.orc_set_populations, L.orc_get_after_collision):
            f.argtypes = [C.c_void_p, dp]
        L.orc_get_macro.argtypes = [C.c_void_p, dp, dp]
        L.orc_lift_drag.argtypes = [C.c_void_p, i64p, i32p, C.c_int64, dp, dp]
        L.orc_reduce.argtypes = [C.c_void_p, dp]
        L.orc_set_by_value_copies.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_partial_bounce_back.argtypes = [C.c_void_p, C.c_double, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


# ---- single-node kernels ------------------------------------------------

def feq(rho, ux, uy):
    e = np.zeros(9)
    lib().orc_feq(rho, ux, uy, _dp(e))
    return e


def moments(p):
    p = np.ascontiguousarray(p, dtype=np.float64)
    r, ux, uy = C.c_double(), C.c_double(), C.c_double()
    lib().orc_moments(_dp(p), C.byref(r), C.byref(ux), C.byref(uy))
    return r.value, ux.value, uy.value


def collide_bgk(tau, p, e):
    p = np.ascontiguousarray(p, dtype=np.float64)
    e = np.ascontiguousarray(e, dtype=np.float64)
    o = np.zeros(9)
    lib().orc_collide_bgk_node(tau, _dp(p), _dp(e), _dp(o))
    return o


def collide_trt(te, to, p, e):
    p = np.ascontiguousarray(p, dtype=np.float64)
    e = np.ascontiguousarray(e, dtype=np.float64)
    o = np.zeros(9)
    lib().orc_collide_trt_node(te, to, _dp(p), _dp(e), _dp(o))
    return o


def zouhe_velocity(btype, p, ux, uy):
    q = np.array(p, dtype=np.float64)
    rho = lib().orc_zouhe_velocity_node(btype, _dp(q), ux, uy)
    return q, rho


def zouhe_pressure(btype, p, rho):
    q = np.array(p, dtype=np.float64)
    un = lib().orc_zouhe_pressure_node(btype, _dp(q), rho)
    return q, un


# ---- whole lattice ------------------------------------------------------

class OracleLattice:
    """Serial-policy lattice driven like ``Lattice::perform_lbm``."""

    def __init__(self, node_class, collision=BGK, tau=1.0, tau_even=1.0, tau_odd=1.0,
                 inlet_policy=ZOUHE_VELOCITY, outlet_policy=NOOP):
        cls = np.ascontiguousarray(node_class, dtype=np.uint8)  # shape (NC, NR): ColMajor plane
        self.nc, self.nr = cls.shape
        err = C.c_int(0)
        h = lib().orc_create(self.nr, self.nc, cls.ctypes.data, collision, tau, tau_even, tau_odd,
                             inlet_policy, outlet_policy, C.byref(err))
        if not h:
            raise ValueError(f"orc_create failed with code {err.value}")
        self._h = h

    def close(self):
        if self._h:
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def count(self, which):
        return int(lib().orc_count(self._h, which))

    def node_list(self, which):
        n = self.count(which)
        ij = np.zeros((max(n, 1), 2), dtype=np.int64)
        tag = np.zeros(max(n, 1), dtype=np.int32)
        lib().orc_get_list(self._h, which, ij.ctypes.data_as(C.POINTER(C.c_int64)),
                           tag.ctypes.data_as(C.POINTER(C.c_int32)))
        return ij[:n], tag[:n]

    def set_open_values(self, which, ux=None, uy=None, rho=None):
        n = self.count(LIST_INLET + which)
        arrs = []
        for a in (ux, uy, rho):
            if a is None:
                arrs.append(None)
            else:
                a = np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (n,)))
                arrs.append(a)
        lib().orc_set_open_values(self._h, which, *[(_dp(a) if a is not None else None) for a in arrs])

    def init(self):
        lib().orc_init(self._h)

    def step(self, n=1):
        lib().orc_step(self._h, int(n))

    @property
    def time(self):
        return int(lib().orc_get_step(self._h))

    def populations(self):
        out = np.empty((9, self.nc, self.nr))
        lib().orc_get_populations(self._h, _dp(out))
        return out

    def set_populations(self, p):
        p = np.ascontiguousarray(p, dtype=np.float64)
        assert p.shape == (9, self.nc, self.nr)
        lib().orc_set_populations(self._h, _dp(p))

    def after_collision(self):
        out = np.empty((9, self.nc, self.nr))
        lib().orc_get_after_collision(self._h, _dp(out))
        return out

    def macro(self):
        rho = np.empty((self.nc, self.nr))
        u = np.empty((2, self.nc, self.nr))
        lib().orc_get_macro(self._h, _dp(rho), _dp(u))
        return rho, u

    def lift_drag(self, points=None, masks=None, per_point=False):
        ld = np.zeros(2)
        if points is None:
            n = self.count(LIST_OBSTACLE)
            ij_p, m_p = None, None
        else:
            pts = np.ascontiguousarray(points, dtype=np.int64)
            msk = np.ascontiguousarray(masks, dtype=np.int32)
            n = len(pts)
            ij_p = pts.ctypes.data_as(C.POINTER(C.c_int64))
            m_p = msk.ctypes.data_as(C.POINTER(C.c_int32))
        pp = np.zeros((max(n, 1), 2))
        lib().orc_lift_drag(self._h, ij_p, m_p, n, _dp(ld), _dp(pp) if per_point else None)
        if per_point:
            return ld[0], ld[1], pp[:n]
        return ld[0], ld[1]

    def set_partial_bounce_back(self, tau, b=None):
        """PSBounceBackPolicy<2> on the obstacle list: the policy's tau and one
        weight per wet obstacle in list order (None = 1); tau <= 0 = off."""
        arr = None if b is None else np.ascontiguousarray(b, dtype=np.float64)
        lib().orc_set_partial_bounce_back(self._h, float(tau), None if arr is None else _dp(arr))

    def set_by_value_copies(self, on=True):
        """CPU timing baseline only: make the reference's by-value copies of
        the rho / u fields in every boundary and update_nodes call
        (OMPPolicy.hpp:50-72,139).  Results are unchanged."""
        lib().orc_set_by_value_copies(self._h, int(bool(on)))

    def reduce(self):
        out = np.zeros(5)
        lib().orc_reduce(self._h, _dp(out))
        return out
