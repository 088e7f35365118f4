"""Python front end of the MI355X D2Q9 engine (ctypes over include/lbm_amd.h).

The engine itself is the HIP library ``liblbm_amd.so`` next to this file;
this module only marshals arrays.  There is no CPU fallback: if the library
is missing or no GPU is present, the calls fail with an error.

Arrays use the reference's Eigen ColMajor layout, expressed as C-ordered
numpy arrays with the row index last:
    node classes   (NC, NR)      uint8   (values: FLUID..OBSTACLE)
    populations    (9, NC, NR)   float64 (element (i, j, d) at [d, j, i])
    density        (NC, NR)      float64
    velocity       (2, NC, NR)   float64
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblbm_amd.so")

# lbm_node_class = reference InputNodeType (aliases.hpp:41-48)
FLUID, SOLID, BOUNDARY, INLET, OUTLET, OBSTACLE = range(6)
BGK, TRT = 0, 1
ZOUHE_VELOCITY, ZOUHE_PRESSURE, OPEN_NOOP = 0, 1, 2
F64, F32_SHIFTED, F16_SHIFTED = 0, 1, 2
STEP_MACRO, STEP_CHECK, STEP_NO_GRAPH = 1, 2, 4
OPT_F16_ROUND_TO_NEAREST = 1
OPT_PLANAR_LAYOUT = 2
LAYOUT_BLOCKED, LAYOUT_PLANES = 0, 1
# lbm_set_launch shapes: one wave per run of up to 64 consecutive wet rows;
# the same runs with each lane's addressing read from a table
LAUNCH_WET_RUNS = 64
LAUNCH_RUN_TABLE = 65
ERR_NAMES = {0: "OK", 1: "INVALID", 2: "HIP", 3: "NAN", 4: "STATE", 5: "UNSUPPORTED", 6: "NOMEM", 7: "RANGE"}
ERR_INVALID, ERR_STATE, ERR_NAN, ERR_UNSUPPORTED, ERR_RANGE = 1, 4, 3, 5, 7

EXPORTED = [
    "lbm_version", "lbm_last_error", "lbm_create", "lbm_destroy", "lbm_set_relaxation",
    "lbm_set_partial_bounce_back", "lbm_set_geometry",
    "lbm_node_counts", "lbm_open_nodes", "lbm_set_open_values", "lbm_init_equilibrium",
    "lbm_upload_populations", "lbm_download_populations", "lbm_download_after_collision", "lbm_step",
    "lbm_get_time", "lbm_download_macro", "lbm_download_macro_async", "lbm_wait", "lbm_host_alloc",
    "lbm_host_free", "lbm_lift_drag", "lbm_lift_drag_async", "lbm_reduce", "lbm_reduce_device", "lbm_synchronize",
    "lbm_get_stream",
    "lbm_get_scale", "lbm_geometry_cavity", "lbm_geometry_channel", "lbm_geometry_add_disc",
    "lbm_geometry_porous", "lbm_set_launch", "lbm_get_launch", "lbm_set_occupancy", "lbm_get_occupancy",
    "lbm_get_layout",
    "lbm_autotune", "lbm_batch_create", "lbm_batch_destroy", "lbm_batch_set_launch", "lbm_batch_get_launch",
    "lbm_batch_autotune", "lbm_batch_step", "lbm_batch_reduce", "lbm_batch_info", "lbm_batch_get_stream",
    "lbm_batch_synchronize", "lbm_copy_bandwidth", "lbm_pattern_bandwidth",
]


class LbmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"lbm error {ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


class lbm_desc(C.Structure):
    _fields_ = [
        ("n_rows", C.c_int64), ("n_cols", C.c_int64),
        ("collision", C.c_int32), ("storage", C.c_int32),
        ("tau", C.c_double), ("tau_even", C.c_double), ("tau_odd", C.c_double),
        ("inlet_policy", C.c_int32), ("outlet_policy", C.c_int32),
        ("f16_scale", C.c_double),
        ("device", C.c_int32), ("options", C.c_int32),
    ]


_lib = None


def build() -> None:
    """Compiles liblbm_amd.so for gfx950 in place (hipcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run build() (hipcc, gfx950) first")
        L = C.CDLL(LIB_PATH)
        vp, dp = C.c_void_p, C.POINTER(C.c_double)
        i64p, i32p = C.POINTER(C.c_int64), C.POINTER(C.c_int32)
        L.lbm_version.restype = C.c_char_p
        L.lbm_last_error.restype = C.c_char_p
        sig = {
            "lbm_create": [C.POINTER(lbm_desc), C.POINTER(vp)],
            "lbm_destroy": [vp],
            "lbm_set_relaxation": [vp, C.c_double, C.c_double, C.c_double],
            "lbm_set_geometry": [vp, vp],
            "lbm_set_partial_bounce_back": [vp, C.c_double, dp],
            "lbm_node_counts": [vp, i64p],
            "lbm_open_nodes": [vp, C.c_int, i64p, i64p],
            "lbm_set_open_values": [vp, C.c_int, dp, dp, dp],
            "lbm_init_equilibrium": [vp],
            "lbm_upload_populations": [vp, dp],
            "lbm_download_populations": [vp, dp],
            "lbm_download_after_collision": [vp, dp],
            "lbm_step": [vp, C.c_int64, C.c_uint32],
            "lbm_get_time": [vp, i64p],
            "lbm_download_macro": [vp, dp, dp],
            "lbm_download_macro_async": [vp, dp, dp, i64p],
            "lbm_wait": [vp, C.c_int64],
            "lbm_host_alloc": [C.c_int64, C.POINTER(vp)],
            "lbm_host_free": [vp],
            "lbm_lift_drag": [vp, i64p, i32p, C.c_int64, dp, dp],
            "lbm_lift_drag_async": [vp, i64p, i32p, C.c_int64, dp, dp, i64p],
            "lbm_reduce": [vp, dp],
            "lbm_reduce_device": [vp, vp],
            "lbm_synchronize": [vp],
            "lbm_get_stream": [vp, C.POINTER(vp)],
            "lbm_get_scale": [vp, dp],
            "lbm_geometry_cavity": [C.c_int64, C.c_int64, vp],
            "lbm_geometry_channel": [C.c_int64, C.c_int64, vp],
            "lbm_geometry_add_disc": [C.c_int64, C.c_int64, vp, C.c_int64, C.c_int64, C.c_int64, i64p],
            "lbm_geometry_porous": [C.c_int64, C.c_int64, vp, C.c_uint64, C.c_int64, C.c_int64, C.c_double, dp,
                                    i64p],
            "lbm_set_launch": [vp, C.c_int32],
            "lbm_get_launch": [vp, i32p],
            "lbm_set_occupancy": [vp, C.c_int32],
            "lbm_get_layout": [vp, i32p],
            "lbm_get_occupancy": [vp, i32p],
            "lbm_autotune": [vp, C.c_int32, i32p, dp],
            "lbm_batch_create": [C.POINTER(vp), C.c_int64, C.POINTER(vp)],
            "lbm_batch_destroy": [vp],
            "lbm_batch_set_launch": [vp, C.c_int32, C.c_int64],
            "lbm_batch_get_launch": [vp, i32p, i64p],
            "lbm_batch_autotune": [vp, C.c_int32, i32p, i64p, dp],
            "lbm_batch_step": [vp, C.c_int64, C.c_uint32],
            "lbm_batch_reduce": [vp, dp],
            "lbm_batch_info": [vp, i64p, i64p],
            "lbm_batch_get_stream": [vp, C.POINTER(vp)],
            "lbm_batch_synchronize": [vp],
            "lbm_copy_bandwidth": [C.c_int32, C.c_int64, C.c_int32, dp],
            "lbm_pattern_bandwidth": [C.c_int32, C.c_int32, C.c_int32, C.c_int64, C.c_int64, C.c_int32, dp],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        _lib = L
    return _lib


def _check(code):
    if code != 0:
        raise LbmError(code, lib().lbm_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def version() -> str:
    return lib().lbm_version().decode()


# ---- geometry helpers ----------------------------------------------------

def geometry_cavity(nr, nc):
    cls = np.zeros((nc, nr), dtype=np.uint8)
    _check(lib().lbm_geometry_cavity(nr, nc, cls.ctypes.data))
    return cls


def geometry_channel(nr, nc):
    cls = np.zeros((nc, nr), dtype=np.uint8)
    _check(lib().lbm_geometry_channel(nr, nc, cls.ctypes.data))
    return cls


def add_disc(cls, ci, cj, r):
    nc, nr = cls.shape
    assert cls.flags.c_contiguous and cls.dtype == np.uint8
    added = C.c_int64(0)
    _check(lib().lbm_geometry_add_disc(nr, nc, cls.ctypes.data, ci, cj, r, C.byref(added)))
    return added.value


def geometry_porous(cls, seed, rmin, rmax, target_solid):
    nc, nr = cls.shape
    assert cls.flags.c_contiguous and cls.dtype == np.uint8
    solid, nd = C.c_double(0), C.c_int64(0)
    _check(lib().lbm_geometry_porous(nr, nc, cls.ctypes.data, seed, rmin, rmax, target_solid, C.byref(solid),
                                     C.byref(nd)))
    return solid.value, nd.value


# ---- measurement -----------------------------------------------------------

def copy_bandwidth(nbytes, reps=20, device=0):
    """Practical HBM ceiling of a plain stream: the best rate (read + write
    GB/s) of the library's copy kernel over nbytes (lbm_copy_bandwidth)."""
    gbs = C.c_double(0)
    _check(lib().lbm_copy_bandwidth(device, int(nbytes), reps, C.byref(gbs)))
    return gbs.value


# access patterns of lbm_pattern_bandwidth
PATTERN_AB, PATTERN_AA, PATTERN_AA_EVEN, PATTERN_AA_ODD, PATTERN_BLOCKED = 0, 1, 2, 3, 4
PATTERN_NAMES = {PATTERN_AB: "ab_pull", PATTERN_AA: "aa", PATTERN_AA_EVEN: "aa_even", PATTERN_AA_ODD: "aa_odd",
                 PATTERN_BLOCKED: "blocked_pull"}


def pattern_bandwidth(pattern, elem_bytes, nr, nc, reps=20, device=0):
    """Rate (read + write GB/s, 18 elements per node) of a D2Q9 step's access
    pattern without arithmetic on a dense nr x nc periodic lattice
    (lbm_pattern_bandwidth)."""
    gbs = C.c_double(0)
    _check(lib().lbm_pattern_bandwidth(device, pattern, elem_bytes, int(nr), int(nc), reps, C.byref(gbs)))
    return gbs.value


# ---- lattice ---------------------------------------------------------------

class Lattice:
    """A device-resident D2Q9 lattice (one lbm_ctx)."""

    def __init__(self, node_class, collision=BGK, tau=1.0, tau_even=1.0, tau_odd=1.0,
                 inlet_policy=ZOUHE_VELOCITY, outlet_policy=OPEN_NOOP, storage=F64, f16_scale=0.0, device=0,
                 options=0):
        cls = np.ascontiguousarray(node_class, dtype=np.uint8)
        self.nc, self.nr = cls.shape
        d = lbm_desc(n_rows=self.nr, n_cols=self.nc, collision=collision, storage=storage, tau=tau,
                     tau_even=tau_even, tau_odd=tau_odd, inlet_policy=inlet_policy, outlet_policy=outlet_policy,
                     f16_scale=f16_scale, device=device, options=options)
        h = C.c_void_p()
        _check(lib().lbm_create(C.byref(d), C.byref(h)))
        self._h = h
        self.storage = storage
        try:
            _check(lib().lbm_set_geometry(self._h, cls.ctypes.data))
        except Exception:
            self.close()
            raise

    def close(self):
        if getattr(self, "_h", None):
            _check(lib().lbm_destroy(self._h))   # refused while the lattice belongs to a Batch
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_relaxation(self, tau=1.0, tau_even=1.0, tau_odd=1.0):
        _check(lib().lbm_set_relaxation(self._h, tau, tau_even, tau_odd))

    def set_geometry(self, node_class):
        """Installs a new class plane of the same dimensions (lbm_set_geometry):
        the populations are dropped (call init_equilibrium or
        upload_populations next); the launch shape is kept."""
        cls = np.ascontiguousarray(node_class, dtype=np.uint8)
        if cls.shape != (self.nc, self.nr):
            raise ValueError(f"class plane must be {(self.nc, self.nr)} (NC, NR), got {cls.shape}")
        _check(lib().lbm_set_geometry(self._h, cls.ctypes.data))

    def set_partial_bounce_back(self, tau, b=None):
        """Partially saturated bounce-back on the wet obstacles (FP64):
        the policy's tau and one weight per wet obstacle in list order
        (None = 1 for all); tau <= 0 switches it off."""
        arr = None if b is None else np.ascontiguousarray(b, dtype=np.float64)
        _check(lib().lbm_set_partial_bounce_back(self._h, float(tau), None if arr is None else _dp(arr)))

    def node_counts(self):
        out = np.zeros(7, dtype=np.int64)
        _check(lib().lbm_node_counts(self._h, out.ctypes.data_as(C.POINTER(C.c_int64))))
        return out

    def open_nodes(self, which):
        n = C.c_int64(0)
        _check(lib().lbm_open_nodes(self._h, which, None, C.byref(n)))
        ij = np.zeros((max(n.value, 1), 2), dtype=np.int64)
        _check(lib().lbm_open_nodes(self._h, which, ij.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(n)))
        return ij[: n.value]

    def set_open_values(self, which, ux=None, uy=None, rho=None):
        n = len(self.open_nodes(which))
        arrs = [None if a is None else np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float64), (n,)))
                for a in (ux, uy, rho)]
        _check(lib().lbm_set_open_values(self._h, which, *[None if a is None else _dp(a) for a in arrs]))

    def init_equilibrium(self):
        _check(lib().lbm_init_equilibrium(self._h))

    def upload_populations(self, p):
        p = np.ascontiguousarray(p, dtype=np.float64)
        assert p.shape == (9, self.nc, self.nr)
        _check(lib().lbm_upload_populations(self._h, _dp(p)))

    def populations(self):
        out = np.empty((9, self.nc, self.nr))
        _check(lib().lbm_download_populations(self._h, _dp(out)))
        return out

    def after_collision(self):
        out = np.empty((9, self.nc, self.nr))
        _check(lib().lbm_download_after_collision(self._h, _dp(out)))
        return out

    def step(self, n=1, flags=0):
        _check(lib().lbm_step(self._h, int(n), int(flags)))

    @property
    def time(self):
        t = C.c_int64(0)
        _check(lib().lbm_get_time(self._h, C.byref(t)))
        return t.value

    def macro(self):
        rho = np.empty((self.nc, self.nr))
        u = np.empty((2, self.nc, self.nr))
        _check(lib().lbm_download_macro(self._h, _dp(rho), _dp(u)))
        return rho, u

    def macro_async(self, rho, u):
        """Queues the download of the last LBM_STEP_MACRO step's fields into
        rho (NC, NR) and u (2, NC, NR) (float64, C-contiguous, kept alive by
        the caller until wait(ticket)); returns the ticket."""
        assert rho.shape == (self.nc, self.nr) and u.shape == (2, self.nc, self.nr)
        assert rho.flags.c_contiguous and u.flags.c_contiguous and rho.dtype == u.dtype == np.float64
        t = C.c_int64(0)
        _check(lib().lbm_download_macro_async(self._h, _dp(rho), _dp(u), C.byref(t)))
        return t.value

    def wait(self, ticket):
        _check(lib().lbm_wait(self._h, int(ticket)))

    def lift_drag(self, points=None, masks=None, per_point=False):
        ld = np.zeros(2)
        if points is None:
            n, ij_p, m_p = 0, None, None
            npts = int(self.node_counts()[6])
        else:
            pts = np.ascontiguousarray(points, dtype=np.int64)
            msk = np.ascontiguousarray(masks, dtype=np.int32)
            n = npts = len(pts)
            ij_p = pts.ctypes.data_as(C.POINTER(C.c_int64))
            m_p = msk.ctypes.data_as(C.POINTER(C.c_int32))
        pp = np.zeros((max(npts, 1), 2))
        _check(lib().lbm_lift_drag(self._h, ij_p, m_p, n, _dp(ld), _dp(pp) if per_point else None))
        if per_point:
            return ld[0], ld[1], pp[:npts]
        return ld[0], ld[1]

    def reduce(self):
        out = np.zeros(5)
        _check(lib().lbm_reduce(self._h, _dp(out)))
        return out

    def reduce_device(self, out_ptr):
        """lbm_reduce_device: queues the five sums into 5 doubles of device
        memory at out_ptr (no allocation, no wait)."""
        _check(lib().lbm_reduce_device(self._h, C.c_void_p(int(out_ptr))))

    def synchronize(self):
        _check(lib().lbm_synchronize(self._h))

    def set_launch(self, segments):
        _check(lib().lbm_set_launch(self._h, segments))

    def launch(self):
        r = C.c_int32(0)
        _check(lib().lbm_get_launch(self._h, C.byref(r)))
        return r.value

    def layout(self):
        """LAYOUT_BLOCKED or LAYOUT_PLANES (lbm_get_layout)."""
        r = C.c_int32(0)
        _check(lib().lbm_get_layout(self._h, C.byref(r)))
        return r.value

    def set_occupancy(self, groups_per_cu):
        """Work-groups per compute unit of the step kernels (2..7), 0 = no
        limit; until set or tuned, the storage format's default (occupancy())."""
        _check(lib().lbm_set_occupancy(self._h, groups_per_cu))

    def occupancy(self):
        r = C.c_int32(0)
        _check(lib().lbm_get_occupancy(self._h, C.byref(r)))
        return r.value

    def autotune(self, steps=5):
        """Fastest launch shape and occupancy (lbm_autotune); returns the shape
        and its ms per step (occupancy() gives the occupancy)."""
        r, ms = C.c_int32(0), C.c_double(0)
        _check(lib().lbm_autotune(self._h, steps, C.byref(r), C.byref(ms)))
        return r.value, ms.value

    def stream_handle(self):
        s = C.c_void_p()
        _check(lib().lbm_get_stream(self._h, C.byref(s)))
        return s.value

    @property
    def scale(self):
        s = C.c_double(0)
        _check(lib().lbm_get_scale(self._h, C.byref(s)))
        return s.value


# ---- ensembles -------------------------------------------------------------

class Batch:
    """An ensemble of Lattice objects advanced together (one lbm_batch).

    The lattices stay usable for downloads, reductions and new prescribed
    values, but are stepped only through the batch while it exists."""

    def __init__(self, lattices):
        self.lattices = list(lattices)
        arr = (C.c_void_p * len(self.lattices))(*[lat._h.value for lat in self.lattices])
        h = C.c_void_p()
        _check(lib().lbm_batch_create(arr, len(self.lattices), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().lbm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_launch(self, segments, blocks=0):
        _check(lib().lbm_batch_set_launch(self._h, segments, blocks))

    def launch(self):
        r, b = C.c_int32(0), C.c_int64(0)
        _check(lib().lbm_batch_get_launch(self._h, C.byref(r), C.byref(b)))
        return r.value, b.value

    def autotune(self, steps=5):
        r, b, ms = C.c_int32(0), C.c_int64(0), C.c_double(0)
        _check(lib().lbm_batch_autotune(self._h, steps, C.byref(r), C.byref(b), C.byref(ms)))
        return r.value, b.value, ms.value

    def step(self, n=1, flags=0):
        _check(lib().lbm_batch_step(self._h, n, flags))

    def reduce(self):
        out = np.zeros((len(self.lattices), 5))
        _check(lib().lbm_batch_reduce(self._h, _dp(out)))
        return out

    def info(self):
        w, t = C.c_int64(0), C.c_int64(0)
        _check(lib().lbm_batch_info(self._h, C.byref(w), C.byref(t)))
        return w.value, t.value

    def stream_handle(self):
        s = C.c_void_p()
        _check(lib().lbm_batch_get_stream(self._h, C.byref(s)))
        return s.value

    def synchronize(self):
        _check(lib().lbm_batch_synchronize(self._h))
