"""lbm_reduce_device, the allocation-free variant of lbm_reduce: its five
sums, written to device memory in stream order from a fixed per-device
scratch array, are bit for bit those of lbm_reduce -- on two calls in a
row, the second on a smaller lattice (whose partials occupy less of the
shared scratch), in FP64 and FP32 storage; lattices beyond the scratch's
block count are refused.

The output buffer is allocated by the test through the HIP runtime the
library itself was loaded with (found in this process's mappings), so no
second runtime enters the process."""
import ctypes as C

import numpy as np
import pytest

import lbm_amd as L
import workloads as W

pytestmark = pytest.mark.gpu


class DeviceDoubles:
    """n doubles of device memory from the library's HIP runtime."""

    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            L.lib()
            with open("/proc/self/maps") as f:
                path = next(ln.split()[-1] for ln in f if "libamdhip64" in ln)
            h = C.CDLL(path)
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            cls._hip = h
        return cls._hip

    def __init__(self, n):
        self.n = n
        self.ptr = C.c_void_p()
        assert self.hip().hipMalloc(C.byref(self.ptr), 8 * n) == 0
        self.put(np.full(n, np.nan))

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert self.hip().hipMemcpy(self.ptr, a.ctypes.data, 8 * self.n, 1) == 0   # host to device

    def get(self):
        out = np.empty(self.n)
        assert self.hip().hipMemcpy(out.ctypes.data, self.ptr, 8 * self.n, 2) == 0  # device to host
        return out

    def free(self):
        if self.ptr:
            self.hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()


def _device_sums(lat, buf):
    buf.put(np.full(5, np.nan))
    lat.reduce_device(buf.ptr.value)
    lat.synchronize()
    return buf.get()


@pytest.mark.parametrize("storage", [L.F64, L.F32_SHIFTED])
def test_reduce_device_matches_reduce(storage):
    buf = DeviceDoubles(5)
    big = W.porous(300, seed=9, rmin=3, rmax=8, nc=420).make_lattice(storage)
    small = W.cavity(33).make_lattice(storage)
    big.step(40)
    small.step(25)
    for lat, name in ((big, "porous 300x420"), (small, "cavity 33x33")):
        want = lat.reduce()
        got = _device_sums(lat, buf)
        print(f"{name} storage {storage}: lbm_reduce {want.tolist()} lbm_reduce_device {got.tolist()}", flush=True)
        assert np.array_equal(got, want), name
    # the collide-only state right after init (no gather) as well
    fresh = W.cavity(33).make_lattice(storage)
    assert np.array_equal(_device_sums(fresh, buf), fresh.reduce())
    for lat in (big, small, fresh):
        lat.close()
    buf.free()


def test_reduce_device_refuses_lattices_beyond_the_scratch():
    # ceil(300 / 256) * 65600 = 131200 reduction blocks > 131072
    cls = np.zeros((65600, 300), dtype=np.uint8)
    cls[:, 0] = cls[:, -1] = L.BOUNDARY
    cls[0, :] = cls[-1, :] = L.BOUNDARY
    lat = L.Lattice(cls, storage=L.F32_SHIFTED)
    lat.init_equilibrium()
    buf = DeviceDoubles(5)
    with pytest.raises(L.LbmError) as e:
        lat.reduce_device(buf.ptr.value)
    assert e.value.code == L.ERR_UNSUPPORTED
    lat.close()
    buf.free()
