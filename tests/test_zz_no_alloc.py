"""The allocation-free resampling entry point (dust2_amd_resample_noalloc)
against the allocating one (dust2_amd_resample) on the same weights: the
same ancestors and the same [increment, resampled, fallback] bit for bit,
on calls in a row (the second with fewer weights, so stale contents of the
shared scratch would show), through the multi-workgroup path and its
sequential fallback, and the single-workgroup path; sizes beyond the
scratch are refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def hip():
    h = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


def allocating(lib, w, u):
    from dust2_amd import _lib
    idx = np.zeros(w.size, dtype=np.uint32)
    out3 = np.zeros(3)
    _lib.call("dust2_amd_resample", _lib.dptr(w), w.size, float(u), _lib.u32ptr(idx),
              _lib.dptr(out3))
    return idx, out3


class DeviceBuffers:
    """Caller-owned device buffers (allocated by the test, outside the call)."""

    def __init__(self, hip, n_max):
        self.hip = hip
        self.w, self.idx, self.out3 = C.c_void_p(), C.c_void_p(), C.c_void_p()
        for p, nb in [(self.w, 8 * n_max), (self.idx, 4 * n_max), (self.out3, 24)]:
            assert hip.hipMalloc(C.byref(p), nb) == 0

    def free(self):
        for p in [self.w, self.idx, self.out3]:
            self.hip.hipFree(p)


def no_alloc(hip, bufs, w, u):
    from dust2_amd import _lib
    w = np.ascontiguousarray(w, dtype=np.float64)
    assert hip.hipMemcpy(bufs.w, w.ctypes.data, w.nbytes, H2D) == 0
    _lib.call("dust2_amd_resample_noalloc", bufs.w, w.size, float(u), bufs.idx, bufs.out3,
              None)
    idx = np.zeros(w.size, dtype=np.uint32)
    out3 = np.zeros(3)
    assert hip.hipMemcpy(idx.ctypes.data, bufs.idx, idx.nbytes, D2H) == 0
    assert hip.hipMemcpy(out3.ctypes.data, bufs.out3, 24, D2H) == 0
    return idx, out3


def test_no_alloc_equals_allocating_entry_point(gpu_lib, hip):
    rng = np.random.default_rng(7)
    bufs = DeviceBuffers(hip, 1 << 20)
    try:
        cases = [
            (rng.normal(-50, 3, 300000), 0.37),       # multi-workgroup
            (rng.normal(-50, 3, 70000), 0.81),        # second call, smaller
            (np.zeros(40000) - 2.5, 0.0),             # equal weights: points on boundaries,
                                                      # the sequential fallback
            (rng.normal(0, 1, 1500), 0.25),           # one workgroup
            (rng.normal(-5, 2, 1 << 20), 0.5),        # the largest accepted size
        ]
        cases[1][0][::97] = -np.inf
        cases[0][0][5] = np.nan
        for w, u in cases:
            i_ref, o_ref = allocating(gpu_lib, w, u)
            i_got, o_got = no_alloc(hip, bufs, w, u)
            assert np.array_equal(i_got, i_ref), w.size
            assert np.array_equal(o_got, o_ref), (w.size, o_got, o_ref)
        assert allocating(gpu_lib, cases[2][0], 0.0)[1][2] == 1   # fallback taken
    finally:
        bufs.free()


def test_no_alloc_refuses_sizes_beyond_scratch(gpu_lib):
    from dust2_amd import _lib
    with pytest.raises(_lib.DustError, match="allocation-free scratch holds"):
        _lib.call("dust2_amd_resample_noalloc", C.c_void_p(1), (1 << 20) + 1, 0.5,
                  C.c_void_p(1), C.c_void_p(1), None)
