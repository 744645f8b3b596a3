"""dem_sq_net_gradient_noalloc: the allocation-free variant of
dem_sq_net_gradient (device pointers in and out, partial sums in a fixed
per-device workspace) against the allocating entry point on the same network
and samples, through the C ABI, on two calls in a row, the second with a
smaller batch; plus its argument checks (batch above the workspace limit,
uninitialised network)."""
import ctypes

import numpy as np
import pytest

import pydem

pytestmark = pytest.mark.gpu


def _inputs(rng, n):
    X = rng.standard_normal((n, 18)).astype(np.float32)
    T = (0.3 * rng.standard_normal((n, 3))).astype(np.float32)
    return X, T


class Hip:
    """Device buffers through the HIP runtime the library itself links (the
    test's inputs and outputs; the entry point under test allocates nothing)."""

    def __init__(self):
        self.rt = ctypes.CDLL("libamdhip64.so", mode=ctypes.RTLD_GLOBAL)
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        self.bufs.append(p)
        assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # hipMemcpyHostToDevice
        return p.value

    def get(self, ptr, like):
        out = np.empty_like(like)
        assert self.rt.hipMemcpy(out.ctypes.data, ctypes.c_void_p(ptr), out.nbytes, 2) == 0  # DeviceToHost
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _device_gradient(s, hip, X, T, sentinel):
    dX = hip.put(X)
    dT = hip.put(T)
    # outputs start as a sentinel, so a result that was not written shows
    g0 = np.full(pydem.SQ_NUM_PARAMS, sentinel, dtype=np.float32)
    l0 = np.full(1, sentinel, dtype=np.float32)
    dg = hip.put(g0)
    dl = hip.put(l0)
    rc = s.net_gradient_noalloc(X.shape[0], dX, dT, dg, dl)
    assert rc == 0, s.lib.dem_sq_last_error(s.h).decode()
    g, l = hip.get(dg, g0), hip.get(dl, l0)
    hip.free()
    return g, float(l[0])


def test_noalloc_gradient_equals_allocating_entry_point():
    s = pydem.SqSolver(0)
    hip = Hip()
    try:
        s.net_init(7)
        rng = np.random.default_rng(11)
        # first call near the workspace limit, the second smaller: the second
        # must not see the first call's partial sums
        for n, sentinel in ((pydem.DEM_SQ_GRAD_NOALLOC_MAX - 37, 1e30), (1000, -7.0)):
            X, T = _inputs(rng, n)
            g_ref, l_ref = s.net_gradient(X, T)
            g, l = _device_gradient(s, hip, X, T, sentinel)
            print(f"n={n}: max |grad - ref| {np.abs(g - g_ref).max():.3e} (ref scale "
                  f"{np.abs(g_ref).max():.3e}), loss {l:.6e} vs {l_ref:.6e}", flush=True)
            # same kernels, same block partition and fixed-order sums: bitwise
            assert np.array_equal(g, g_ref)
            assert l == l_ref
    finally:
        s.close()


def test_noalloc_gradient_rejects_oversized_batch_and_missing_network():
    s = pydem.SqSolver(0)
    hip = Hip()
    try:
        p = hip.put(np.zeros(64, dtype=np.float32))
        # no network yet
        assert s.net_gradient_noalloc(1, p, p, p, 0) == pydem.DEM_ERR_STATE
        s.net_init(3)
        # beyond the workspace: refused before any launch
        rc = s.net_gradient_noalloc(pydem.DEM_SQ_GRAD_NOALLOC_MAX + 1, p, p, p, 0)
        assert rc == pydem.DEM_ERR_ARG
        assert "workspace" in s.lib.dem_sq_last_error(s.h).decode()
        assert s.net_gradient_noalloc(0, p, p, p, 0) == pydem.DEM_ERR_ARG
    finally:
        hip.free()
        s.close()
