"""aifmm_ls_dense_matvec_noalloc: the dense product with caller device
buffers and the library's static scratch (no device allocation), against
aifmm_ls_dense_matvec (host buffers, scratch from an arena) on the same
inputs.  Two calls in a row, the second on a smaller problem, so the second
runs over the scratch contents the first left.  Both entry points run the
same kernels with the same leaf-chunk count at these sizes, so the products
are expected bitwise equal; the bar is 1e-14 relative."""
import ctypes
import os

import numpy as np
import pytest

import aifmm_ls_amd as A

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        path = next(p for p in ("/opt/rocm/lib/libamdhip64.so", "libamdhip64.so") if os.path.exists(p) or "/" not in p)
        _hip = ctypes.CDLL(path)
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return _hip


def device_copy(a):
    a = np.ascontiguousarray(a)
    p = ctypes.c_void_p()
    assert hip().hipMalloc(ctypes.byref(p), a.nbytes) == 0
    assert hip().hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
    return p


def noalloc_product(g, X):
    L = A.lib()
    nrhs = X.shape[1]
    xc = np.ascontiguousarray(X.T).astype(np.complex128)  # column-major N x nrhs
    dx = device_copy(xc)
    dy = device_copy(np.zeros_like(xc))
    try:
        rc = L.aifmm_ls_dense_matvec_noalloc(g.h, nrhs, dx, dy, None)
        if rc != 0:
            raise A.AifmmError(L.aifmm_ls_last_error().decode())
        out = np.zeros_like(xc)
        assert hip().hipMemcpy(out.ctypes.data, dy, out.nbytes, 2) == 0  # device to host
    finally:
        hip().hipFree(dx)
        hip().hipFree(dy)
    return out.T


def test_noalloc_product_matches_the_allocating_one():
    rng = np.random.default_rng(3)
    bumps = [[1.0, 0.15, 0.1, 0.2], [0.7, 0.1, -0.3, -0.2]]
    for L, kap, nrhs in ((4, 20.0, 2), (3, 30.0, 1)):  # the second call on a smaller input
        print(f"\n[no-alloc] L={L} kappa={kap} nrhs={nrhs}", flush=True)
        g = A.LSProblem(L, kap, 7, 6, bumps=bumps)
        X = rng.standard_normal((g.N, nrhs)) + 1j * rng.standard_normal((g.N, nrhs))
        Yref = g.dense_matvec(X)
        Y = noalloc_product(g, X)
        d = np.linalg.norm(Y - Yref) / np.linalg.norm(Yref)
        print(f"  no-alloc vs allocating: {d:.2e} ({'bitwise equal' if np.array_equal(Y, Yref) else 'rounding'})",
              flush=True)
        assert d < 1e-14
        g.close()
