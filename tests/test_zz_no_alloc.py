"""rpd_knn_device_noalloc, the entry point that allocates nothing, against
rpd_knn (the same kernels with context buffers): bit-identical permutation
and neighbour lists, on two calls in a row (the second with fewer seeds, so
the shared scratch keeps the first call's contents), then a size above its
limit refused with RPD_E_LIMIT."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime librpd is linked with (the test's own buffers; the
    entry point under test allocates nothing)."""
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return hip


def _noalloc(rpd, ctx, x, K, prec):
    hip = _hip()
    n = x.shape[0]
    x = np.ascontiguousarray(x, np.float64)
    perm = np.full(n, -7, np.int32)
    knn = np.full(n * K, -7, np.int32)
    ptrs = []
    for a in (x, perm, knn):
        p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        ptrs.append(p)
    f = rpd.lib().rpd_knn_device_noalloc
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                  ctypes.c_void_p, ctypes.c_void_p]
    rc = f(ctx._h, ptrs[0], n, K, prec, ptrs[1], ptrs[2])
    for a, p in ((perm, ptrs[1]), (knn, ptrs[2])):
        assert hip.hipMemcpy(a.ctypes.data, p, a.nbytes, 2) == 0  # device to host
    for p in ptrs:
        hip.hipFree(p)
    return rc, perm, knn.reshape(n, K)


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_knn_matches_rpd_knn(rpd, ctx, prec):
    for n, K, seed in ((200000, 64, 3), (37000, 20, 4)):
        h = (1.0 / n) ** (1 / 3)
        x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=seed, wmax=0.25 * h * h * 1e6)
        rc, perm, knn = _noalloc(rpd, ctx, x, K, prec)
        assert rc == 0, rc
        ctx.set_seeds(x)
        want = ctx.knn(K, precision=prec)
        print("noalloc knn", prec, n, K, flush=True)
        assert np.array_equal(perm, ctx.perm())
        assert np.array_equal(knn, want)
    # above the limit: refused, nothing written
    n = (1 << 18) + 1
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=5)
    rc, perm, knn = _noalloc(rpd, ctx, x, 8, prec)
    assert rc == -6  # RPD_E_LIMIT
    assert (perm == -7).all() and (knn == -7).all()
