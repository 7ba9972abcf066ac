"""The allocation-free multi-dot (N_VDotProdMultiNoAlloc_Hip) against the
allocating one (N_VDotProdMulti_Hip), through the C ABI: bitwise equal
results on two calls in a row, the second with shorter vectors and fewer
products (the shared device scratch keeps the first call's partials; the
second must not read them), and the size limit refused without a launch."""
import ctypes as C
import math

import numpy as np
import pytest

import llg_amd as la

pytestmark = pytest.mark.gpu


def _vecs(n, k, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(n)
    Y = [rng.standard_normal(n) for _ in range(k)]
    return x, Y


def _arr(vs):
    return (C.c_void_p * len(vs))(*[v.v for v in vs])


def test_noalloc_multidot_equals_allocating_one():
    L = la.lib()
    for n, k, seed in ((2_359_296, 16, 1), (100_003, 5, 2), (7, 3, 3)):
        x, Y = _vecs(n, k, seed)
        vx = la.HipVector(n, None, x)
        vY = [la.HipVector(n, None, a) for a in Y]
        d_ref, d_na = (C.c_double * k)(), (C.c_double * k)()
        assert L.N_VDotProdMulti_Hip(k, vx.v, _arr(vY), d_ref) == 0
        assert L.N_VDotProdMultiNoAlloc_Hip(k, vx.v, _arr(vY), d_na) == 0
        print(f"[gpu] no-alloc multi-dot n={n} k={k}: "
              f"{'bit-equal' if list(d_na) == list(d_ref) else 'DIFFERENT'}", flush=True)
        assert list(d_na) == list(d_ref)
        for j in range(k):
            t = x * Y[j]
            assert abs(d_na[j] - math.fsum(t)) <= 1e-14 * math.fsum(np.abs(t))


def test_noalloc_multidot_single_product_and_limits():
    L = la.lib()
    n = 50_000
    x, Y = _vecs(n, 17, 4)
    vx = la.HipVector(n, None, x)
    vY = [la.HipVector(n, None, a) for a in Y]
    d = (C.c_double * 17)()
    assert L.N_VDotProdMultiNoAlloc_Hip(1, vx.v, _arr(vY[:1]), d) == 0
    t = x * Y[0]
    assert abs(d[0] - math.fsum(t)) <= 1e-14 * math.fsum(np.abs(t))
    d[0] = 12345.0
    assert L.N_VDotProdMultiNoAlloc_Hip(17, vx.v, _arr(vY), d) == -1
    assert d[0] == 12345.0
    short = la.HipVector(n - 1, None, x[:-1])
    assert L.N_VDotProdMultiNoAlloc_Hip(2, vx.v, _arr([vY[0], short]), d) == -1
    assert L.N_VDotProdMultiNoAlloc_Hip(0, vx.v, _arr(vY[:1]), d) == 0
