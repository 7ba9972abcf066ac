"""GPU parity of the HIP N_Vector ops table (standard, fused and
vector-array ops, the integrator's one-pass Nordsieck / error-weight
kernels), called through the C ABI, against exact host references.

Semantics follow SUNDIALS' N_Vector contract as the reference uses it
(NVECTOR_CUDA on every variant; the fused multi-dot and linear combination
of src/irregular/2d_fft/2d_i4/deferred_nvector.cu:95-180 and
src/2d/2d_heter/deferred_nvector.cu:135-151,209-236).

Tolerances: products, quotients and copies within 1 ulp of the host
value; a*x+b*y forms within 2 ulp of |a*x|+|b*y| (the device contracts
them into one fused multiply-add, as the reference's nvcc build does);
norms within 1e-13 relative of a correctly rounded host sum (math.fsum),
signed dot products within 1e-14 of the sum of |terms|; min/max/compare/
mask ops exact.
Reductions must also be bitwise repeatable for a fixed launch shape.
"""
import ctypes as C
import math

import numpy as np
import pytest

import llg_amd as la

pytestmark = pytest.mark.gpu

SIZES = [1, 7, 1000, 262147, 2_359_296]   # the last is 3 x 1536 x 512
RTOL_RED = 1e-13


def ulp_close(a, b, nulp=2):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    tol = nulp * np.spacing(np.maximum(np.abs(a), np.abs(b))) + 1e-300
    return bool(np.all(np.abs(a - b) <= tol))


def sum_close(got, ref, *terms):
    """a*x + b*y style results: a fused multiply-add rounds once where the
    host rounds twice, so the bound scales with the terms, not the result."""
    mag = sum(np.abs(t) for t in terms)
    return bool(np.all(np.abs(np.asarray(got) - ref) <= 2 * np.spacing(mag) + 1e-300))


def vec(x):
    return la.HipVector(len(x), None, x)


def arr(vs):
    return (C.c_void_p * len(vs))(*[v.v for v in vs])


def fsum_rel(got, ref):
    """sums of non-negative terms (norms): relative bound"""
    return abs(got - ref) <= RTOL_RED * max(abs(ref), 1e-300)


def dot_close(got, a, b):
    """signed sums cancel: bound by the sum of |terms| (a summation tree of
    depth d errs by at most d * 2^-53 * sum|terms|; d < 64 here)"""
    t = a * b
    return abs(got - math.fsum(t)) <= 1e-14 * math.fsum(np.abs(t))


@pytest.fixture(autouse=True)
def _default_reduce_config():
    L = la.lib()
    yield
    L.NVHip_SetReduceConfig(1024, 1)   # the library default


@pytest.mark.parametrize("n", SIZES)
def test_elementwise_ops(n):
    L = la.lib()
    rng = np.random.default_rng(n)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    y[y == 0] = 1.0
    vx, vy, vz = vec(x), vec(y), vec(np.zeros(n))
    L.N_VLinearSum_Hip(0.7, vx.v, -1.3, vy.v, vz.v)
    assert sum_close(vz.get(), 0.7 * x - 1.3 * y, 0.7 * x, 1.3 * y)
    L.N_VLinearSum_Hip(1.0, vx.v, 1.0, vy.v, vz.v)
    assert ulp_close(vz.get(), x + y)
    L.N_VConst_Hip(2.5, vz.v)
    assert np.all(vz.get() == 2.5)
    L.N_VProd_Hip(vx.v, vy.v, vz.v)
    assert np.array_equal(vz.get(), x * y)
    L.N_VDiv_Hip(vx.v, vy.v, vz.v)
    assert ulp_close(vz.get(), x / y, 1)
    L.N_VScale_Hip(-3.0, vx.v, vz.v)
    assert np.array_equal(vz.get(), -3.0 * x)
    L.N_VAbs_Hip(vx.v, vz.v)
    assert np.array_equal(vz.get(), np.abs(x))
    L.N_VInv_Hip(vy.v, vz.v)
    assert ulp_close(vz.get(), 1.0 / y, 1)
    L.N_VAddConst_Hip(vx.v, 0.25, vz.v)
    assert np.array_equal(vz.get(), x + 0.25)
    L.N_VCompare_Hip(0.5, vx.v, vz.v)
    assert np.array_equal(vz.get(), (np.abs(x) >= 0.5).astype(float))
    # in place (z aliases x)
    L.N_VLinearSum_Hip(2.0, vx.v, 1.0, vy.v, vx.v)
    assert sum_close(vx.get(), 2.0 * x + y, 2.0 * x, y)


@pytest.mark.parametrize("n", SIZES)
def test_reductions(n):
    L = la.lib()
    rng = np.random.default_rng(100 + n)
    x, w = rng.standard_normal(n), rng.uniform(0.5, 2.0, n)
    idm = (rng.uniform(size=n) > 0.3).astype(float)
    vx, vw, vid = vec(x), vec(w), vec(idm)
    assert dot_close(L.N_VDotProd_Hip(vx.v, vw.v), x, w)
    assert L.N_VMaxNorm_Hip(vx.v) == np.max(np.abs(x))
    assert L.N_VMin_Hip(vx.v) == np.min(x)
    wsq = (x * w) ** 2
    assert fsum_rel(L.N_VWSqrSumLocal_Hip(vx.v, vw.v), math.fsum(wsq))
    assert fsum_rel(L.N_VWSqrSumMaskLocal_Hip(vx.v, vw.v, vid.v), math.fsum(wsq[idm > 0]))
    assert fsum_rel(L.N_VWrmsNorm_Hip(vx.v, vw.v), math.sqrt(math.fsum(wsq) / n))
    assert fsum_rel(L.N_VWrmsNormMask_Hip(vx.v, vw.v, vid.v),
                    math.sqrt(math.fsum(wsq[idm > 0]) / n))
    assert fsum_rel(L.N_VWL2Norm_Hip(vx.v, vw.v), math.sqrt(math.fsum(wsq)))
    assert fsum_rel(L.N_VL1Norm_Hip(vx.v), math.fsum(np.abs(x)))


def test_invtest_constrmask_minquotient():
    L = la.lib()
    n = 5000
    rng = np.random.default_rng(3)
    x = rng.standard_normal(n)
    vx, vz = vec(x), vec(np.zeros(n))
    assert L.N_VInvTest_Hip(vx.v, vz.v) == 1
    assert ulp_close(vz.get(), 1.0 / x, 1)
    x0 = x.copy()
    x0[17] = 0.0
    vx0 = vec(x0)
    assert L.N_VInvTest_Hip(vx0.v, vz.v) == 0
    # constraints: 2 -> x > 0, 1 -> x >= 0, -1 -> x <= 0, -2 -> x < 0, 0 -> none
    c = rng.choice([-2.0, -1.0, 0.0, 1.0, 2.0], n)
    xx = rng.choice([-1.0, 0.0, 1.0], n)
    bad = ((c == 2) & (xx <= 0)) | ((c == 1) & (xx < 0)) | ((c == -1) & (xx > 0)) | \
          ((c == -2) & (xx >= 0))
    vc, vxx, vm = vec(c), vec(xx), vec(np.zeros(n))
    ok = L.N_VConstrMask_Hip(vc.v, vxx.v, vm.v)
    assert ok == (0 if bad.any() else 1)
    assert np.array_equal(vm.get(), bad.astype(float))
    num, den = rng.standard_normal(n), rng.standard_normal(n)
    den[::3] = 0.0
    # keep the handles alive across the calls (a temporary's .v would dangle)
    vnum, vden, vzero = vec(num), vec(den), vec(np.zeros(n))
    q = L.N_VMinQuotient_Hip(vnum.v, vden.v)
    assert q == np.min(num[den != 0] / den[den != 0])
    assert L.N_VMinQuotient_Hip(vnum.v, vzero.v) == np.finfo(float).max


@pytest.mark.parametrize("k", [1, 2, 3, 5, 6, 16, 19])
def test_fused_lincomb_scaleaddmulti_dotmulti(k):
    L = la.lib()
    n = 100_003
    rng = np.random.default_rng(k)
    X = [rng.standard_normal(n) for _ in range(k)]
    c = rng.standard_normal(k)
    vX = [vec(a) for a in X]
    vz = vec(np.zeros(n))
    cc = (C.c_double * k)(*c)
    assert L.N_VLinearCombination_Hip(k, cc, arr(vX), vz.v) == 0
    ref = sum(c[j] * X[j] for j in range(k))
    assert np.max(np.abs(vz.get() - ref)) <= 1e-14 * k * np.max(np.abs(ref))
    # in place: z == X[0] (the SUNDIALS contract allows it)
    if k > 1:
        assert L.N_VLinearCombination_Hip(k, cc, arr(vX), vX[0].v) == 0
        assert np.max(np.abs(vX[0].get() - ref)) <= 1e-14 * k * np.max(np.abs(ref))
        vX[0].set(X[0])
    x = rng.standard_normal(n)
    vx = vec(x)
    vZ = [vec(np.zeros(n)) for _ in range(k)]
    assert L.N_VScaleAddMulti_Hip(k, cc, vx.v, arr(vX), arr(vZ)) == 0
    for j in range(k):
        assert sum_close(vZ[j].get(), c[j] * x + X[j], c[j] * x, X[j])
    d = (C.c_double * k)()
    assert L.N_VDotProdMulti_Hip(k, vx.v, arr(vX), d) == 0
    for j in range(k):
        assert dot_close(d[j], x, X[j])


@pytest.mark.parametrize("k", [1, 4, 17])
def test_vector_array_ops(k):
    L = la.lib()
    n = 50_001
    rng = np.random.default_rng(50 + k)
    X = [rng.standard_normal(n) for _ in range(k)]
    Y = [rng.standard_normal(n) for _ in range(k)]
    W = [rng.uniform(0.5, 2, n) for _ in range(k)]
    vX, vY, vW = [vec(a) for a in X], [vec(a) for a in Y], [vec(a) for a in W]
    vZ = [vec(np.zeros(n)) for _ in range(k)]
    assert L.N_VLinearSumVectorArray_Hip(k, 0.5, arr(vX), -2.0, arr(vY), arr(vZ)) == 0
    for j in range(k):
        assert sum_close(vZ[j].get(), 0.5 * X[j] - 2.0 * Y[j], 0.5 * X[j], 2.0 * Y[j])
    c = rng.standard_normal(k)
    assert L.N_VScaleVectorArray_Hip(k, (C.c_double * k)(*c), arr(vX), arr(vZ)) == 0
    for j in range(k):
        assert np.array_equal(vZ[j].get(), c[j] * X[j])
    assert L.N_VConstVectorArray_Hip(k, -1.5, arr(vZ)) == 0
    for j in range(k):
        assert np.all(vZ[j].get() == -1.5)
    nrm = (C.c_double * k)()
    assert L.N_VWrmsNormVectorArray_Hip(k, arr(vX), arr(vW), nrm) == 0
    for j in range(k):
        assert fsum_rel(nrm[j], math.sqrt(math.fsum((X[j] * W[j]) ** 2) / n))
    idm = (rng.uniform(size=n) > 0.5).astype(float)
    vid = vec(idm)
    assert L.N_VWrmsNormMaskVectorArray_Hip(k, arr(vX), arr(vW), vid.v, nrm) == 0
    for j in range(k):
        ref = math.sqrt(math.fsum(((X[j] * W[j]) ** 2)[idm > 0]) / n)
        assert fsum_rel(nrm[j], ref)


@pytest.mark.parametrize("q", [1, 2, 3, 4, 5])
def test_pascal_matches_linear_sum_sequence(q):
    """NVHip_Pascal(q, zn, +1) == CVODE's predict loop of N_VLinearSum;
    sign -1 undoes it (cvRestore)."""
    L = la.lib()
    n = 30_001
    rng = np.random.default_rng(q)
    Z = [rng.standard_normal(n) for _ in range(q + 1)]
    vZ = [vec(a) for a in Z]
    assert L.NVHip_Pascal(q, arr(vZ), 1.0) == 0
    R = [a.copy() for a in Z]
    for k in range(1, q + 1):
        for j in range(q, k - 1, -1):
            R[j - 1] = R[j - 1] + R[j]
    for j in range(q + 1):
        assert np.array_equal(vZ[j].get(), R[j])
    assert L.NVHip_Pascal(q, arr(vZ), -1.0) == 0
    for j in range(q + 1):
        assert np.max(np.abs(vZ[j].get() - Z[j])) <= 1e-12 * (2 ** q)


def test_ewt_matches_abs_scale_inv():
    L = la.lib()
    n = 70_001
    rng = np.random.default_rng(9)
    y, av = rng.standard_normal(n), rng.uniform(1e-6, 1e-3, n)
    vy, vav, vw = vec(y), vec(av), vec(np.zeros(n))
    assert L.NVHip_EwtSet(vy.v, 1e-4, 1e-5, None, vw.v) == 0
    assert ulp_close(vw.get(), 1.0 / (1e-4 * np.abs(y) + 1e-5), 3)
    assert L.NVHip_EwtSet(vy.v, 1e-4, 0.0, vav.v, vw.v) == 0
    assert ulp_close(vw.get(), 1.0 / (1e-4 * np.abs(y) + av), 3)


@pytest.mark.parametrize("blocks,mode", [(512, 1), (1, 1), (64, 0), (2048, 0), (4096, 1)])
def test_reduction_launch_shapes_agree_and_repeat(blocks, mode):
    L = la.lib()
    assert L.NVHip_SetReduceConfig(blocks, mode) == 0
    n = 2_359_296
    rng = np.random.default_rng(77)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    vx, vy = vec(x), vec(y)
    a = L.N_VDotProd_Hip(vx.v, vy.v)
    b = L.N_VDotProd_Hip(vx.v, vy.v)
    assert a == b
    assert dot_close(a, x, y)
    d = (C.c_double * 3)()
    assert L.N_VDotProdMulti_Hip(3, vx.v, arr([vx, vy, vx]), d) == 0
    assert dot_close(d[1], x, y)
    assert fsum_rel(d[0], math.fsum(x * x))
    assert L.NVHip_SetReduceConfig(0, 1) == -1
    assert L.NVHip_SetReduceConfig(512, 2) == -1


# ---------------------------------------------------------------------------
# One-pass kernels for the integrator: each must reproduce, bit for bit, the
# standard op sequence it replaces (same arithmetic, same reduction order).

N_FUSED = 2_359_296


def _vecs(rng, k, n=N_FUSED, lo=None):
    if lo is None:
        return [vec(rng.standard_normal(n)) for _ in range(k)]
    return [vec(rng.uniform(lo, 2.0, n)) for _ in range(k)]


def test_scalediv_equals_scale_then_div():
    L = la.lib()
    rng = np.random.default_rng(1)
    x, = _vecs(rng, 1)
    s, = _vecs(rng, 1, lo=0.5)
    xa, va = vec(x.get()), vec(np.zeros(N_FUSED))
    L.N_VScale_Hip(0.37, xa.v, xa.v)
    L.N_VDiv_Hip(xa.v, s.v, va.v)
    xb, vb = vec(x.get()), vec(np.zeros(N_FUSED))
    assert L.NVHip_ScaleDiv(0.37, xb.v, s.v, xb.v, vb.v) == 0
    assert np.array_equal(xa.get(), xb.get()) and np.array_equal(va.get(), vb.get())


@pytest.mark.parametrize("k", [1, 2, 4, 6])
def test_proddotmulti_equals_prod_then_dotmulti(k):
    L = la.lib()
    rng = np.random.default_rng(10 + k)
    s, = _vecs(rng, 1, lo=0.5)
    t, = _vecs(rng, 1)
    Y = _vecs(rng, k - 1)
    za, zb = vec(np.zeros(N_FUSED)), vec(np.zeros(N_FUSED))
    L.N_VProd_Hip(s.v, t.v, za.v)
    da = (C.c_double * k)()
    assert L.N_VDotProdMulti_Hip(k, za.v, arr(Y + [za]), da) == 0
    db = (C.c_double * k)()
    assert L.NVHip_ProdDotMulti(s.v, t.v, zb.v, k, arr(Y + [zb]), db) == 0
    assert np.array_equal(za.get(), zb.get())
    assert list(da) == list(db)


@pytest.mark.parametrize("k", [2, 3, 6])
def test_gsproject_equals_cgs_sequence(k):
    """prod, multi-dot, combination with the negated dots, norm -- one wait"""
    L = la.lib()
    rng = np.random.default_rng(20 + k)
    s, = _vecs(rng, 1, lo=0.5)
    t, = _vecs(rng, 1)
    Y = _vecs(rng, k - 1)
    za = vec(np.zeros(N_FUSED))
    L.N_VProd_Hip(s.v, t.v, za.v)
    da = (C.c_double * k)()
    L.N_VDotProdMulti_Hip(k, za.v, arr(Y + [za]), da)
    c = (C.c_double * k)(1.0, *[-da[j] for j in range(k - 1)])
    L.N_VLinearCombination_Hip(k, c, arr([za] + Y), za.v)
    zza = L.N_VDotProd_Hip(za.v, za.v)
    zb = vec(np.zeros(N_FUSED))
    db, zzb = (C.c_double * k)(), C.c_double()
    assert L.NVHip_GSProject(s.v, t.v, zb.v, k, arr(Y + [zb]), db, C.byref(zzb)) == 0
    assert list(da) == list(db)
    assert zza == zzb.value
    assert np.array_equal(za.get(), zb.get())
    # the last Y must be z itself
    assert L.NVHip_GSProject(s.v, t.v, zb.v, k, arr(Y + [s]), db, C.byref(zzb)) == -1


@pytest.mark.parametrize("k", [1, 3, 5])
def test_lincombdot_and_lincombdiv(k):
    L = la.lib()
    rng = np.random.default_rng(30 + k)
    X = _vecs(rng, k)
    s, = _vecs(rng, 1, lo=0.5)
    cf = rng.standard_normal(k)
    c = (C.c_double * k)(*cf)
    za, zb = vec(np.zeros(N_FUSED)), vec(np.zeros(N_FUSED))
    L.N_VLinearCombination_Hip(k, c, arr(X), za.v)
    zza = L.N_VDotProd_Hip(za.v, za.v)
    zzb = C.c_double()
    assert L.NVHip_LinCombDot(k, c, arr(X), zb.v, C.byref(zzb)) == 0
    assert np.array_equal(za.get(), zb.get()) and zza == zzb.value
    # x = (0 + sum c_j X_j) / s  ==  N_VConst 0, N_VLinearCombination, N_VDiv
    xc, xa, xb = vec(np.zeros(N_FUSED)), vec(np.zeros(N_FUSED)), vec(np.zeros(N_FUSED))
    c1 = (C.c_double * (k + 1))(1.0, *cf)
    L.N_VLinearCombination_Hip(k + 1, c1, arr([xc] + X), xc.v)
    L.N_VDiv_Hip(xc.v, s.v, xa.v)
    assert L.NVHip_LinCombDiv(k, c, arr(X), s.v, xb.v) == 0
    assert np.array_equal(xa.get(), xb.get())


@pytest.mark.parametrize("it", [0, 1])
def test_newton_update_equals_linear_sums_and_norms(it):
    L = la.lib()
    rng = np.random.default_rng(40 + it)
    ycor0 = rng.standard_normal(N_FUSED) * (0 if it == 0 else 1e-3)
    dl, zn0 = _vecs(rng, 2)
    w, = _vecs(rng, 1, lo=0.5)
    ya, yb = vec(ycor0), vec(ycor0)
    outa, outb = vec(np.zeros(N_FUSED)), vec(np.zeros(N_FUSED))
    L.N_VLinearSum_Hip(1.0, ya.v, 1.0, dl.v, ya.v)
    L.N_VLinearSum_Hip(1.0, zn0.v, 1.0, ya.v, outa.v)
    if it == 0:
        dna, yna = L.N_VWrmsNorm_Hip(dl.v, w.v), None
    else:
        nr = (C.c_double * 2)()
        L.N_VWrmsNormVectorArray_Hip(2, arr([dl, ya]), arr([w, w]), nr)
        dna, yna = nr[0], nr[1]
    dn, yn = C.c_double(), C.c_double(-1.0)
    assert L.NVHip_NewtonUpdate(it, yb.v, dl.v, zn0.v, w.v, outb.v, C.byref(dn), C.byref(yn)) == 0
    assert np.array_equal(ya.get(), yb.get()) and np.array_equal(outa.get(), outb.get())
    assert dn.value == dna
    if it > 0:
        assert yn.value == yna


@pytest.mark.parametrize("q,weights", [(1, None), (3, "scalar"), (5, "vector")])
def test_predict_init_equals_pascal_const_linear_sum(q, weights):
    """NVHip_PredictInit == NVHip_EwtSet on the old zn[0] (when asked),
    NVHip_Pascal(+1), N_VConst(0, acor) and N_VLinearSum(1, zn[0], 1, acor,
    y), bit for bit (a -0 in zn[0] gives +0 in y, as the sum does)"""
    L = la.lib()
    n = 30_001
    rng = np.random.default_rng(60 + q)
    Z = [rng.standard_normal(n) for _ in range(q + 1)]
    Z[0][:7] = -0.0
    for j in range(1, q + 1):
        Z[j][:7] = 0.0
    za, zb = [vec(a) for a in Z], [vec(a) for a in Z]
    acor_a, y_a = vec(rng.standard_normal(n)), vec(rng.standard_normal(n))
    acor_b, y_b = vec(rng.standard_normal(n)), vec(rng.standard_normal(n))
    av = vec(rng.uniform(1e-6, 1e-3, n)) if weights == "vector" else None
    w_a, w_b = vec(np.zeros(n)), vec(np.zeros(n))
    rtol, atol = 1e-4, (0.0 if weights == "vector" else 1e-5)
    if weights:
        assert L.NVHip_EwtSet(za[0].v, rtol, atol, av.v if av else None, w_a.v) == 0
    assert L.NVHip_Pascal(q, arr(za), 1.0) == 0
    L.N_VConst_Hip(0.0, acor_a.v)
    L.N_VLinearSum_Hip(1.0, za[0].v, 1.0, acor_a.v, y_a.v)
    assert L.NVHip_PredictInit(q, arr(zb), acor_b.v, y_b.v, rtol, atol, av.v if av else None,
                               w_b.v if weights else None) == 0
    for j in range(q + 1):
        assert np.array_equal(za[j].get(), zb[j].get()), j
    assert np.array_equal(acor_a.get(), acor_b.get())
    ya, yb = y_a.get(), y_b.get()
    assert np.array_equal(ya, yb) and np.array_equal(np.signbit(ya), np.signbit(yb))
    if weights:
        assert np.array_equal(w_a.get(), w_b.get())
    assert L.NVHip_PredictInit(q, arr(zb), zb[0].v, y_b.v, rtol, atol, None,
                               None) == -1   # acor aliases zn[0]


@pytest.mark.parametrize("q,weights,scaled", [(1, None, False), (2, "scalar", True),
                                              (4, "vector", False), (5, "scalar", True)])
def test_predict_complete_equals_completion_rescale_then_predict_init(q, weights, scaled):
    """NVHip_PredictComplete == N_VScaleAddMulti(q+1, l, acor, zn, zn), then
    (scaled) N_VScaleVectorArray(q, c[1..q], zn[1..q]), then NVHip_PredictInit,
    bit for bit (zn[q] written too; acor read before it is zeroed)"""
    L = la.lib()
    n = 30_001
    rng = np.random.default_rng(90 + q)
    Z = [rng.standard_normal(n) for _ in range(q + 1)]
    Z[0][:5] = -0.0
    A = rng.standard_normal(n)
    A[5:9] = 0.0
    za, zb = [vec(a) for a in Z], [vec(a) for a in Z]
    acor_a, acor_b = vec(A), vec(A)
    y_a, y_b = vec(np.zeros(n)), vec(np.zeros(n))
    av = vec(rng.uniform(1e-6, 1e-3, n)) if weights == "vector" else None
    w_a, w_b = vec(np.zeros(n)), vec(np.zeros(n))
    rtol, atol = 1e-4, (0.0 if weights == "vector" else 1e-5)
    lv = (C.c_double * 6)(*([1.0] + list(rng.uniform(-2.0, 2.0, 5))))
    eta = 0.8371
    cv = (C.c_double * 6)(*[eta ** j for j in range(6)])
    assert L.N_VScaleAddMulti_Hip(q + 1, lv, acor_a.v, arr(za), arr(za)) == 0
    if scaled:
        cs = (C.c_double * 5)(*[cv[j] for j in range(1, q + 1)])
        assert L.N_VScaleVectorArray_Hip(q, cs, arr(za[1:]), arr(za[1:])) == 0
    assert L.NVHip_PredictInit(q, arr(za), acor_a.v, y_a.v, rtol, atol, av.v if av else None,
                               w_a.v if weights else None) == 0
    assert L.NVHip_PredictComplete(q, arr(zb), acor_b.v, lv, cv if scaled else None, y_b.v,
                                   rtol, atol, av.v if av else None,
                                   w_b.v if weights else None) == 0
    for j in range(q + 1):
        assert np.array_equal(za[j].get(), zb[j].get()), j
    assert np.array_equal(acor_a.get(), acor_b.get())
    assert np.array_equal(y_a.get(), y_b.get())
    if weights:
        assert np.array_equal(w_a.get(), w_b.get())
    assert L.NVHip_PredictComplete(q, arr(zb), acor_b.v, None, None, y_b.v, rtol, atol, None,
                                   None) == -1   # no completion coefficients


@pytest.mark.parametrize("q,scaled", [(2, False), (4, True)])
def test_predict_complete_save_equals_copy_then_predict_complete(q, scaled):
    """NVHip_PredictCompleteSave == N_VScale(1, acor, save) then
    NVHip_PredictComplete, bit for bit; save may not alias zn, acor or y"""
    L = la.lib()
    n = 20_011
    rng = np.random.default_rng(120 + q)
    Z = [rng.standard_normal(n) for _ in range(q + 1)]
    A = rng.standard_normal(n)
    za, zb = [vec(a) for a in Z], [vec(a) for a in Z]
    acor_a, acor_b = vec(A), vec(A)
    y_a, y_b = vec(np.zeros(n)), vec(np.zeros(n))
    w_a, w_b = vec(np.zeros(n)), vec(np.zeros(n))
    sa, sb = vec(np.full(n, 3.0)), vec(np.full(n, 3.0))
    lv = (C.c_double * 6)(*([1.0] + list(rng.uniform(-2.0, 2.0, 5))))
    cv = (C.c_double * 6)(*[0.91 ** j for j in range(6)])
    c = cv if scaled else None
    L.N_VScale_Hip(1.0, acor_a.v, sa.v)
    assert L.NVHip_PredictComplete(q, arr(za), acor_a.v, lv, c, y_a.v, 1e-4, 1e-5, None,
                                   w_a.v) == 0
    assert L.NVHip_PredictCompleteSave(q, arr(zb), acor_b.v, lv, c, y_b.v, 1e-4, 1e-5, None,
                                       w_b.v, sb.v) == 0
    assert np.array_equal(sb.get(), A)
    assert np.array_equal(sa.get(), sb.get())
    for j in range(q + 1):
        assert np.array_equal(za[j].get(), zb[j].get()), j
    assert np.array_equal(acor_a.get(), acor_b.get())
    assert np.array_equal(y_a.get(), y_b.get())
    assert np.array_equal(w_a.get(), w_b.get())
    for bad in (zb[0], acor_b, y_b, w_b):
        assert L.NVHip_PredictCompleteSave(q, arr(zb), acor_b.v, lv, c, y_b.v, 1e-4, 1e-5,
                                           None, w_b.v, bad.v) == -1


def test_wrms_norm_pair_comp_equals_completion_then_pair():
    """NVHip_WrmsNormPairComp(x, c, xc, ...) == N_VScaleAddMulti(1, c, xc, x)
    then NVHip_WrmsNormPair, bit for bit, x and xc left unchanged"""
    L = la.lib()
    rng = np.random.default_rng(71)
    x, xc, yv, zv = _vecs(rng, 4)
    w, = _vecs(rng, 1, lo=0.5)
    x0, xc0 = x.get(), xc.get()
    a, b, c = -0.734, 1.0, 0.4166
    xt = vec(x0)
    cc = (C.c_double * 1)(c)
    assert L.N_VScaleAddMulti_Hip(1, cc, xc.v, arr([xt]), arr([xt])) == 0
    na, nb = (C.c_double * 2)(), (C.c_double * 2)()
    assert L.NVHip_WrmsNormPair(xt.v, a, yv.v, b, zv.v, w.v, na) == 0
    assert L.NVHip_WrmsNormPairComp(x.v, c, xc.v, a, yv.v, b, zv.v, w.v, nb) == 0
    assert list(na) == list(nb)
    assert np.array_equal(x.get(), x0) and np.array_equal(xc.get(), xc0)


def test_wrms_norm_pair_queued_equals_pair_and_keeps_its_slot():
    """NVHip_WrmsNormPairCompQueue then NVHip_WrmsNormPairWait ==
    NVHip_WrmsNormPairComp bit for bit; reductions waited for in between
    (first result slot) leave the queued pair alone; a later queue replaces
    it, and the wait for the replaced ticket reports -1."""
    L = la.lib()
    rng = np.random.default_rng(72)
    x, xc, yv, zv = _vecs(rng, 4)
    w, = _vecs(rng, 1, lo=0.5)
    a, b, c = -0.734, 1.0, 0.4166
    ref = (C.c_double * 2)()
    assert L.NVHip_WrmsNormPairComp(x.v, c, xc.v, a, yv.v, b, zv.v, w.v, ref) == 0
    t1 = C.c_uint(0)
    assert L.NVHip_WrmsNormPairCompQueue(x.v, c, xc.v, a, yv.v, b, zv.v, w.v, C.byref(t1)) == 0
    d = L.N_VDotProd_Hip(yv.v, zv.v)          # a reduction waited for in between
    assert dot_close(d, yv.get(), zv.get())
    nr = (C.c_double * 2)()
    assert L.NVHip_WrmsNormPairWait(x.v, t1.value, nr) == 0
    assert list(nr) == list(ref)
    t2 = C.c_uint(0)
    assert L.NVHip_WrmsNormPairCompQueue(x.v, -c, xc.v, b, yv.v, a, zv.v, w.v, C.byref(t2)) == 0
    assert t2.value != t1.value
    assert L.NVHip_WrmsNormPairWait(x.v, t2.value, nr) == 0
    bad = (C.c_double * 2)()
    assert L.NVHip_WrmsNormPairWait(x.v, t1.value, bad) == -1
    ref2 = (C.c_double * 2)()
    assert L.NVHip_WrmsNormPairComp(x.v, -c, xc.v, b, yv.v, a, zv.v, w.v, ref2) == 0
    assert list(nr) == list(ref2)


def test_wrms_norm_pair_equals_two_norms():
    """NVHip_WrmsNormPair == N_VWrmsNorm(x, w) and N_VLinearSum(a, y, b, z,
    t) + N_VWrmsNorm(t, w), bit for bit, with one host wait"""
    L = la.lib()
    rng = np.random.default_rng(70)
    x, yv, zv = _vecs(rng, 3)
    w, = _vecs(rng, 1, lo=0.5)
    t = vec(np.zeros(N_FUSED))
    a, b = -0.734, 1.0
    n1 = L.N_VWrmsNorm_Hip(x.v, w.v)
    L.N_VLinearSum_Hip(a, yv.v, b, zv.v, t.v)
    n2 = L.N_VWrmsNorm_Hip(t.v, w.v)
    nr = (C.c_double * 2)()
    s0 = L.NVHip_NumSyncs()
    assert L.NVHip_WrmsNormPair(x.v, a, yv.v, b, zv.v, w.v, nr) == 0
    assert L.NVHip_NumSyncs() - s0 == 1
    assert nr[0] == n1 and nr[1] == n2


def test_newton_update_comb_equals_lincombdiv_then_update():
    """NVHip_NewtonUpdateComb == NVHip_LinCombDiv then NVHip_NewtonUpdate(0),
    bit for bit (ycor, y and the norm; the solution vector not written)"""
    L = la.lib()
    rng = np.random.default_rng(80)
    X = _vecs(rng, 4)
    sv, = _vecs(rng, 1, lo=0.5)
    zn0, = _vecs(rng, 1)
    w, = _vecs(rng, 1, lo=0.5)
    ycor0 = rng.standard_normal(N_FUSED) * 1e-3
    c = (C.c_double * 4)(0.5, -0.25, 1.5e-3, 7.0)
    xa = vec(np.zeros(N_FUSED))
    ya, yb = vec(ycor0), vec(ycor0)
    outa, outb = vec(np.zeros(N_FUSED)), vec(np.zeros(N_FUSED))
    assert L.NVHip_LinCombDiv(4, c, arr(X), sv.v, xa.v) == 0
    dna, yna = C.c_double(), C.c_double(-1.0)
    assert L.NVHip_NewtonUpdate(0, ya.v, xa.v, zn0.v, w.v, outa.v, C.byref(dna),
                                C.byref(yna)) == 0
    dnb = C.c_double()
    assert L.NVHip_NewtonUpdateComb(yb.v, 4, c, arr(X), sv.v, zn0.v, w.v, outb.v,
                                    C.byref(dnb)) == 0
    assert np.array_equal(ya.get(), yb.get()) and np.array_equal(outa.get(), outb.get())
    assert dna.value == dnb.value
    assert L.NVHip_NewtonUpdateComb(yb.v, 4, c, arr(X), sv.v, zn0.v, w.v, X[0].v,
                                    C.byref(dnb)) == -1   # y aliases an input


class _NVecObj(C.Structure):
    _fields_ = [("content", C.c_void_p), ("ops", C.POINTER(C.c_void_p * 56)),
                ("sunctx", C.c_void_p)]


def _ops(v):
    return list(C.cast(v, C.POINTER(_NVecObj)).contents.ops.contents)


def test_fused_nvec_init_propagates_to_clones():
    """FusedNVec_Init (2d_i4/deferred_nvector.cu:182-235) installs the
    one-launch multi-dot and linear combination on a vector, and CVODE's
    internal clones (N_VClone through the ops table, then clones of clones)
    carry the same 56 slots; the XBraid buffer slots stay NULL.  The fused
    results equal the standard ops' results, and K > 16 (the reference's
    fallback point, deferred_nvector.cu:151-155) is still exact."""
    L = la.lib()
    n = 200_003
    rng = np.random.default_rng(77)
    v = vec(rng.standard_normal(n))
    base = _ops(v.v)
    assert base[32] is None and base[30] is None      # created without fused ops
    assert base[54] is not None and base[55] is not None
    L.FusedNVec_Init(v.v)
    ops = _ops(v.v)
    assert ops[30] is not None and ops[32] is not None   # lincomb, dotprodmulti
    assert ops[51] is None and ops[52] is None and ops[53] is None
    clone_fn = C.CFUNCTYPE(C.c_void_p, C.c_void_p)(ops[1])
    c1 = clone_fn(v.v)
    c2 = C.CFUNCTYPE(C.c_void_p, C.c_void_p)(_ops(c1)[1])(c1)
    try:
        assert _ops(c1) == ops and _ops(c2) == ops
        assert L.N_VGetLength_Hip(c2) == n
        x = rng.standard_normal(n)
        vx = vec(x)
        for k in (3, 16, 19):
            Y = [rng.standard_normal(n) for _ in range(k)]
            vY = [vec(a) for a in Y]
            dotmulti = C.CFUNCTYPE(C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p)(ops[32])
            d = (C.c_double * k)()
            assert dotmulti(k, vx.v, C.cast(arr(vY), C.c_void_p), C.cast(d, C.c_void_p)) == 0
            for j in range(k):
                assert dot_close(d[j], x, Y[j])
            c = rng.standard_normal(k)
            lincomb = C.CFUNCTYPE(C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p)(ops[30])
            assert lincomb(k, C.cast((C.c_double * k)(*c), C.c_void_p),
                           C.cast(arr(vY), C.c_void_p), c2) == 0
            got = np.empty(n)
            assert L.N_VGetToHost_Hip(c2, got.ctypes.data_as(C.POINTER(C.c_double))) == 0
            ref = sum(c[j] * Y[j] for j in range(k))
            assert np.max(np.abs(got - ref)) <= 1e-14 * k * np.max(np.abs(ref))
    finally:
        L.N_VDestroy_Hip(c2)
        L.N_VDestroy_Hip(c1)
