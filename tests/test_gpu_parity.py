"""GPU parity of the plug-ins (called through the C ABI) against the CPU
oracle, on seeded inputs at sizes the oracle finishes in seconds.

Tolerance: 1e-12 relative (max |diff| / max |ref|) for single FP64 f / Jv /
Psetup / Psolve evaluations -- the bar stated for FP64 RHS/Jv; masks and
compact id lists bit-exact.
"""
import ctypes as C

import numpy as np
import pytest

import llg_amd as la
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
TOL = 1e-12


def rel(a, b):
    d = np.max(np.abs(a - b))
    s = np.max(np.abs(b))
    return d / s if s > 0 else d


def perturbed(y, active, seed, amp=0.05):
    rng = np.random.default_rng(seed)
    n = active.size
    yy = y + amp * rng.standard_normal(y.size)
    yy *= np.tile(active.astype(float), 3)
    return yy


CASES = [
    # name, preset kwargs, geometry builder, ic
    ("i4_periodic", dict(name="i4", ng=64, ny=48, demag=4.0),
     lambda: la.Geometry.square_hole(64, 48, halfside_frac=0.22), "two_domain"),
    ("i4_padded", dict(name="square_hole", ng=48, ny=32, demag=4.0),
     lambda: la.Geometry.square_hole(48, 32, halfside_frac=0.2), "split_x"),
    ("i6_compact", dict(name="i6", ng=64, ny=64, demag=4.0),
     lambda: la.Geometry.ellipse(64, 64, 0.3, 0.3), "uniform"),
    ("i2_regularised", dict(name="i2", ng=40, ny=30, demag=4.0),
     lambda: la.Geometry.square_hole(40, 30, halfside_frac=0.2), "uniform"),
    ("i5_rectangle", dict(name="i5", ng=60, ny=40, demag=4.0),
     lambda: la.Geometry.rectangle(60, 40, 0.5, 0.5), "uniform"),
    ("i1_dmi_exact_norm", dict(name="i1", ng=48, ny=48, demag=2.0),
     lambda: la.Geometry.full(48, 48), "stripes"),
    ("plain128_neumann", dict(name="plain128", ng=40, ny=24),
     lambda: la.Geometry.full(40, 24), "tilted_z"),
    ("i3_polycrystal_windowed", dict(name="i3", ng=48, ny=32, demag=1.0),
     lambda: la.Geometry.polycrystal(48, 32, num_grains=8, dead_frac=0.25), "grain_bumps"),
    ("i4_jok_two_phase", dict(name="i4", ng=32, ny=24, demag=4.0, pc_mode=la.PC_JOK),
     lambda: la.Geometry.square_hole(32, 24, halfside_frac=0.2), "two_domain"),
    # grids whose demag FFT lengths take the register-resident kernels
    # (rows Px/2 and columns Py in the plan tables)
    ("i4_padded_regfft_128x64", dict(name="square_hole", ng=128, ny=64, demag=4.0),
     lambda: la.Geometry.square_hole(128, 64, halfside_frac=0.2), "split_x"),
    ("i1_periodic_regfft_256x128", dict(name="i1", ng=256, ny=128, demag=2.0),
     lambda: la.Geometry.full(256, 128), "stripes"),
    ("i3_windowed_regfft_256x128", dict(name="i3", ng=256, ny=128, demag=1.0),
     lambda: la.Geometry.polycrystal(256, 128, num_grains=12, dead_frac=0.2), "grain_bumps"),
]


def make(case):
    name, kw, gb, ic = case
    kw = dict(kw)
    geom = gb()
    cfg = la.preset(kw.pop("name"), **kw)
    y0 = {"two_domain": geom.ic_two_domain, "split_x": geom.ic_split_x,
          "uniform": geom.ic_uniform, "stripes": geom.ic_stripes,
          "tilted_z": geom.ic_tilted_z, "grain_bumps": geom.ic_grain_bumps}[ic]()
    return cfg, geom, y0


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_rhs_jv_precond_parity(case):
    cfg, geom, y0 = make(case)
    print(f"[gpu] case {case[0]} {cfg.ng}x{cfg.ny}", flush=True)
    sim = la.Simulation(cfg, geom)
    ref = orc.OracleSim(cfg, geom)
    # mask / lists bit-exact
    assert np.array_equal(sim.mask(), ref.mask())
    a1, h1 = sim.ids()
    a2, h2 = ref.ids()
    assert np.array_equal(a1, a2) and np.array_equal(h1, h2)
    y = perturbed(y0, geom.active, 1)
    f_gpu, yn_gpu = sim.rhs(y)
    f_ref, yn_ref = ref.rhs(y)
    assert rel(yn_gpu, yn_ref) < TOL          # in-place normalisation of y
    assert rel(f_gpu, f_ref) < TOL
    assert rel(sim.hdemag(), ref.hdemag()) < TOL
    v = np.random.default_rng(2).standard_normal(y.size)
    assert rel(sim.jtv(yn_ref, v), ref.jtv(yn_ref, v)) < TOL
    gamma = 0.037
    assert sim.psetup(yn_ref, gamma) == ref.psetup(yn_ref, gamma)
    if cfg.pc_mode == la.PC_JOK:
        # jok=1 reuses the stored J and only refactors with the new gamma
        assert sim.psetup(yn_ref * 0.5, 0.02, jok=1) == ref.psetup(yn_ref * 0.5, 0.02, jok=1) == 0
    assert rel(sim.pinv(), ref.pinv()) < TOL or cfg.exec_ == la.EXEC_COMPACT
    r = np.random.default_rng(3).standard_normal(y.size)
    assert rel(sim.psolve(r), ref.psolve(r)) < TOL


def test_demag_padded_matches_direct_sum():
    nx, ny = 24, 20
    rng = np.random.default_rng(5)
    y = rng.standard_normal(3 * nx * ny)
    d = la.lib().Demag_InitPadded(nx, ny, 1.0, 4.0)
    sim_vec_y = la.HipVector(3 * nx * ny, None, y)
    sim_vec_h = la.HipVector(3 * nx * ny, None)
    la.lib().Demag_SetStream(d, la.lib().N_VGetStream_Hip(sim_vec_y.v))
    la.lib().Demag_Apply(d, la.lib().N_VGetDeviceArrayPointer_Hip(sim_vec_y.v),
                         la.lib().N_VGetDeviceArrayPointer_Hip(sim_vec_h.v))
    h = sim_vec_h.get()
    la.lib().Demag_Destroy(d)
    h_ref = orc.demag_direct(nx, ny, 1.0, 4.0, y, padded=True)
    assert rel(h, h_ref) < TOL


def test_demag_periodic_matches_direct_sum_odd_rows():
    nx, ny = 30, 15
    rng = np.random.default_rng(6)
    y = rng.standard_normal(3 * nx * ny)
    d = la.lib().Demag_Init(nx, ny, 1.0, 2.0)
    vy = la.HipVector(3 * nx * ny, None, y)
    vh = la.HipVector(3 * nx * ny, None)
    la.lib().Demag_SetStream(d, la.lib().N_VGetStream_Hip(vy.v))
    la.lib().Demag_Apply(d, la.lib().N_VGetDeviceArrayPointer_Hip(vy.v),
                         la.lib().N_VGetDeviceArrayPointer_Hip(vh.v))
    h = vh.get()
    la.lib().Demag_Destroy(d)
    h_ref = orc.demag_direct(nx, ny, 1.0, 2.0, y, padded=False)
    assert rel(h, h_ref) < TOL


@pytest.mark.parametrize("padded", [False, True])
def test_demag_fp32_spectral_error(padded):
    """FP32 spectral demag against the FP64 pipeline: per-field max error
    relative to the field's max magnitude stays at FP32 level."""
    nx, ny = 96, 64
    geom = la.Geometry.square_hole(nx, ny, halfside_frac=0.2)
    y = geom.ic_split_x()
    out = {}
    for fp32 in (0, 1):
        cfg = la.preset("square_hole" if padded else "i4", ng=nx, ny=ny, demag=4.0)
        cfg.demag_fp32 = fp32
        sim = la.Simulation(cfg, geom)
        sim.rhs(y)
        out[fp32] = sim.hdemag().reshape(3, -1)
    for c in range(3):
        scale = np.max(np.abs(out[0][c])) + 1e-30
        err = np.max(np.abs(out[1][c] - out[0][c])) / scale
        print(f"[gpu] fp32 demag comp {c} rel err {err:.2e}")
        assert err < 2e-5


def test_integrate_square_hole_small():
    cfg = la.preset("square_hole", ng=48, ny=32)
    geom = la.Geometry.square_hole(48, 32, halfside_frac=0.2)
    y0 = geom.ic_split_x()
    sim = la.Simulation(cfg, geom)
    run = sim.integrator(y0, rtol=1e-4, atol=1e-4)
    out = run.advance(1.0)
    assert out["retval"] >= 0 and abs(out["t"] - 1.0) < 1e-9
    y = run.state()
    assert np.all(np.isfinite(y))
    assert np.all(y.reshape(3, -1)[:, geom.active == 0] == 0.0)


@pytest.mark.parametrize("case", [CASES[1], CASES[2], CASES[7]], ids=["padded", "compact", "i3"])
def test_atimes_equals_jtv_then_linear_sum(case):
    """LLG_Atimes == JtvProduct + N_VLinearSum(1, v, -gamma, Jv), bit for bit"""
    cfg, geom, y0 = make(case)
    sim = la.Simulation(cfg, geom)
    L = la.lib()
    rng = np.random.default_rng(5)
    vy = sim.vec(perturbed(y0, geom.active, 3))
    vv = sim.vec(rng.standard_normal(sim.neq))
    za, zb = sim.vec(np.zeros(sim.neq)), sim.vec(np.zeros(sim.neq))
    gamma = 0.0371
    assert L.JtvProduct(vv.v, za.v, 0.0, vy.v, None, sim.ctx, None) == 0
    L.N_VLinearSum_Hip(1.0, vv.v, -gamma, za.v, za.v)
    assert L.LLG_Atimes(vv.v, zb.v, gamma, 0.0, vy.v, None, sim.ctx) == 0
    assert np.array_equal(za.get(), zb.get())


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_rhs_residual_equals_rhs_then_linear_combination(case):
    """LLG_RhsResidual == LLG_Rhs into f, then N_VLinearCombination(3,
    {-rl1, -1, gamma}, {zn1, ycor, f}), bit for bit (the Newton residual of
    CVODE's cvNlsResidual); compact grids return 2 and leave res alone"""
    cfg, geom, y0 = make(case)
    sim = la.Simulation(cfg, geom)
    L = la.lib()
    rng = np.random.default_rng(21)
    n = sim.neq
    y = perturbed(y0, geom.active, 6)
    ya, yb = sim.vec(y), sim.vec(y)   # the RHS may normalise y in place
    zn1 = sim.vec(rng.standard_normal(n))
    ycor = sim.vec(1e-3 * rng.standard_normal(n))
    f, ra = sim.vec(np.zeros(n)), sim.vec(np.zeros(n))
    rb = sim.vec(np.full(n, 7.0))
    c = (C.c_double * 3)(-0.8125, -1.0, 0.0371)
    assert L.LLG_Rhs(0.0, ya.v, f.v, sim.ctx) == 0
    L.N_VLinearCombination_Hip(3, c, (C.c_void_p * 3)(zn1.v, ycor.v, f.v), ra.v)
    rc = L.LLG_RhsResidual(0.0, yb.v, c, zn1.v, ycor.v, rb.v, sim.ctx)
    if cfg.exec_ == la.EXEC_COMPACT:
        assert rc == 2 and np.all(rb.get() == 7.0)
        return
    assert rc == 0
    assert np.array_equal(ya.get(), yb.get())
    assert np.array_equal(ra.get(), rb.get())
    assert L.LLG_RhsResidual(0.0, yb.v, c, zn1.v, ycor.v, ycor.v, sim.ctx) == -1


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_krylov_chain_equals_separate_calls(case):
    """LLG_Krylov == NVHip_ScaleDiv + LLG_Atimes + PrecondSolve + N_VProd and
    NVHip_GSProjectScaled == (N_VScale of the basis vector) + NVHip_GSProject,
    bit for bit; the chain leaves the basis vector it reads unscaled"""
    cfg, geom, y0 = make(case)
    sim = la.Simulation(cfg, geom)
    L = la.lib()
    rng = np.random.default_rng(6)
    n = sim.neq
    y = perturbed(y0, geom.active, 4)
    sim.rhs(y)
    gamma, c = 0.0371, 0.4375
    sim.psetup(y, gamma)
    vy = sim.vec(y)
    s = sim.vec(1.0 / (1e-4 * np.abs(rng.standard_normal(n)) + 1e-4))
    x = rng.standard_normal(n) * np.tile(geom.active.astype(float), 3)
    # separate calls (the unfused SPGMR iteration)
    vl_a, u, z, t, vk_a = sim.vec(x), sim.vec(), sim.vec(), sim.vec(), sim.vec()
    assert L.NVHip_ScaleDiv(c, vl_a.v, s.v, vl_a.v, u.v) == 0
    assert L.LLG_Atimes(u.v, z.v, gamma, 0.0, vy.v, None, sim.ctx) == 0
    assert L.PrecondSolve(0.0, vy.v, None, z.v, t.v, gamma, 0.0, la.PREC_LEFT, sim.ctx) == 0
    L.N_VProd_Hip(s.v, t.v, vk_a.v)
    # the chain
    vl_b, vk_b = sim.vec(x), sim.vec()
    assert L.LLG_Krylov(c, vl_b.v, s.v, vk_b.v, gamma, 0.0, vy.v, None, sim.ctx) == 0
    assert np.array_equal(vl_b.get(), x)
    assert np.array_equal(vk_a.get(), vk_b.get())
    assert L.LLG_Krylov(c, vl_b.v, s.v, vl_b.v, gamma, 0.0, vy.v, None, sim.ctx) == -1
    # Gram-Schmidt against [w0, w1, vl, vk], vl normalised in the pass
    w = [rng.standard_normal(n) for _ in range(2)]
    wa, wb = [sim.vec(a) for a in w], [sim.vec(a) for a in w]
    Ya = (C.c_void_p * 4)(wa[0].v, wa[1].v, vl_a.v, vk_a.v)
    Yb = (C.c_void_p * 4)(wb[0].v, wb[1].v, vl_b.v, vk_b.v)
    da, db = (C.c_double * 4)(), (C.c_double * 4)()
    za, zb = C.c_double(), C.c_double()
    assert L.NVHip_GSProject(s.v, t.v, vk_a.v, 4, Ya, da, C.byref(za)) == 0
    assert L.NVHip_GSProjectScaled(vk_b.v, 4, Yb, 2, c, db, C.byref(zb)) == 0
    assert list(da) == list(db) and za.value == zb.value
    assert np.array_equal(vl_a.get(), vl_b.get())
    assert np.array_equal(vk_a.get(), vk_b.get())
    print(f"[gpu] {case[0]}: |vk|^2 after projection {zb.value:.6e}")


def _dot_ok(got, a, b, tol=1e-13):
    """|got - a.b| within tol of the Cauchy-Schwarz scale |a| |b|"""
    ref = float(a @ b)
    return abs(got - ref) <= tol * np.linalg.norm(a) * np.linalg.norm(b), (got, ref)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_arnoldi_step_equals_separate_calls(case):
    """LLG_Arnoldi, one SPGMR iteration with lagged normalisation in one pass:
    the first step's u_0 equals PrecondSolve + N_VProd and every step's u
    equals NVHip_ProjNorm's combination, bit for bit; w = A u equals
    LLG_Krylov(1, u) bit for bit; the sums |u|^2, <w, U_j>, <w, u>, |w|^2
    (and the first step's sum (b s)^2) agree with numpy to 1e-13 of |a| |b|;
    compact execution and aliasing
    are refused without a launch"""
    cfg, geom, y0 = make(case)
    print(f"[gpu] case {case[0]} {cfg.ng}x{cfg.ny}", flush=True)
    sim = la.Simulation(cfg, geom)
    L = la.lib()
    rng = np.random.default_rng(8)
    n = sim.neq
    y = perturbed(y0, geom.active, 4)
    sim.rhs(y)
    gamma = 0.0371
    sim.psetup(y, gamma)
    vy = sim.vec(y)
    s = sim.vec(1.0 / (1e-4 * np.abs(rng.standard_normal(n)) + 1e-4))
    vb = sim.vec(rng.standard_normal(n) * np.tile(geom.active.astype(float), 3))
    nl = 5
    U = [sim.vec() for _ in range(nl + 1)]
    Up = (C.c_void_p * (nl + 1))(*[u.v for u in U])
    w = [sim.vec(), sim.vec()]
    r = (C.c_double * 16)()
    rc = L.LLG_Arnoldi(0, 1, 1.0, None, vb.v, Up, w[0].v, s.v, gamma, 0.0, vy.v, None, sim.ctx,
                       r)
    if cfg.exec_ == la.EXEC_COMPACT:
        assert rc == -1
        return
    assert rc == 0
    t, u0, wk = sim.vec(), sim.vec(), sim.vec()
    assert L.PrecondSolve(0.0, vy.v, None, vb.v, t.v, gamma, 0.0, la.PREC_LEFT, sim.ctx) == 0
    L.N_VProd_Hip(s.v, t.v, u0.v)
    assert np.array_equal(U[0].get(), u0.get())
    assert L.LLG_Krylov(1.0, u0.v, s.v, wk.v, gamma, 0.0, vy.v, None, sim.ctx) == 0
    assert np.array_equal(w[0].get(), wk.get())
    u, wv = U[0].get(), w[0].get()
    bs = vb.get() * s.get()
    for k, (a, b) in enumerate([(u, u), (wv, u), (wv, wv), (bs, bs)]):
        ok, vals = _dot_ok(r[k], a, b)
        assert ok, (0, k, vals)
    # without the preconditioner solve, step 0 applies A to U[0] as it is
    assert L.LLG_Arnoldi(0, 0, 1.0, None, None, Up, w[1].v, s.v, gamma, 0.0, vy.v, None,
                         sim.ctx, r) == 0
    assert np.array_equal(w[1].get(), wk.get())
    uref = sim.vec()
    for l in range(1, nl + 1):
        a = float(rng.uniform(0.5, 2.0))
        b = (C.c_double * 8)(*rng.uniform(-1.0, 1.0, 8))
        win, wout = w[(l - 1) % 2], w[l % 2]
        assert L.LLG_Arnoldi(l, 0, a, b, win.v, Up, win.v, s.v, gamma, 0.0, vy.v, None,
                             sim.ctx, r) == -1          # wout aliases win
        assert L.LLG_Arnoldi(l, 0, a, b, win.v, Up, wout.v, s.v, gamma, 0.0, vy.v, None,
                             sim.ctx, r) == 0
        nn = C.c_double()
        assert L.NVHip_ProjNorm(l, a, b, win.v, Up, uref.v, C.byref(nn)) == 0
        assert np.array_equal(U[l].get(), uref.get()), l
        assert L.LLG_Krylov(1.0, uref.v, s.v, wk.v, gamma, 0.0, vy.v, None, sim.ctx) == 0
        assert np.array_equal(wout.get(), wk.get()), l
        u, wv = U[l].get(), wout.get()
        pairs = [(u, u)] + [(wv, U[j].get()) for j in range(l)] + [(wv, u), (wv, wv)]
        for k, (x1, x2) in enumerate(pairs):
            ok, vals = _dot_ok(r[k], x1, x2)
            assert ok, (l, k, vals)
        ok, vals = _dot_ok(nn.value, u, u)
        assert ok, (l, "projnorm", vals)
    print(f"[gpu] {case[0]}: step {nl} |u|^2 {r[0]:.6e} |w|^2 {r[nl + 2]:.6e}")


@pytest.mark.parametrize("case", CASES[:2], ids=[c[0] for c in CASES[:2]])
def test_arnoldi_first_pass_with_residual_formed_where_read(case):
    """LLG_ArnoldiFirst (b = sum c_k X_k formed where the first pass reads
    it) == N_VLinearCombination then LLG_Arnoldi(0, pre), bit for bit"""
    cfg, geom, y0 = make(case)
    sim = la.Simulation(cfg, geom)
    L = la.lib()
    rng = np.random.default_rng(12)
    n = sim.neq
    y = perturbed(y0, geom.active, 4)
    sim.rhs(y)
    gamma = 0.0371
    sim.psetup(y, gamma)
    vy = sim.vec(y)
    s = sim.vec(1.0 / (1e-4 * np.abs(rng.standard_normal(n)) + 1e-4))
    X = [sim.vec(rng.standard_normal(n)) for _ in range(3)]
    Xp = (C.c_void_p * 3)(*[x.v for x in X])
    c = (C.c_double * 3)(-1.25, -1.0, gamma)
    b = sim.vec()
    L.N_VLinearCombination_Hip(3, c, Xp, b.v)
    Ua, Ub = [sim.vec()], [sim.vec()]
    wa, wb = sim.vec(), sim.vec()
    ra, rb = (C.c_double * 8)(), (C.c_double * 8)()
    assert L.LLG_Arnoldi(0, 1, 1.0, None, b.v, (C.c_void_p * 1)(Ua[0].v), wa.v, s.v, gamma,
                         0.0, vy.v, None, sim.ctx, ra) == 0
    assert L.LLG_ArnoldiFirst(3, c, Xp, (C.c_void_p * 1)(Ub[0].v), wb.v, s.v, gamma, 0.0,
                              vy.v, None, sim.ctx, rb) == 0
    assert np.array_equal(Ua[0].get(), Ub[0].get()) and np.array_equal(wa.get(), wb.get())
    assert list(ra)[:4] == list(rb)[:4]


def test_gated_arnoldi_step_equals_direct():
    """LLG_ArnoldiQueue behind an NVHip_GateOpen gate, released with the
    coefficients, gives LLG_Arnoldi's vectors and sums bit for bit; released
    with stop, the pass and its finish are skipped (the wait reports the
    missing result); a gate the host does not release in time (1 s) stops
    by itself; the stream then works as before"""
    L = la.lib()
    L.NVHip_SetSyncMode(1)   # gates need the polling wait
    try:
        _gated_arnoldi_checks(L)
    finally:
        L.NVHip_SetSyncMode(0)


def _gated_arnoldi_checks(L):
    cfg, geom, y0 = make(CASES[1])
    sim = la.Simulation(cfg, geom)
    rng = np.random.default_rng(9)
    n = sim.neq
    y = perturbed(y0, geom.active, 4)
    sim.rhs(y)
    gamma = 0.0371
    sim.psetup(y, gamma)
    vy = sim.vec(y)
    s = sim.vec(1.0 / (1e-4 * np.abs(rng.standard_normal(n)) + 1e-4))
    U = [sim.vec(rng.standard_normal(n)) for _ in range(3)]
    Up = (C.c_void_p * 3)(*[u.v for u in U])
    win = sim.vec(rng.standard_normal(n))
    wa, wb = sim.vec(), sim.vec()
    l, a = 2, 0.8125
    b = (C.c_double * 8)(0.25, -0.375)
    ra, rb = (C.c_double * 16)(), (C.c_double * 16)()
    assert L.LLG_Arnoldi(l, 0, a, b, win.v, Up, wa.v, s.v, gamma, 0.0, vy.v, None, sim.ctx,
                         ra) == 0
    ua = U[l].get()
    cdev, sdev, tk = C.c_void_p(), C.c_void_p(), C.c_uint()
    # released with the coefficients
    assert L.NVHip_GateOpen(wb.v, C.byref(cdev), C.byref(sdev)) == 0
    assert L.NVHip_GateOpen(wb.v, C.byref(cdev), C.byref(sdev)) == -1   # one at a time
    assert L.LLG_ArnoldiQueue(l, 0, 0.0, None, cdev, sdev, win.v, Up, wb.v, s.v, gamma, 0.0,
                              vy.v, None, sim.ctx, C.byref(tk)) == 0
    cf = (C.c_double * 3)(a, b[0], b[1])
    assert L.NVHip_GateRelease(wb.v, 0, cf, 3) == 0
    assert L.NVHip_ReduceWait(wb.v, tk.value, l + 3, rb) == 0
    assert np.array_equal(wb.get(), wa.get()) and np.array_equal(U[l].get(), ua)
    assert list(rb)[:l + 3] == list(ra)[:l + 3]
    # released with stop: nothing runs, the result never arrives
    wc = sim.vec(np.full(n, 7.0))
    assert L.NVHip_GateOpen(wc.v, C.byref(cdev), C.byref(sdev)) == 0
    assert L.LLG_ArnoldiQueue(l, 0, 0.0, None, cdev, sdev, win.v, Up, wc.v, s.v, gamma, 0.0,
                              vy.v, None, sim.ctx, C.byref(tk)) == 0
    assert L.NVHip_GateRelease(wc.v, 1, None, 0) == 0
    assert L.NVHip_ReduceWait(wc.v, tk.value, l + 3, rb) == -1
    assert np.all(wc.get() == 7.0)
    # not released in time: the gate stops by itself after 1 s
    assert L.NVHip_GateOpen(wc.v, C.byref(cdev), C.byref(sdev)) == 0
    assert L.LLG_ArnoldiQueue(l, 0, 0.0, None, cdev, sdev, win.v, Up, wc.v, s.v, gamma, 0.0,
                              vy.v, None, sim.ctx, C.byref(tk)) == 0
    assert L.NVHip_ReduceWait(wc.v, tk.value, l + 3, rb) == -1
    assert np.all(wc.get() == 7.0)
    assert L.NVHip_GateRelease(wc.v, 1, None, 0) == 0      # late release is harmless
    # and the stream works as before
    assert L.LLG_Arnoldi(l, 0, a, b, win.v, Up, wc.v, s.v, gamma, 0.0, vy.v, None, sim.ctx,
                         rb) == 0
    assert np.array_equal(wc.get(), wa.get())


@pytest.mark.parametrize("ng,ny", [(96, 64), (128, 64)])
def test_demag_autotune_keeps_the_field(ng, ny):
    """the autotuner changes launch shapes, LDS layout and the FFT kernel
    family: within a family the field must not change by a single bit, and
    across families it stays within 1e-12 relative"""
    cfg = la.preset("square_hole", ng=ng, ny=ny)
    geom = la.Geometry.square_hole(ng, ny, halfside_frac=0.2)
    sim = la.Simulation(cfg, geom)
    L = la.lib()
    d = L.LLG_GetDemag(sim.ctx)
    y = sim.vec(geom.ic_split_x())
    h0, h1 = sim.vec(), sim.vec()
    rv0, cv0 = C.c_int(), C.c_int()
    assert L.Demag_GetFftKernels(d, C.byref(rv0), C.byref(cv0)) == 0
    L.Demag_Apply(d, L.N_VGetDeviceArrayPointer_Hip(y.v), L.N_VGetDeviceArrayPointer_Hip(h0.v))
    choice = (C.c_int * 6)()
    assert L.Demag_Autotune(d, 3, choice) == 0
    print(f"[gpu] {ng}x{ny} kernels ({rv0.value},{cv0.value}) autotune choice {list(choice)}")
    L.Demag_Apply(d, L.N_VGetDeviceArrayPointer_Hip(y.v), L.N_VGetDeviceArrayPointer_Hip(h1.v))
    if (choice[4], choice[5]) == (rv0.value, cv0.value):
        assert np.array_equal(h0.get(), h1.get())
    else:
        assert rel(h1.get(), h0.get()) < TOL
    # back on the original kernels, whatever the tuned launch shapes
    assert L.Demag_SetFftKernels(d, rv0.value, cv0.value) == 0
    L.Demag_Apply(d, L.N_VGetDeviceArrayPointer_Hip(y.v), L.N_VGetDeviceArrayPointer_Hip(h1.v))
    assert np.array_equal(h0.get(), h1.get())


def _demag_field(L, d, n3, y):
    vy = la.HipVector(n3, None, y)
    vh = la.HipVector(n3, None)
    L.Demag_SetStream(d, L.N_VGetStream_Hip(vy.v))
    L.Demag_Apply(d, L.N_VGetDeviceArrayPointer_Hip(vy.v), L.N_VGetDeviceArrayPointer_Hip(vh.v))
    return vh.get()


@pytest.mark.parametrize("nx,ny,padded", [(128, 64, True), (256, 128, False)])
def test_register_fft_demag_matches_direct_sum(nx, ny, padded):
    """the register-resident FFT kernels against the O(N^2) direct sum"""
    L = la.lib()
    rng = np.random.default_rng(7)
    y = rng.standard_normal(3 * nx * ny)
    d = (L.Demag_InitPadded if padded else L.Demag_Init)(nx, ny, 1.0, 3.0)
    rv, cv = C.c_int(), C.c_int()
    L.Demag_GetFftKernels(d, C.byref(rv), C.byref(cv))
    assert (rv.value, cv.value) == (0, 0)    # the register kernels are the default here
    h = _demag_field(L, d, 3 * nx * ny, y)
    L.Demag_Destroy(d)
    h_ref = orc.demag_direct(nx, ny, 1.0, 3.0, y, padded=padded)
    print(f"[gpu] {nx}x{ny} padded={padded}: rel err vs direct sum {rel(h, h_ref):.2e}")
    assert rel(h, h_ref) < TOL


REG_SHAPES = [
    # (nx, ny, padded): bench grid, reference 512^2 reading, ensemble and
    # polycrystal grids
    (1536, 512, True), (512, 512, False), (256, 256, False), (512, 256, False),
    (768, 512, False), (1024, 512, False), (1024, 1024, False), (1024, 1024, True),
    (2048, 512, True),
]
# the oracle's FFT demag (tensor on the host) is affordable up to this many
# displacement cells
ORACLE_MAX_CELLS = 1 << 21


@pytest.mark.parametrize("nx,ny,padded", REG_SHAPES,
                         ids=[f"{a}x{b}{'_padded' if p else ''}" for a, b, p in REG_SHAPES])
def test_register_fft_variants_match_stockham(nx, ny, padded):
    """every FFT kernel family for the grid (register variants, LDS
    Stockham) against the oracle's FFT demag where affordable, else
    against the Stockham kernels: FP64 within 1e-12 relative per field
    component; the FP32 spectral pipeline within 2e-5"""
    L = la.lib()
    n3 = 3 * nx * ny
    rng = np.random.default_rng(nx + ny)
    y = rng.standard_normal(n3)
    d = (L.Demag_InitPadded if padded else L.Demag_Init)(nx, ny, 1.0, 4.0)
    cells = (2 * nx) * (2 * ny) if padded else nx * ny
    if cells <= ORACLE_MAX_CELLS:
        h_ref = orc.Demag(nx, ny, 1.0, 4.0, padded=padded).apply(y).reshape(3, -1)
        combos = [(r, c) for r in (-1, 0) for c in (-1, 0, 1, 2, 3, 4)]
        ref_name = "oracle"
    else:
        assert L.Demag_SetFftKernels(d, -1, -1) == 0
        h_ref = _demag_field(L, d, n3, y).reshape(3, -1)
        combos = [(r, c) for r in (-1, 0) for c in (-1, 0, 1, 2, 3, 4) if (r, c) != (-1, -1)]
        ref_name = "stockham"
    ran, worst = 0, {}
    for fp32 in (0, 1):
        assert L.Demag_SetPrecision(d, fp32) == 0
        for r, c in combos:
            if L.Demag_SetFftKernels(d, r, c) != 0:
                continue
            h = _demag_field(L, d, n3, y).reshape(3, -1)
            errs = [rel(h[k], h_ref[k]) for k in range(3)]
            worst[(fp32, r, c)] = max(errs)
            print(f"[gpu] {nx}x{ny} padded={padded} fp32={fp32} kernels ({r},{c}) vs "
                  f"{ref_name}: rel err {max(errs):.2e}", flush=True)
            ran += 1
    L.Demag_Destroy(d)
    assert ran >= 4
    for (fp32, r, c), e in worst.items():
        assert e < (2e-5 if fp32 else TOL), (fp32, r, c, e)


CB_SHAPES = [(1536, 512, True), (1024, 1024, True), (256, 256, True), (512, 512, False),
             (1024, 1024, False), (1024, 512, False)]


@pytest.mark.parametrize("fp32", [0, 1])
@pytest.mark.parametrize("nx,ny,padded", CB_SHAPES,
                         ids=[f"{a}x{b}{'_padded' if p else ''}" for a, b, p in CB_SHAPES])
def test_blocked_column_variants_equal_one_column_kernel_bitwise(nx, ny, padded, fp32):
    """Column variants 2, 3, 4 (2, 4, 8 adjacent columns per workgroup, the
    spectrum staged through LDS, the padded tensor read for ky <= Py/2 only)
    run variant 0's transforms on the same values: the field must be the
    same bits.  Hx = Px/2 + 1 is odd, so the last block is always partial."""
    L = la.lib()
    n3 = 3 * nx * ny
    y = np.random.default_rng(nx * 7 + ny).standard_normal(n3)
    d = (L.Demag_InitPadded if padded else L.Demag_Init)(nx, ny, 1.0, 4.0)
    assert L.Demag_SetPrecision(d, fp32) == 0
    assert L.Demag_SetFftKernels(d, 0, 0) == 0
    h0 = _demag_field(L, d, n3, y)
    ran = 0
    for cv in (2, 3, 4):
        if L.Demag_SetFftKernels(d, 0, cv) != 0:
            continue
        h = _demag_field(L, d, n3, y)
        print(f"[gpu] {nx}x{ny} padded={padded} fp32={fp32} column variant {cv}: "
              f"max|dh| {np.max(np.abs(h - h0)):.1e}", flush=True)
        assert np.array_equal(h, h0), cv
        ran += 1
    L.Demag_Destroy(d)
    assert ran >= 2


WAVE_SHAPES = [(1536, 512, True), (512, 512, True), (40, 512, True), (512, 1024, False),
               (96, 1024, False)]


@pytest.mark.parametrize("fp32", [0, 1])
@pytest.mark.parametrize("nx,ny,padded", WAVE_SHAPES,
                         ids=[f"{a}x{b}{'_padded' if p else ''}" for a, b, p in WAVE_SHAPES])
def test_wave_column_variants_equal_one_column_kernel_bitwise(nx, ny, padded, fp32):
    """Column variants 5 and 6 run each 1024-point column transform on one
    wave (16 values per lane, exchanges split into real and imaginary
    parts, x and y waves swapping their spectra through LDS) with the
    arithmetic of variant 1's one-column kernel: the field must be the same
    bits.  The shapes leave partial x/y and z workgroups (Hx odd, and 21 or
    49 columns for the narrow grids)."""
    L = la.lib()
    n3 = 3 * nx * ny
    y = np.random.default_rng(nx * 11 + ny).standard_normal(n3)
    d = (L.Demag_InitPadded if padded else L.Demag_Init)(nx, ny, 1.0, 4.0)
    assert L.Demag_SetPrecision(d, fp32) == 0
    rv = 0 if L.Demag_SetFftKernels(d, 0, 1) == 0 else -1   # rows without a register plan
    assert L.Demag_SetFftKernels(d, rv, 1) == 0
    h1 = _demag_field(L, d, n3, y)
    for cv in (5, 6):
        assert L.Demag_SetFftKernels(d, rv, cv) == 0, cv
        h = _demag_field(L, d, n3, y)
        print(f"[gpu] {nx}x{ny} padded={padded} fp32={fp32} column variant {cv}: "
              f"max|dh| {np.max(np.abs(h - h1)):.1e}", flush=True)
        assert np.array_equal(h, h1), cv
    L.Demag_Destroy(d)


def test_padded_tensor_spectrum_is_mirror_symmetric():
    """the zero-padded grid's tensor spectra are stored with K(Py - ky) =
    K(ky) (xx, yy, zz) and -K(ky) (xy) exactly, which the blocked column
    kernels rely on; the field still matches the direct sum"""
    nx, ny = 40, 24
    y = np.random.default_rng(17).standard_normal(3 * nx * ny)
    L = la.lib()
    d = L.Demag_InitPadded(nx, ny, 1.0, 2.0)
    h = _demag_field(L, d, 3 * nx * ny, y)
    L.Demag_Destroy(d)
    assert rel(h, orc.demag_direct(nx, ny, 1.0, 2.0, y, padded=True)) < TOL


@pytest.mark.parametrize("fp32", [0, 1])
@pytest.mark.parametrize("shape", ["square_hole_1536x512_padded", "full_512x512_periodic"])
def test_compile_time_fft_plans_equal_general_plans_bitwise(shape, fp32):
    """the FFT kernels specialised for the benchmarked grid shapes run the
    same arithmetic as the general ones: the field must agree bit for bit"""
    if shape.startswith("square_hole"):
        ng, ny = 1536, 512
        geom = la.Geometry.square_hole(ng, ny, halfside_frac=222.0 / 512.0)
        cfg = la.preset("square_hole", ng=ng, ny=ny, padded=True, demag_fp32=fp32)
        y0 = geom.ic_split_x()
    else:
        ng = ny = 512
        geom = la.Geometry.full(ng, ny)
        cfg = la.preset("i1", ng=ng, ny=ny, demag=2.0, demag_fp32=fp32)
        y0 = geom.ic_stripes(left=0.25, right=0.75, eps=0.01)
    sim = la.Simulation(cfg, geom)
    L = la.lib()
    d = L.LLG_GetDemag(sim.ctx)
    assert L.Demag_SetFftKernels(d, -1, -1) == 0  # the LDS Stockham kernels
    y = sim.vec(perturbed(y0, geom.active, 11, amp=0.3))
    h_ct, h_gen = sim.vec(), sim.vec()
    assert L.Demag_SetSpecialized(d, 1) == 2      # rows and columns both specialised
    L.Demag_Apply(d, L.N_VGetDeviceArrayPointer_Hip(y.v), L.N_VGetDeviceArrayPointer_Hip(h_ct.v))
    assert L.Demag_SetSpecialized(d, 0) == 0
    L.Demag_Apply(d, L.N_VGetDeviceArrayPointer_Hip(y.v), L.N_VGetDeviceArrayPointer_Hip(h_gen.v))
    a, b = h_ct.get(), h_gen.get()
    print(f"[gpu] {shape} fp32={fp32}: max|h| {np.max(np.abs(b)):.3e}")
    assert np.max(np.abs(b)) > 0
    assert np.array_equal(a, b)


def test_context_autotune_keeps_callbacks_bitwise():
    cfg, geom, y0 = make(CASES[1])
    sim = la.Simulation(cfg, geom)
    L = la.lib()
    rng = np.random.default_rng(4)
    y = perturbed(y0, geom.active, 8)
    v = rng.standard_normal(sim.neq)
    f0, _ = sim.rhs(y)
    j0 = sim.jtv(y, v)
    choice = (C.c_int * 8)()
    assert L.LLG_Autotune(sim.ctx, 3, choice) == 0
    print(f"[gpu] context autotune choice {list(choice)}")
    f1, _ = sim.rhs(y)
    j1 = sim.jtv(y, v)
    # the demag FFT kernel family may change (rounding only)
    assert rel(f1, f0) < TOL
    assert np.array_equal(np.asarray(j0), np.asarray(j1))
    assert L.LLG_SetStencilBlock(sim.ctx, 50, 4) == -1   # 200 threads: not whole waves
