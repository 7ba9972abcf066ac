"""The benchmark workloads at their real sizes on the GPU.

* 1536x512-cell square hole with the zero-padded 3072x1024 demag (the
  throughput workload): f, the demag field, Jv, P^-1 and the P-solve
  against the CPU oracle at full size, 1e-12 relative; the FP32 spectral
  demag's per-component error against the same oracle field.
* 64-instance ensemble on the 256x256 ... 1024x512 grids (ellipse / ring
  geometries, in-plane fields, periodic demag 2.0): every member bit-equal
  to its single-instance run; the two smallest members' trajectories
  against the integrator on the oracle's callbacks (rtol = atol = 1e-6,
  max |dm| < 1e-5).
* 4096x4096 polycrystal (400 grains, 10 % dead, padded 8192x8192 demag):
  compact and dense-mask execution equal for f, Jv and the P-solve; the
  torque identity m.f = g alpha (m.h)(1 - |m|^2) with h restated in numpy
  from the i3 formulas (2d_i3/2d_fft.cu:451-627) plus the device's demag
  field; a two-stage field ramp (LLG_RunRamp) at full size, and a ramp on a
  small polycrystal against the same ramp on the oracle's callbacks.
"""
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


def _perturbed(y, active, seed, amp=0.05):
    rng = np.random.default_rng(seed)
    return (y + amp * rng.standard_normal(y.size)) * np.tile(active.astype(float), 3)


# ------------------------------------------------------------- 1536 x 512
def test_square_hole_1536x512_padded_matches_oracle_at_full_size():
    ng, ny = 1536, 512
    geom = la.Geometry.square_hole(ng, ny, halfside_frac=222.0 / 512.0)
    cfg = la.preset("square_hole", ng=ng, ny=ny, padded=True)
    y = _perturbed(geom.ic_split_x(max_angle=0.05, seed=1), geom.active, 31)
    print(f"[gpu] square hole {ng}x{ny}, padded {2 * ng}x{2 * ny} demag", flush=True)
    sim = la.Simulation(cfg, geom)
    ref = orc.OracleSim(cfg, geom)
    f_gpu, _ = sim.rhs(y)
    f_ref, _ = ref.rhs(y)
    hd_ref = ref.hdemag().reshape(3, -1)
    hd_gpu = sim.hdemag().reshape(3, -1)
    errs = [rel(hd_gpu[k], hd_ref[k]) for k in range(3)]
    print(f"[gpu] demag field rel err per component {['%.2e' % e for e in errs]}, "
          f"f rel err {rel(f_gpu, f_ref):.2e}", flush=True)
    assert max(errs) < TOL
    assert rel(f_gpu, f_ref) < TOL
    v = np.random.default_rng(2).standard_normal(y.size)
    assert rel(sim.jtv(y, v), ref.jtv(y, v)) < TOL
    gamma = 0.0371
    assert sim.psetup(y, gamma) == ref.psetup(y, gamma) == 1
    assert rel(sim.pinv(), ref.pinv()) < TOL
    r = np.random.default_rng(3).standard_normal(y.size)
    assert rel(sim.psolve(r), ref.psolve(r)) < TOL
    # the FP32 spectral form of the same field, per component
    L = la.lib()
    d = L.LLG_GetDemag(sim.ctx)
    assert L.Demag_SetPrecision(d, 1) == 0
    sim.rhs(y)
    h32 = sim.hdemag().reshape(3, -1)
    e32 = [rel(h32[k], hd_ref[k]) for k in range(3)]
    print(f"[gpu] FP32 spectral demag rel err per component "
          f"{['%.2e' % e for e in e32]}", flush=True)
    assert max(e32) < 2e-5
    ref.close()
    sim.close()


# --------------------------------------------------------- 64 instances
def test_ensemble_64_instances_bitwise_and_against_oracle():
    members = la.ensemble_members(n=64, seed=4)     # the ensemble workload
    grids = sorted({(c.ng, c.ny) for c, _, _ in members})
    ncell = sum(c.ng * c.ny for c, _, _ in members)
    print(f"[gpu] ensemble of {len(members)} instances on {grids}, {ncell} cells", flush=True)
    assert len(grids) == 4
    kw = dict(rtol=1e-4, atol=1e-4, t_out=0.1, mxstep=100000)
    t_end = 0.2
    ens = la.Ensemble(members, nthreads=0, **kw)
    out = ens.advance(t_end)
    assert out["failed"] == 0 and abs(out["inst_time"] - 64 * t_end) < 1e-9
    print(f"[gpu] ensemble advanced: nst={out['nst']} nli={out['nli']}", flush=True)
    for i, (cfg, geom, y0) in enumerate(members):
        sim = la.Simulation(cfg, geom)
        run = sim.integrator(y0, **kw)
        single = run.advance(t_end)
        r = ens.result(i)
        assert (r["nst"], r["nli"], r["nfe"]) == (single["nst"], single["nli"], single["nfe"]), i
        assert np.array_equal(ens.state(i), run.state()), i
        run.close()
        sim.close()
    print("[gpu] all 64 members bit-equal to their single runs", flush=True)
    # batched evaluation: same-shape members share one launch per pass of f
    bat = la.Ensemble(members, nthreads=0, **kw)
    assert bat.set_batched(1) == 4
    outb = bat.advance(t_end)
    st = bat.batch_stats()
    print(f"[gpu] batched: {st['calls']} f calls in {st['launches']} group evaluations", flush=True)
    assert outb["failed"] == 0 and st["launches"] < st["calls"]
    for i in range(len(members)):
        a, b = bat.result(i), ens.result(i)
        assert (a["nst"], a["nli"], a["nfe"]) == (b["nst"], b["nli"], b["nfe"]), i
        assert np.array_equal(bat.state(i), ens.state(i)), i
    bat.close()
    ens.close()
    small = [k for k, (c, _, _) in enumerate(members) if (c.ng, c.ny) == (256, 256)][:2]
    assert len(small) == 2
    for k in small:
        cfg, geom, y0 = members[k]
        sim = la.Simulation(cfg, geom)
        run = sim.integrator(y0, rtol=1e-6, atol=1e-6, t_out=0.1)
        g = run.advance(t_end)
        y_gpu = run.state()
        y_cpu, st, _ = orc.cpu_integrate(la.lib(), cfg, geom, y0, t_end, t_out=0.1, rtol=1e-6,
                                         atol=1e-6, mxstep=5000)
        dm = np.max(np.abs(y_gpu - y_cpu))
        print(f"[gpu] member {k} 256x256: max|dm| vs oracle-driven run {dm:.2e} "
              f"(nst {g['nst']} / {st['nst']})", flush=True)
        assert dm < 1e-5
        run.close()
        sim.close()


# ------------------------------------------------------- 4096 x 4096
def _config5(ng, ny, exec_, happ=(0.0, 0.0, 0.0)):
    """Polycrystal physics (i3 constants: window-weighted exchange, per-grain
    axes, DMI; fixed-axis Jv and preconditioner) with dead grains as hard
    holes, the zero-padded demag of strength 1, an applied field in every
    term set, regularised normalisation."""
    rhs = la.terms(che=4.0, chk=1.0, cha=0.0, chb=0.6, happ=happ, aniso=la.ANISO_LINEAR,
                   cell_axes=1, window=1, gate=la.GATE_MASK)
    fixed = la.terms(che=4.0, chk=4.0, cha=0.0, chb=0.3, msk=(1, 0, 0), nsk=(1, 0, 0),
                     happ=happ, aniso=la.ANISO_LINEAR, gate=la.GATE_MASK)
    return la.make_config(ng, ny, rhs, fixed, fixed, exec_=exec_,
                          normalize=la.NORM_REGULARIZED, demag=la.DEMAG_PADDED,
                          demag_strength=1.0)


def _i3_field_np(m, geom, hd, happ, che=4.0, chk=1.0, cha=0.0, chb=0.6):
    """h of the windowed polycrystal RHS, term by term from
    2d_i3/2d_fft.cu:576-597 (neighbours enter as w*m with periodic wrap),
    plus the applied field and the device's demag field."""
    ng, ny = geom.ng, geom.ny
    M = m.reshape(3, ny, ng)
    W = geom.w.reshape(ny, ng)
    MS = geom.msk.reshape(3, ny, ng)
    NS = geom.nsk.reshape(3, ny, ng)
    mdotmsk = (M * MS).sum(0)
    h = np.empty_like(M)
    for a in range(3):
        t = W * M[a]
        L, R = np.roll(t, 1, 1), np.roll(t, -1, 1)
        U, D = np.roll(t, 1, 0), np.roll(t, -1, 0)
        h[a] = (happ[a] + che * (L + R + U + D) + MS[a] * (chk * mdotmsk + cha)
                + chb * NS[a] * (L + R) + hd.reshape(3, ny, ng)[a])
    return h.reshape(3, -1)


def test_polycrystal_4096_compact_equals_dense_and_torque_identity():
    ng = ny = 4096
    geom = la.Geometry.polycrystal(ng, ny, num_grains=400, dead_frac=0.10, seed=5)
    n = geom.ncell
    dead = int((geom.active == 0).sum())
    print(f"[gpu] polycrystal 4096^2, 400 grains: {dead} hole cells "
          f"({100.0 * dead / n:.1f} %)", flush=True)
    assert 0 < dead < n
    happ = (-0.2, 0.0, 0.0)
    y0 = geom.ic_random_sphere(seed=3)
    dense = la.Simulation(_config5(ng, ny, la.EXEC_YMSK, happ), geom)
    comp = la.Simulation(_config5(ng, ny, la.EXEC_COMPACT, happ), geom)
    a, hids = comp.ids()
    assert np.array_equal(np.flatnonzero(geom.active), a)
    assert len(a) + len(hids) == n
    f_d, yn_d = dense.rhs(y0)
    f_c, yn_c = comp.rhs(y0)
    print(f"[gpu] f compact vs dense {rel(f_c, f_d):.2e}", flush=True)
    assert rel(yn_c, yn_d) < 1e-13 and rel(f_c, f_d) < 1e-13
    assert np.all(f_c.reshape(3, -1)[:, geom.active == 0] == 0.0)
    rng = np.random.default_rng(5)
    v = rng.standard_normal(3 * n)
    assert rel(comp.jtv(yn_d, v), dense.jtv(yn_d, v)) < 1e-13
    assert comp.psetup(yn_d, 0.05) == dense.psetup(yn_d, 0.05)
    r = rng.standard_normal(3 * n)
    assert rel(comp.psolve(r), dense.psolve(r)) < 1e-13
    # torque identity with the regularised |m| = |y| / (|y| + 1e-3) != 1
    hd = dense.hdemag()
    h = _i3_field_np(yn_d, geom, hd, happ)
    m = yn_d.reshape(3, -1)
    g = geom.active.astype(float)
    alpha = dense.cfg.alpha
    mh = (m * h).sum(0)
    expect = g * alpha * mh * (1.0 - (m * m).sum(0))
    mf = (m * f_d.reshape(3, -1)).sum(0)
    err = np.max(np.abs(mf - expect))
    print(f"[gpu] torque identity: max|m.f - g a (m.h)(1-|m|^2)| {err:.2e}, "
          f"max|expected| {np.max(np.abs(expect)):.2e}", flush=True)
    assert np.max(np.abs(expect)) > 1e-3
    assert err < 1e-11 * np.max(np.abs(h))
    comp.close()
    # two quasi-static stages of the field ramp at full size
    run = dense.integrator(y0, rtol=1e-4, atol=1e-4, t_out=0.05, mxstep=100000)
    fields = np.array([[-0.2, 0.0, 0.0], [-0.19, 0.0, 0.0]])
    mavg, st = run.ramp(fields, t_stage=0.1, which=la.TERMS_RHS | la.TERMS_JV | la.TERMS_PC)
    print(f"[gpu] 2-stage ramp at 4096^2: m_avg {mavg.tolist()}, nst={st['nst']} "
          f"nli={st['nli']} wall {st['wall_s']:.2f} s", flush=True)
    assert np.all(np.isfinite(mavg)) and st["nst"] > 0
    assert dense.applied_field(la.TERMS_JV) == (-0.19, 0.0, 0.0)
    y_end = run.state()
    assert np.array_equal(mavg[-1], dense.mean_magnetization(y_end))
    run.close()
    dense.close()


def test_applied_field_step_changes_f_by_the_zeeman_torque():
    """LLG_SetAppliedField: f(h + dh) - f(h) = g [chg dh x m + alpha (dh -
    (m.dh) m)] for a uniform dh (the RHS is affine in the field; the
    reference's precession term is (m3 h2 - m2 h3, ...) = h x m,
    2d_i3/2d_fft.cu:603-605)."""
    ng, ny = 96, 64
    geom = la.Geometry.polycrystal(ng, ny, num_grains=12, dead_frac=0.2, seed=9)
    sim = la.Simulation(_config5(ng, ny, la.EXEC_YMSK), geom)
    y = geom.ic_random_sphere(seed=2)
    f0, yn = sim.rhs(y)
    dh = np.array([0.13, -0.07, 0.02])
    sim.set_applied_field(dh, which=la.TERMS_RHS)
    assert sim.applied_field(la.TERMS_RHS) == tuple(dh)
    assert sim.applied_field(la.TERMS_PC) == (0.0, 0.0, 0.0)
    f1, _ = sim.rhs(y)
    m = yn.reshape(3, -1)
    g = geom.active.astype(float)
    cross = np.cross(dh, m.T).T
    expect = g * (sim.cfg.chg * cross + sim.cfg.alpha * (dh[:, None] - (m.T @ dh) * m))
    assert np.max(np.abs((f1 - f0).reshape(3, -1) - expect)) < 1e-12 * np.max(np.abs(f0))
    with pytest.raises(ValueError):
        sim.set_applied_field(dh, which=8)


def test_field_ramp_matches_oracle_ramp():
    ng, ny = 64, 48
    geom = la.Geometry.polycrystal(ng, ny, num_grains=8, dead_frac=0.2, seed=13)
    cfg = _config5(ng, ny, la.EXEC_YMSK)
    y0 = geom.ic_random_sphere(seed=4)
    fields = np.array([[-0.2, 0.0, 0.0], [-0.1, 0.0, 0.0], [0.0, 0.0, 0.0]])
    which = la.TERMS_RHS | la.TERMS_JV | la.TERMS_PC
    sim = la.Simulation(cfg, geom)
    run = sim.integrator(y0, rtol=1e-6, atol=1e-6, t_out=0.1, mxstep=100000)
    mavg, st = run.ramp(fields, t_stage=0.3, which=which)
    y_gpu = run.state()
    y_cpu, states = orc.cpu_ramp(la.lib(), cfg, geom, y0, fields, 0.3, which=which, t_out=0.1,
                                 rtol=1e-6, atol=1e-6)
    dm = np.max(np.abs(y_gpu - y_cpu))
    act = geom.active == 1
    m_cpu = [s.reshape(3, -1)[:, act].mean(1) for s in states]
    print(f"[gpu] 3-stage ramp 64x48: max|dm| vs oracle ramp {dm:.2e}; m_avg gpu "
          f"{mavg.tolist()} cpu {[x.tolist() for x in m_cpu]}", flush=True)
    assert dm < 1e-5
    assert np.max(np.abs(mavg - np.array(m_cpu))) < 1e-5
    run.close()
    sim.close()
