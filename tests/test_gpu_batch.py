"""One launch per pass over many instances (include/llg/llg_batch.h) and
the ensemble's batched mode.  Every instance of a batched evaluation must
get exactly the bits of its own call (same per-cell and per-transform
arithmetic), and a batched ensemble exactly the trajectories and
integrator statistics of the independent one.  No reference counterpart
(the reference runs one instance per process)."""
import ctypes as C

import numpy as np
import pytest

import llg_amd as la

pytestmark = pytest.mark.gpu


def _members():
    """Mixed shapes and physics: the periodic i4 demag on two grids of
    register-FFT plans (each plan shared by two instances with different
    fields and geometries), the windowed polycrystal (i3, dense mask) on one
    of those plans, the zero-padded demag on a register plan (two
    instances) and on a Stockham-only grid, an FP32-spectral demag, exact
    normalisation with DMI (i1, Stockham), and the Neumann plain form
    without demag.  The last four run their own passes inside the batch."""
    out = []
    for k, (ng, ny) in enumerate([(256, 128), (256, 128), (256, 256), (256, 256)]):
        cfg = la.preset("i4", ng=ng, ny=ny, demag=2.0)
        cfg.rhs.happ[0], cfg.rhs.happ[1] = 0.03 * (k + 1), -0.02 * k
        geom = la.Geometry.ellipse(ng, ny, 0.3 + 0.05 * k, 0.35) if k % 2 else \
            la.Geometry.ring(ng, ny, 0.9, 0.95, 0.4, 0.5)
        out.append((cfg, geom, None))
    geom = la.Geometry.polycrystal(256, 128, num_grains=9, dead_frac=0.2, seed=7)
    out.append((la.preset("i3", ng=256, ny=128), geom, None))
    for k in range(2):
        cfg = la.preset("square_hole", ng=128, ny=64, demag=4.0)
        cfg.rhs.happ[1] = 0.05 * k
        out.append((cfg, la.Geometry.square_hole(128, 64, halfside_frac=0.2 + 0.05 * k), None))
    cfg = la.preset("square_hole", ng=48, ny=40, demag=4.0)
    out.append((cfg, la.Geometry.square_hole(48, 40), None))
    cfg = la.preset("i4", ng=256, ny=128, demag=2.0)
    out.append((cfg, la.Geometry.ellipse(256, 128, 0.4, 0.3), "fp32"))
    out.append((la.preset("i1", ng=80, ny=32), la.Geometry.full(80, 32), None))
    out.append((la.preset("plain128", ng=40, ny=36), la.Geometry.full(40, 36), None))
    return out


def _sims():
    sims = []
    for cfg, geom, tag in _members():
        s = la.Simulation(cfg, geom)
        if tag == "fp32":
            assert la.lib().Demag_SetPrecision(la.lib().LLG_GetDemag(s.ctx), 1) == 0
        sims.append(s)
    return sims


def _state(sim, seed):
    rng = np.random.default_rng(seed)
    y = rng.normal(size=sim.neq)
    m = y.reshape(3, -1)
    m /= np.linalg.norm(m, axis=0) * rng.uniform(0.9, 1.1, size=m.shape[1])
    return y


def test_rhs_batch_bitwise_equals_single_calls():
    sims = _sims()
    n = len(sims)
    ys = [_state(s, 100 + i) for i, s in enumerate(sims)]
    single = []
    for s, y in zip(sims, ys):
        f, yn = s.rhs(y)
        single.append((f, yn, s.hdemag()))
    batch = la.Batch(sims)
    vy = [s.vec(y) for s, y in zip(sims, ys)]
    vf = [s.vec(np.full(s.neq, 7.0)) for s in sims]
    skip = 2
    vy_skip = [v if i != skip else None for i, v in enumerate(vy)]
    batch.rhs(vy_skip, vf)
    for i, s in enumerate(sims):
        if i == skip:
            assert np.all(vf[i].get() == 7.0)      # not evaluated
            continue
        f, yn, hd = single[i]
        assert np.array_equal(vf[i].get(), f), i
        assert np.array_equal(vy[i].get(), yn), i       # normalised in place as f() does
        assert np.array_equal(s.hdemag(), hd), i
    batch.close()
    for s in sims:
        s.close()


def test_rhs_batch_residual_form_bitwise():
    """Members in the Newton-residual form (LLG_RhsResidual) and in the plain
    form in one call."""
    L = la.lib()
    sims = _sims()
    ys = [_state(s, 200 + i) for i, s in enumerate(sims)]
    rng = np.random.default_rng(5)
    zn1 = [s.vec(rng.normal(size=s.neq)) for s in sims]
    ycor = [s.vec(rng.normal(size=s.neq)) for s in sims]
    coef = [(-0.7 - 0.01 * i, -1.0, 0.013 * (i + 1)) for i in range(len(sims))]
    use_res = [i % 2 == 0 for i in range(len(sims))]
    expect = []
    for i, (s, y) in enumerate(zip(sims, ys)):
        vy, vr = s.vec(y), s.vec()
        if use_res[i]:
            c = (C.c_double * 3)(*coef[i])
            assert L.LLG_RhsResidual(0.0, vy.v, c, zn1[i].v, ycor[i].v, vr.v, s.ctx) == 0
        else:
            assert L.LLG_Rhs(0.0, vy.v, vr.v, s.ctx) == 0
        expect.append(vr.get())
    batch = la.Batch(sims)
    vy = [s.vec(y) for s, y in zip(sims, ys)]
    vo = [s.vec() for s in sims]
    res = [(zn1[i], ycor[i], *coef[i]) if use_res[i] else None for i in range(len(sims))]
    batch.rhs(vy, vo, res=res)
    for i in range(len(sims)):
        assert np.array_equal(vo[i].get(), expect[i]), i
    batch.close()


def test_jtv_batch_bitwise_equals_single_calls():
    sims = _sims()
    ys = [_state(s, 300 + i) for i, s in enumerate(sims)]
    rng = np.random.default_rng(9)
    vs = [rng.normal(size=s.neq) for s in sims]
    expect = [s.jtv(y, v) for s, y, v in zip(sims, ys, vs)]
    batch = la.Batch(sims)
    vy = [s.vec(y) for s, y in zip(sims, ys)]
    vv = [s.vec(v) for s, v in zip(sims, vs)]
    vj = [s.vec() for s in sims]
    batch.jtv(vv, vj, vy)
    for i in range(len(sims)):
        assert np.array_equal(vj[i].get(), expect[i]), i
    batch.close()


def test_batch_rejects_compact_and_sees_a_new_applied_field():
    L = la.lib()
    cfg = la.preset("i6", ng=64, ny=48)
    s = la.Simulation(cfg, la.Geometry.ellipse(64, 48))
    with pytest.raises(RuntimeError):
        la.Batch([s])
    s.close()
    sims = [la.Simulation(la.preset("i4", ng=64, ny=48, demag=2.0),
                          la.Geometry.ellipse(64, 48, 0.3 + 0.1 * k, 0.3)) for k in range(2)]
    batch = la.Batch(sims)
    y = _state(sims[0], 1)
    sims[1].set_applied_field((0.05, -0.04, 0.02))
    expect = [s.rhs(y)[0] for s in sims]
    vy = [s.vec(y) for s in sims]
    vo = [s.vec() for s in sims]
    batch.rhs(vy, vo)
    for i in range(2):
        assert np.array_equal(vo[i].get(), expect[i]), i
    batch.close()


def test_ensemble_batched_equals_independent_ensemble():
    members = la.ensemble_members(n=7, seed=21, grids=((256, 128), (256, 256)),
                                  frac=(0.3, 0.7))
    kw = dict(rtol=1e-4, atol=1e-4)
    ref = la.Ensemble(members, **kw)
    ref.advance(0.3)
    ref.advance(0.6)
    ens = la.Ensemble(members, **kw)
    assert ens.set_batched(1) == 2
    ens.advance(0.3)
    out = ens.advance(0.6)
    st = ens.batch_stats()
    print(f"[gpu] batched ensemble: {out}, {st}", flush=True)
    assert out["failed"] == 0
    assert st["launches"] < st["calls"]          # calls shared launches
    for i in range(len(members)):
        a, b = ens.result(i), ref.result(i)
        assert (a["nst"], a["nfe"], a["nli"], a["nni"]) == (b["nst"], b["nfe"], b["nli"], b["nni"]), i
        assert np.array_equal(ens.state(i), ref.state(i)), i
    # back to independent evaluation
    assert ens.set_batched(0) == 0
    ens.advance(0.8)
    ref.advance(0.8)
    for i in range(len(members)):
        assert np.array_equal(ens.state(i), ref.state(i)), i
    ens.close()
    ref.close()


def test_ensemble_batched_small_shapes_only():
    """mode 2: the small grid's instances evaluate f together, the large
    grid's on their own; trajectories unchanged."""
    members = la.ensemble_members(n=5, seed=5, grids=((256, 128), (512, 512)), frac=(0.3, 0.7))
    kw = dict(rtol=1e-4, atol=1e-4)
    ref = la.Ensemble(members, **kw)
    ref.advance(0.3)
    ens = la.Ensemble(members, **kw)
    assert ens.set_batched(2) == 1
    ens.advance(0.3)
    st = ens.batch_stats()
    assert 0 < st["launches"] <= st["calls"]
    for i in range(len(members)):
        assert ens.result(i)["nst"] == ref.result(i)["nst"], i
        assert np.array_equal(ens.state(i), ref.state(i)), i
    ens.close()
    ref.close()
