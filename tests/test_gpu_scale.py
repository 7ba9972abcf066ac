"""GPU checks at the full sizes of the scaling workloads, through properties
that do not depend on the grid size (the oracle cannot run at these sizes
in a test's time).

- 4096 x 4096 cells, zero-padded 8192 x 8192 demag (column transforms of
  8192 points, register-resident kernels only): the padded demag is an
  open-boundary convolution, so the field of a magnetised patch near the
  patch does not depend on how far the grid extends.  The same patch in a
  256 x 256 padded grid (covered by the oracle and the Stockham kernels)
  must give the same field in the overlapping window: 1e-12 relative in
  FP64, 2e-5 in the FP32 spectral form.
- 1024 x 1024 Voronoi polycrystal (72 grains, 16 % dead, per-grain axes,
  window-weighted exchange, windowed demag 1.0): the compact active-cell
  execution (id lists) and the dense execution gated by the active flag
  evaluate the same operator: f, Jv and the preconditioner solve agree to
  1e-13 relative, hole rows are exactly zero; id lists are ascending and
  partition the grid.
"""
import ctypes as C

import numpy as np
import pytest

import llg_amd as la
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def rel(a, b):
    d = np.max(np.abs(a - b))
    s = np.max(np.abs(b))
    return d / s if s > 0 else d


def _field(L, d, nx, ny, y):
    n3 = 3 * nx * ny
    vy = la.HipVector(n3, None, y)
    vh = la.HipVector(n3, None)
    L.Demag_SetStream(d, L.N_VGetStream_Hip(vy.v))
    L.Demag_Apply(d, L.N_VGetDeviceArrayPointer_Hip(vy.v), L.N_VGetDeviceArrayPointer_Hip(vh.v))
    return vh.get().reshape(3, ny, nx)


def test_padded_demag_4096_is_translation_invariant():
    L = la.lib()
    big, small = 4096, 256
    pw, ph = 64, 48                 # patch size
    ob = (1000, 2000)               # patch origin in the big grid (x, y)
    os_ = (100, 120)                # and in the small grid
    rng = np.random.default_rng(40)
    patch = rng.standard_normal((3, ph, pw))
    ys = np.zeros((3, small, small))
    ys[:, os_[1]:os_[1] + ph, os_[0]:os_[0] + pw] = patch
    yb = np.zeros((3, big, big))
    yb[:, ob[1]:ob[1] + ph, ob[0]:ob[0] + pw] = patch
    dx, dy = ob[0] - os_[0], ob[1] - os_[1]

    # small grid: oracle FFT demag and the library
    ds = L.Demag_InitPadded(small, small, 1.0, 1.0)
    assert ds
    h_small = _field(L, ds, small, small, ys.ravel())
    h_orc = orc.Demag(small, small, 1.0, 1.0, padded=True).apply(ys.ravel()).reshape(3, small,
                                                                                        small)
    L.Demag_Destroy(ds)
    print(f"[gpu] 256^2 padded vs oracle: {rel(h_small, h_orc):.2e}", flush=True)
    assert rel(h_small, h_orc) < 1e-12

    db = L.Demag_InitPadded(big, big, 1.0, 1.0)
    assert db, L.Demag_LastError()
    px, py, pad = C.c_int(), C.c_int(), C.c_int()
    L.Demag_GetPlanShape(db, C.byref(px), C.byref(py), C.byref(pad))
    assert (px.value, py.value, pad.value) == (8192, 8192, 1)
    # 8192-point columns only exist as register kernels
    assert L.Demag_SetFftKernels(db, -1, -1) == -1
    for fp32 in (0, 1):
        assert L.Demag_SetPrecision(db, fp32) == 0
        for cv in (0, 1):
            assert L.Demag_SetFftKernels(db, 0, cv) == 0
            hb = _field(L, db, big, big, yb.ravel())
            win = hb[:, dy:dy + small, dx:dx + small]
            e = rel(win, h_orc)
            # nothing outside reaches a magnitude the window would miss
            print(f"[gpu] 4096^2 padded fp32={fp32} col variant {cv}: window vs 256^2 oracle "
                  f"{e:.2e}, max|h| {np.max(np.abs(hb)):.3e}", flush=True)
            assert e < (2e-5 if fp32 else 1e-12)
            assert np.all(np.isfinite(hb))
    L.Demag_Destroy(db)


def _config3(exec_):
    """1024^2 polycrystal workload: i3 physics (window-weighted exchange, DMI,
    per-grain uniaxial axes, windowed demag 1.0, regularised normalisation)
    with dead grains as hard holes.  Compact execution skips the holes; the
    dense-mask equivalent gates the torque by the active flag (the i3
    preset's soft-window gate is a different operator)."""
    rhs = la.terms(che=4.0, chk=1.0, cha=0.0, chb=0.6, aniso=la.ANISO_LINEAR, cell_axes=1,
                   window=1, gate=la.GATE_MASK)
    fixed = la.terms(che=4.0, chk=4.0, cha=0.0, chb=0.3, msk=(1, 0, 0), nsk=(1, 0, 0),
                     aniso=la.ANISO_LINEAR, gate=la.GATE_MASK)
    return la.make_config(1024, 1024, rhs, fixed, fixed, exec_=exec_,
                          normalize=la.NORM_REGULARIZED, demag=la.DEMAG_PERIODIC,
                          demag_strength=1.0)


def test_polycrystal_1024_compact_equals_dense():
    geom = la.Geometry.polycrystal(1024, 1024, num_grains=72, dead_frac=0.16)
    n = geom.ncell
    dead = int((geom.active == 0).sum())
    print(f"[gpu] polycrystal 1024^2: {dead} hole cells", flush=True)
    assert 0 < dead < n
    y0 = geom.ic_random_sphere(seed=3)
    dense = la.Simulation(_config3(la.EXEC_YMSK), geom)
    comp = la.Simulation(_config3(la.EXEC_COMPACT), geom)
    a, h = comp.ids()
    assert np.all(np.diff(a) > 0) and np.all(np.diff(h) > 0)
    assert len(a) + len(h) == n
    assert np.array_equal(np.sort(np.concatenate([a, h])), np.arange(n))
    assert np.array_equal(np.flatnonzero(geom.active), a)
    f_d, yn_d = dense.rhs(y0)
    f_c, yn_c = comp.rhs(y0)
    print(f"[gpu] f compact vs dense {rel(f_c, f_d):.2e}", flush=True)
    assert rel(yn_c, yn_d) < 1e-13
    assert rel(f_c, f_d) < 1e-13
    assert np.all(f_c.reshape(3, -1)[:, geom.active == 0] == 0.0)
    v = np.random.default_rng(5).standard_normal(3 * n)
    assert rel(comp.jtv(yn_d, v), dense.jtv(yn_d, v)) < 1e-13
    gamma = 0.05
    assert comp.psetup(yn_d, gamma) == dense.psetup(yn_d, gamma)
    r = np.random.default_rng(6).standard_normal(3 * n)
    assert rel(comp.psolve(r), dense.psolve(r)) < 1e-13
    # torque invariant: m . f = 0 on the unit sphere for the windowed RHS
    m = yn_d.reshape(3, -1)
    act = geom.active == 1
    mf = np.abs(np.sum(m * f_d.reshape(3, -1), axis=0))[act]
    print(f"[gpu] max |m.f| {mf.max():.2e}", flush=True)
