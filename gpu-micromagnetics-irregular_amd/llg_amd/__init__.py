"""llg_amd -- MI355X-native 2D LLG hot path (CVODE plug-ins on gfx950).

Host-side mirror of the reference's plug-in surface: configurations that
reproduce each reference variant's physics (configs/sim_config.h knobs,
per-variant constants of src/irregular/2d_fft/2d_i{1..6}), geometry/IC
builders, single-callback entry points (f, Jv, PrecondSetup/Solve,
Demag_Apply) and the timed integration loop.  All compute runs in
libllg_amd.so; this module only marshals arguments.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import native
from .native import (llg_config, llg_terms, llg_run_opts, llg_run_result, llg_grain,
                     llg_instance, llg_ensemble_result, llg_res_in,
                     EXEC_YMSK, EXEC_COMPACT, ANISO_LINEAR, ANISO_CUBIC, NORM_NONE,
                     NORM_EXACT, NORM_REGULARIZED, BC_PERIODIC, BC_NEUMANN, DEMAG_OFF,
                     DEMAG_PERIODIC, DEMAG_PADDED, GATE_NONE, GATE_MASK, GATE_WINDOW,
                     PC_ALWAYS, PC_JOK, PREC_LEFT, PREC_RIGHT, MODIFIED_GS, CLASSICAL_GS,
                     TERMS_RHS, TERMS_JV, TERMS_PC)

__all__ = [
    "native", "lib", "terms", "make_config", "preset", "PRESETS", "Geometry", "Simulation",
    "HipVector", "llg_config", "llg_terms", "llg_run_opts",
]


def lib():
    return native.load()


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _ucp(a):
    return a.ctypes.data_as(C.POINTER(C.c_ubyte)) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int)) if a is not None else None


# --------------------------------------------------------------- configs
def terms(che=0.0, chk=0.0, cha=0.0, chb=0.0, msk=(0, 0, 0), nsk=(0, 0, 0), happ=(0, 0, 0),
          aniso=ANISO_CUBIC, cell_axes=0, window=0, gate=GATE_NONE) -> llg_terms:
    t = llg_terms()
    t.che, t.chk, t.cha, t.chb = che, chk, cha, chb
    for a in range(3):
        t.msk[a], t.nsk[a], t.happ[a] = msk[a], nsk[a], happ[a]
    t.aniso, t.cell_axes, t.window, t.gate = aniso, cell_axes, window, gate
    return t


def make_config(ng, ny, rhs, jv=None, pc=None, exec_=EXEC_YMSK, boundary=BC_PERIODIC,
                exch_self=0, normalize=NORM_NONE, demag=DEMAG_OFF, demag_strength=0.0,
                demag_thick=1.0, demag_fp32=0, alpha=0.2, chg=1.0, pc_mode=PC_ALWAYS):
    c = llg_config()
    c.ng, c.ny, c.exec_, c.boundary = ng, ny, exec_, boundary
    c.exch_self, c.normalize, c.demag, c.demag_fp32 = exch_self, normalize, demag, demag_fp32
    c.pc_mode = pc_mode
    c.demag_strength, c.demag_thick, c.alpha, c.chg = demag_strength, demag_thick, alpha, chg
    c.rhs = rhs
    c.jv = jv if jv is not None else rhs
    c.pc = pc if pc is not None else rhs
    return c


def _preset_i4(ng=512, ny=512, demag=4.0, demag_mode=DEMAG_PERIODIC, **kw):
    """2d_i4: square hole, ymsk, Zeeman (-0.2,0,0); Jv without Zeeman/demag;
    preconditioner field with pc_happ1 = -1.0 (2d_fft.cu:175-180, jtv.cu:36-39,
    precond.cu:64-70)."""
    cub = dict(che=20.0, chk=1.0, msk=(1, 1, 1), nsk=(1, 0, 0), aniso=ANISO_CUBIC, gate=GATE_MASK)
    return make_config(ng, ny, terms(happ=(-0.2, 0, 0), **cub), terms(**cub),
                       terms(happ=(-1.0, 0, 0), **cub), demag=demag_mode if demag > 0 else DEMAG_OFF,
                       demag_strength=demag, **kw)


def _preset_i2(ng=512, ny=512, demag=4.0, **kw):
    cub = dict(che=50.0, chk=1.0, msk=(1, 1, 1), nsk=(1, 0, 0), aniso=ANISO_CUBIC, gate=GATE_MASK)
    return make_config(ng, ny, terms(**cub), normalize=NORM_REGULARIZED,
                       demag=DEMAG_PERIODIC if demag > 0 else DEMAG_OFF, demag_strength=demag, **kw)


def _preset_i5(ng=512, ny=512, demag=4.0, **kw):
    cub = dict(che=10.0, chk=1.0, msk=(1, 1, 1), nsk=(1, 0, 0), aniso=ANISO_CUBIC, gate=GATE_MASK)
    return make_config(ng, ny, terms(**cub), demag=DEMAG_PERIODIC if demag > 0 else DEMAG_OFF,
                       demag_strength=demag, **kw)


def _preset_i6(ng=512, ny=512, demag=4.0, **kw):
    cub = dict(che=10.0, chk=1.0, msk=(1, 1, 1), nsk=(1, 0, 0), aniso=ANISO_CUBIC, gate=GATE_NONE)
    return make_config(ng, ny, terms(**cub), exec_=EXEC_COMPACT,
                       demag=DEMAG_PERIODIC if demag > 0 else DEMAG_OFF, demag_strength=demag, **kw)


def _preset_i1(ng=512, ny=512, demag=2.0, **kw):
    lin = dict(che=4.0, chk=4.0, cha=0.0, chb=0.3, msk=(1, 0, 0), nsk=(1, 0, 0),
               aniso=ANISO_LINEAR, gate=GATE_NONE)
    return make_config(ng, ny, terms(**lin), normalize=NORM_EXACT,
                       demag=DEMAG_PERIODIC if demag > 0 else DEMAG_OFF, demag_strength=demag, **kw)


def _preset_i3(ng=768, ny=256, demag=1.0, **kw):
    rhs = terms(che=4.0, chk=1.0, cha=0.0, chb=0.6, aniso=ANISO_LINEAR, cell_axes=1, window=1,
                gate=GATE_WINDOW)
    fixed = terms(che=4.0, chk=4.0, cha=0.0, chb=0.3, msk=(1, 0, 0), nsk=(1, 0, 0),
                  aniso=ANISO_LINEAR, gate=GATE_NONE)
    return make_config(ng, ny, rhs, fixed, fixed, normalize=NORM_REGULARIZED,
                       demag=DEMAG_PERIODIC if demag > 0 else DEMAG_OFF, demag_strength=demag, **kw)


def _preset_plain128(ng=128, ny=128, a=0.02, c_ex=1.0, k_u=0.1, h_ext=(0.0, 0.0, -0.05), **kw):
    """Plain CPU-runnable workload: LL form -(1/(1+a^2))[m x h + a m x (m x h)]
    mapped onto the reference form with chg = 1/(1+a^2), alpha = a/(1+a^2);
    h = c_ex * Laplacian(m) (Neumann) + k_u (m.z) z + h_ext; no demag."""
    t = terms(che=c_ex, chk=k_u, cha=0.0, chb=0.0, msk=(0, 0, 1), happ=h_ext,
              aniso=ANISO_LINEAR, gate=GATE_NONE)
    return make_config(ng, ny, t, boundary=BC_NEUMANN, exch_self=1, alpha=a / (1 + a * a),
                       chg=1.0 / (1 + a * a), **kw)


def _preset_square_hole(ng=1536, ny=512, demag=4.0, padded=True, **kw):
    """1536x512-cell square hole (444^2 cells, 25.07 %), i4 physics constants,
    zero-padded 3072x1024 FFT demag of strength 4, thickness 1."""
    return _preset_i4(ng, ny, demag=demag, demag_mode=DEMAG_PADDED if padded else DEMAG_PERIODIC,
                      **kw)


def _preset_polycrystal(ng=1024, ny=1024, demag=1.0, compact=True, padded=False, happ=(0, 0, 0),
                        **kw):
    """The polycrystal workloads (1024^2, 72 grains, 16 % dead; 4096^2, 400
    grains, 10 % dead): the i3 constants (window-weighted exchange 4, per-
    grain uniaxial anisotropy 1, DMI 0.6 on the x neighbours; the Jv and
    preconditioner mirrors' fixed-axis set: anisotropy 4, DMI 0.3,
    2d_i3/2d_fft.cu:176-181, jtv.cu:30-35, precond.cu:58-63), dead grains as
    hard holes (compact active-cell lists, or the byte mask), windowed FFT
    demag, regularised normalisation (2d_i3/2d_fft.cu:418-428); an applied
    field in all three term sets for field sweeps."""
    rhs = terms(che=4.0, chk=1.0, cha=0.0, chb=0.6, happ=happ, aniso=ANISO_LINEAR, cell_axes=1,
                window=1, gate=GATE_MASK)
    fixed = terms(che=4.0, chk=4.0, cha=0.0, chb=0.3, msk=(1, 0, 0), nsk=(1, 0, 0), happ=happ,
                  aniso=ANISO_LINEAR, gate=GATE_MASK)
    mode = DEMAG_PADDED if padded else DEMAG_PERIODIC
    return make_config(ng, ny, rhs, fixed, fixed, exec_=EXEC_COMPACT if compact else EXEC_YMSK,
                       normalize=NORM_REGULARIZED, demag=mode if demag > 0 else DEMAG_OFF,
                       demag_strength=demag, **kw)


PRESETS = {
    "i1": _preset_i1, "i2": _preset_i2, "i3": _preset_i3, "i4": _preset_i4,
    "i5": _preset_i5, "i6": _preset_i6, "plain128": _preset_plain128,
    "square_hole": _preset_square_hole, "polycrystal": _preset_polycrystal,
}


def preset(name, **kw) -> llg_config:
    return PRESETS[name](**kw)


# -------------------------------------------------------------- geometry
class Geometry:
    """Active mask (+ optional window and per-cell axes) of an ng x ny grid."""

    def __init__(self, ng, ny, active=None, w=None, msk=None, nsk=None, grain_id=None,
                 grains=None):
        self.ng, self.ny = ng, ny
        self.active = active
        self.w, self.msk, self.nsk = w, msk, nsk
        self.grain_id, self.grains = grain_id, grains

    @property
    def ncell(self):
        return self.ng * self.ny

    @classmethod
    def full(cls, ng, ny):
        return cls(ng, ny, np.ones(ng * ny, np.uint8))

    @classmethod
    def square_hole(cls, ng, ny, cx_frac=0.5, cy_frac=0.5, halfside_frac=0.22):
        a = np.zeros(ng * ny, np.uint8)
        lib().LLG_GeomSquareHole(ng, ny, cx_frac, cy_frac, halfside_frac, _ucp(a))
        return cls(ng, ny, a)

    @classmethod
    def rectangle(cls, ng, ny, w_frac=0.5, h_frac=0.5):
        a = np.zeros(ng * ny, np.uint8)
        lib().LLG_GeomRectangle(ng, ny, w_frac, h_frac, _ucp(a))
        return cls(ng, ny, a)

    @classmethod
    def ring(cls, ng, ny, w_frac=0.5, h_frac=0.5, inner_w=0.5, inner_h=0.5):
        a = np.zeros(ng * ny, np.uint8)
        lib().LLG_GeomRing(ng, ny, w_frac, h_frac, inner_w, inner_h, _ucp(a))
        return cls(ng, ny, a)

    @classmethod
    def ellipse(cls, ng, ny, rx_frac=0.25, ry_frac=0.25):
        a = np.zeros(ng * ny, np.uint8)
        lib().LLG_GeomEllipse(ng, ny, rx_frac, ry_frac, _ucp(a))
        return cls(ng, ny, a)

    @classmethod
    def polycrystal(cls, ng, ny, num_grains=72, dead_frac=0.16, seed=20251104, z_bias=1.6,
                    eps_cells=2.2):
        N = ng * ny
        grains = (llg_grain * num_grains)()
        gid = np.zeros(N, np.int32)
        a = np.zeros(N, np.uint8)
        w = np.zeros(N)
        msk = np.zeros(3 * N)
        nsk = np.zeros(3 * N)
        lib().LLG_GeomPolycrystal(ng, ny, num_grains, dead_frac, seed, z_bias, eps_cells, grains,
                                  _ip(gid), _ucp(a), _dp(w), _dp(msk), _dp(nsk))
        return cls(ng, ny, a, w, msk, nsk, gid, grains)

    # ---- initial states
    def ic_uniform(self, m=(1.0, -0.0175, 0.0)):
        y = np.zeros(3 * self.ncell)
        lib().LLG_ICUniform(self.ng, self.ny, _ucp(self.active), m[0], m[1], m[2], _dp(y))
        return y

    def ic_two_domain(self, j_split=None, tail_my=-0.0175):
        y = np.zeros(3 * self.ncell)
        js = (7 * self.ny) // 8 if j_split is None else j_split
        lib().LLG_ICTwoDomain(self.ng, self.ny, _ucp(self.active), js, tail_my, _dp(y))
        return y

    def ic_stripes(self, left=0.25, right=0.75, eps=0.01, seed=12345):
        y = np.zeros(3 * self.ncell)
        lib().LLG_ICStripes(self.ng, self.ny, left, right, eps, seed, _dp(y))
        return y

    def ic_grain_bumps(self, core_mz=0.95, seed=12345):
        y = np.zeros(3 * self.ncell)
        lib().LLG_ICGrainBumps(self.ng, self.ny, self.grains, _ip(self.grain_id), _dp(self.w),
                               core_mz, seed, _dp(y))
        return y

    def ic_tilted_z(self, max_angle=0.1, seed=1):
        y = np.zeros(3 * self.ncell)
        lib().LLG_ICTiltedZ(self.ng, self.ny, _ucp(self.active), max_angle, seed, _dp(y))
        return y

    def ic_split_x(self, max_angle=0.05, seed=1):
        y = np.zeros(3 * self.ncell)
        lib().LLG_ICSplitX(self.ng, self.ny, _ucp(self.active), max_angle, seed, _dp(y))
        return y

    def ic_random_sphere(self, seed=1):
        y = np.zeros(3 * self.ncell)
        lib().LLG_ICRandomSphere(self.ng, self.ny, _ucp(self.active), seed, _dp(y))
        return y


# ---------------------------------------------------------------- vectors
class _SunCtx(C.Structure):
    _fields_ = [("in_use", C.c_int)]


_SUNCTX = _SunCtx()


class HipVector:
    """Owning handle of an N_Vector_Hip with numpy transfer helpers."""

    def __init__(self, n, stream=None, values=None):
        L = lib()
        self.n = n
        self.v = L.N_VNewWithStream_Hip(n, stream, C.byref(_SUNCTX))
        if not self.v:
            raise RuntimeError("N_VNewWithStream_Hip failed")
        if values is not None:
            self.set(values)

    def set(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.n
        if lib().N_VSetFromHost_Hip(self.v, _dp(a)) != 0:
            raise RuntimeError("host->device copy failed")

    def get(self):
        a = np.empty(self.n)
        if lib().N_VGetToHost_Hip(self.v, _dp(a)) != 0:
            raise RuntimeError("device->host copy failed")
        return a

    def fused(self, on=True):
        lib().N_VEnableFusedOps_Hip(self.v, 1 if on else 0)
        return self

    def __del__(self):
        try:
            if self.v:
                lib().N_VDestroy_Hip(self.v)
                self.v = None
        except Exception:
            pass


def _vp_array(vs):
    arr = (C.c_void_p * len(vs))(*[v.v for v in vs])
    return arr


# ------------------------------------------------------------- simulation
class Simulation:
    """A context (the reference's UserData) plus single-callback access and
    the timed integration loop."""

    def __init__(self, cfg: llg_config, geom: Geometry | None = None):
        L = lib()
        self.cfg = cfg
        self.geom = geom or Geometry.full(cfg.ng, cfg.ny)
        g = self.geom
        self.ctx = L.LLG_Create(C.byref(cfg), _ucp(g.active), _dp(g.w), _dp(g.msk), _dp(g.nsk),
                                None)
        if not self.ctx:
            raise RuntimeError("LLG_Create failed: " + L.LLG_LastError().decode())
        self.ncell = cfg.ng * cfg.ny
        self.neq = 3 * self.ncell
        self.stream = L.LLG_GetStream(self.ctx)

    def close(self):
        if getattr(self, "ctx", None):
            lib().LLG_Destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def vec(self, values=None):
        return HipVector(self.neq, self.stream, values)

    # --- single callbacks through the C ABI
    def rhs(self, y, t=0.0):
        vy, vf = self.vec(y), self.vec()
        rc = lib().LLG_Rhs(t, vy.v, vf.v, self.ctx)
        if rc != 0:
            raise RuntimeError("LLG_Rhs failed: " + lib().LLG_LastError().decode())
        return vf.get(), vy.get()

    def jtv(self, y, v, t=0.0):
        vy, vv, vj = self.vec(y), self.vec(v), self.vec()
        rc = lib().JtvProduct(vv.v, vj.v, t, vy.v, None, self.ctx, None)
        if rc != 0:
            raise RuntimeError("JtvProduct failed")
        return vj.get()

    def psetup(self, y, gamma, jok=0, t=0.0):
        vy = self.vec(y)
        jcur = C.c_int(0)
        rc = lib().PrecondSetup(t, vy.v, None, jok, C.byref(jcur), gamma, self.ctx)
        if rc != 0:
            raise RuntimeError("PrecondSetup failed")
        return jcur.value

    def psolve(self, r, gamma=0.0, t=0.0):
        vr, vz = self.vec(r), self.vec()
        rc = lib().PrecondSolve(t, None, None, vr.v, vz.v, gamma, 0.0, PREC_LEFT, self.ctx)
        if rc != 0:
            raise RuntimeError("PrecondSolve failed")
        return vz.get()

    def hdemag(self):
        h = np.empty(self.neq)
        lib().LLG_GetHdemag(self.ctx, _dp(h))
        return h

    def pinv(self):
        p = np.empty(9 * self.ncell)
        lib().LLG_GetPinv(self.ctx, _dp(p))
        return p

    def mask(self):
        m = np.empty(self.neq)
        lib().LLG_GetMask(self.ctx, _dp(m))
        return m

    def ids(self):
        nc, na, ni = C.c_int(), C.c_int(), C.c_int()
        lib().LLG_GetCounts(self.ctx, C.byref(nc), C.byref(na), C.byref(ni))
        a = np.empty(na.value, np.int32)
        h = np.empty(ni.value, np.int32)
        lib().LLG_GetActiveIds(self.ctx, _ip(a), _ip(h))
        return a, h

    def set_applied_field(self, happ, which=TERMS_RHS | TERMS_PC):
        """LLG_SetAppliedField: the Zeeman field of the term sets in `which`"""
        h = (C.c_double * 3)(*happ)
        if lib().LLG_SetAppliedField(self.ctx, which, h) != 0:
            raise ValueError("LLG_SetAppliedField: " + lib().LLG_LastError().decode())

    def applied_field(self, which=TERMS_RHS):
        h = (C.c_double * 3)()
        if lib().LLG_GetAppliedField(self.ctx, which, h) != 0:
            raise ValueError("LLG_GetAppliedField: one term set per call")
        return tuple(h)

    def mean_magnetization(self, y):
        v = self.vec(y)
        m = (C.c_double * 3)()
        if lib().LLG_MeanMagnetization(self.ctx, v.v, m) != 0:
            raise RuntimeError("LLG_MeanMagnetization failed")
        return np.array(m)

    def self_coupling(self):
        d = lib().LLG_GetDemag(self.ctx)
        v = (C.c_double * 3)()
        lib().Demag_GetSelfCoupling(d, C.byref(v, 0), C.byref(v, 8), C.byref(v, 16))
        return tuple(v)

    # --- integration
    @staticmethod
    def run_opts(**kw) -> llg_run_opts:
        o = llg_run_opts()
        lib().LLG_DefaultRunOpts(C.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        return o

    def integrator(self, y0, **kw):
        return Run(self, y0, self.run_opts(**kw))


class Run:
    """The reference's CVode output loop (T1 cadence) on this context."""

    def __init__(self, sim: Simulation, y0, opts: llg_run_opts):
        self.sim = sim
        y0 = np.ascontiguousarray(y0, dtype=np.float64)
        self.h = lib().LLG_RunCreate(sim.ctx, C.byref(opts), _dp(y0))
        if not self.h:
            raise RuntimeError("LLG_RunCreate failed")
        self.opts = opts

    def advance(self, t_end):
        res = llg_run_result()
        rc = lib().LLG_RunAdvance(self.h, t_end, C.byref(res))
        out = {
            "retval": res.retval, "wall_s": res.wall_s, "gpu_ms": res.gpu_ms,
            "sim_per_wall": res.sim_per_wall, "t": res.t_reached, "nout": res.nout,
            **res.st.as_dict(),
        }
        if rc != 0:
            raise RuntimeError(f"integration failed (retval {res.retval}) at t={res.t_reached}")
        return out

    def state(self):
        y = np.empty(self.sim.neq)
        lib().LLG_RunGetState(self.h, _dp(y))
        return y

    def stage(self, happ, t_stage, which=TERMS_RHS | TERMS_PC):
        """LLG_RunStage: set the applied field, re-initialise the integrator
        at the current state and advance t_stage; returns the statistics
        and the stage's mean magnetisation"""
        h = np.ascontiguousarray(happ, dtype=np.float64).reshape(3)
        m = np.zeros(3)
        res = llg_run_result()
        rc = lib().LLG_RunStage(self.h, which, _dp(h), t_stage, C.byref(res), _dp(m))
        out = {"retval": res.retval, "wall_s": res.wall_s, "t": res.t_reached,
               "sim_per_wall": res.sim_per_wall, **res.st.as_dict()}
        if rc != 0:
            raise RuntimeError(f"stage failed (rc {rc}) at t={res.t_reached}")
        return m, out

    def ramp(self, fields, t_stage, which=TERMS_RHS | TERMS_PC):
        """LLG_RunRamp: one quasi-static stage per row of `fields` (n x 3),
        t_stage each; returns the per-stage mean magnetisation (n x 3) and
        the summed statistics"""
        f = np.ascontiguousarray(fields, dtype=np.float64).reshape(-1, 3)
        m = np.zeros_like(f)
        res = llg_run_result()
        rc = lib().LLG_RunRamp(self.h, which, f.shape[0], _dp(f), t_stage, _dp(m), C.byref(res))
        out = {"retval": res.retval, "wall_s": res.wall_s, "t": res.t_reached,
               "sim_per_wall": res.sim_per_wall, **res.st.as_dict()}
        if rc != 0:
            raise RuntimeError(f"ramp failed (rc {rc}) at t={res.t_reached}")
        return m, out

    def close(self):
        if self.h:
            lib().LLG_RunDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------- ensembles
ENSEMBLE_GRIDS = ((256, 256), (512, 256), (768, 512), (1024, 512))


def ensemble_members(n=64, seed=4, demag=2.0, grids=ENSEMBLE_GRIDS, frac=(0.2, 0.8),
                     hmax=0.1):
    """The batched-ensemble workload: n instances, grid drawn from `grids`,
    an ellipse or a rectangular ring of seeded active fraction in `frac`,
    in-plane h_ext of magnitude U[0, hmax] along a random direction, FFT
    demag `demag` (periodic), i4 physics, two-domain initial state.
    Returns a list of (llg_config, Geometry, y0)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        ng, ny = grids[int(rng.integers(len(grids)))]
        f = float(rng.uniform(*frac))
        if rng.random() < 0.5:
            aspect = float(rng.uniform(0.7, 1.3))
            rx = min(0.5, math.sqrt(f * aspect / math.pi))
            ry = min(0.5, f / (math.pi * rx))
            geom = Geometry.ellipse(ng, ny, rx, ry)
        else:
            wf, hf = float(rng.uniform(0.9, 1.0)), float(rng.uniform(0.9, 1.0))
            inner = math.sqrt(max(0.0, 1.0 - min(f, 0.99 * wf * hf) / (wf * hf)))
            geom = Geometry.ring(ng, ny, wf, hf, inner, inner)
        h, phi = float(rng.uniform(0.0, hmax)), float(rng.uniform(0.0, 2.0 * math.pi))
        cfg = preset("i4", ng=ng, ny=ny, demag=demag)
        cfg.rhs.happ[0], cfg.rhs.happ[1], cfg.rhs.happ[2] = h * math.cos(phi), h * math.sin(phi), 0.0
        y0 = geom.ic_split_x(max_angle=0.05, seed=1000 + i)
        out.append((cfg, geom, y0))
    return out


class Ensemble:
    """Independent instances (own grid, geometry, fields, step control)
    integrated concurrently on one GPU: one context, stream, integrator and
    host thread per instance (LLG_EnsembleCreate).  members: list of
    (llg_config, Geometry, y0)."""

    def __init__(self, members, nthreads=0, **run_kw):
        L = lib()
        self.n = len(members)
        self._keep = []
        arr = (llg_instance * self.n)()
        for i, (cfg, geom, y0) in enumerate(members):
            y0 = np.ascontiguousarray(y0, dtype=np.float64)
            self._keep += [geom, y0]
            arr[i].cfg = cfg
            arr[i].active = _ucp(geom.active)
            arr[i].w = _dp(geom.w)
            arr[i].cell_msk = _dp(geom.msk)
            arr[i].cell_nsk = _dp(geom.nsk)
            arr[i].y0 = _dp(y0)
        self.sizes = [3 * cfg.ng * cfg.ny for cfg, _, _ in members]
        opts = Simulation.run_opts(**run_kw)
        self.h = L.LLG_EnsembleCreate(self.n, arr, C.byref(opts), nthreads)
        if not self.h:
            raise RuntimeError("LLG_EnsembleCreate failed: " + L.LLG_LastError().decode())

    def advance(self, t_end):
        res = llg_ensemble_result()
        rc = lib().LLG_EnsembleAdvance(self.h, t_end, C.byref(res))
        out = {k: getattr(res, k) for k, _ in res._fields_ if k != "pad"}
        if rc != 0:
            raise RuntimeError(f"ensemble integration failed (rc {rc}, {res.failed} instances)")
        return out

    def state(self, i):
        y = np.empty(self.sizes[i])
        if lib().LLG_EnsembleGetState(self.h, i, _dp(y)) != 0:
            raise RuntimeError("LLG_EnsembleGetState failed")
        return y

    def set_batched(self, mode=1):
        """mode 1: same-shape instances evaluate f together (one
        LLG_RhsBatch per meeting); returns the number of shape groups."""
        g = lib().LLG_EnsembleSetBatched(self.h, mode)
        if g < 0:
            raise RuntimeError("LLG_EnsembleSetBatched failed: " + lib().LLG_LastError().decode())
        return g

    def batch_stats(self):
        launches, calls = C.c_long(0), C.c_long(0)
        lib().LLG_EnsembleBatchStats(self.h, C.byref(launches), C.byref(calls))
        return {"launches": launches.value, "calls": calls.value}

    def result(self, i):
        r = llg_run_result()
        lib().LLG_EnsembleGetResult(self.h, i, C.byref(r))
        return {"t": r.t_reached, "wall_s": r.wall_s, **r.st.as_dict()}

    def close(self):
        if getattr(self, "h", None):
            lib().LLG_EnsembleDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- batches
class Batch:
    """One launch per pass over several contexts (include/llg/llg_batch.h).
    sims: Simulation objects (dense-mask execution)."""

    def __init__(self, sims):
        self.sims = list(sims)
        arr = (C.c_void_p * len(self.sims))(*[s.ctx for s in self.sims])
        self.h = lib().LLG_BatchCreate(len(self.sims), arr)
        if not self.h:
            raise RuntimeError("LLG_BatchCreate failed: " + lib().LLG_LastError().decode())

    @staticmethod
    def _dev(v):
        return C.cast(lib().N_VGetDeviceArrayPointer_Hip(v.v), C.c_void_p).value

    @classmethod
    def _ptrs(cls, vs):
        return (C.c_void_p * len(vs))(*[cls._dev(v) if v is not None else None for v in vs])

    def rhs(self, ys, outs, res=None, stream=None):
        """ys, outs: HipVector per instance (None skips it); res: per
        instance None or (zn1, ycor, c0, c1, c2) with HipVector zn1, ycor.
        Launches on `stream` (default: the first instance's) and waits for
        the device, so the results are readable from any stream."""
        keep, rp = [], None
        if res is not None:
            rp = (C.c_void_p * len(res))()
            for i, r in enumerate(res):
                if r is None:
                    continue
                zn1, ycor, c0, c1, c2 = r
                st = llg_res_in(self._dev(zn1), self._dev(ycor), c0, c1, c2)
                keep.append(st)
                rp[i] = C.addressof(st)
        rc = lib().LLG_RhsBatch(self.h, self._ptrs(ys), self._ptrs(outs), rp,
                                stream or self.sims[0].stream)
        if rc != 0 or lib().LLG_DeviceSynchronize() != 0:
            raise RuntimeError("LLG_RhsBatch failed: " + lib().LLG_LastError().decode())

    def jtv(self, vs, jvs, ys, stream=None):
        rc = lib().LLG_JtvBatch(self.h, self._ptrs(vs), self._ptrs(jvs), self._ptrs(ys),
                                stream or self.sims[0].stream)
        if rc != 0 or lib().LLG_DeviceSynchronize() != 0:
            raise RuntimeError("LLG_JtvBatch failed: " + lib().LLG_LastError().decode())

    def close(self):
        if getattr(self, "h", None):
            lib().LLG_BatchDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------- learned preconditioner
class LearnedPrecond:
    """The learned correction to block-Jacobi (include/llg/learned_pc.h) on a
    Simulation's grid.  Training pairs come from the solver's own FP64
    operator: x is a sampled smooth field on the active cells, b = (I -
    gamma J(y)) x by LLG_Atimes, u = P_bj^-1 b by PrecondSolve, so x solves
    the Newton system exactly."""

    def __init__(self, sim, periodic=None):
        L = lib()
        self.sim = sim
        if periodic is None:
            periodic = sim.cfg.boundary == BC_PERIODIC and sim.cfg.demag != DEMAG_PADDED
        self.h = L.LLG_LpcCreate(sim.ctx, 1 if periodic else 0)
        if not self.h:
            raise RuntimeError("LLG_LpcCreate failed: " + L.LLG_LpcLastError().decode())
        self.nw = L.LLG_LpcNumWeights(self.h)
        n = sim.neq
        self._y, self._x, self._b, self._u = sim.vec(), sim.vec(), sim.vec(), sim.vec()

    # weights
    def weights(self):
        w = np.zeros(self.nw, np.float32)
        lib().LLG_LpcGetWeights(self.h, w.ctypes.data_as(C.POINTER(C.c_float)))
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == self.nw
        if lib().LLG_LpcSetWeights(self.h, w.ctypes.data_as(C.POINTER(C.c_float))) != 0:
            raise RuntimeError("LLG_LpcSetWeights failed")

    def grad(self):
        g = np.zeros(self.nw, np.float32)
        lib().LLG_LpcGetGrad(self.h, g.ctypes.data_as(C.POINTER(C.c_float)))
        return g

    def init_random(self, seed=0, std=None):
        rng = np.random.default_rng(seed)
        sizes = [9 * 16 * 32, 9 * 32 * 32, 9 * 32 * 16]
        parts = []
        for k, nk in enumerate(sizes):
            fan_in = 9 * (16 if k == 0 else 32)
            sd = std if std is not None else 1.0 / math.sqrt(fan_in)
            parts.append(rng.normal(0.0, sd, nk))
        self.set_weights(np.concatenate(parts).astype(np.float32))

    # pairs
    def smooth_field(self, rng, nmodes=6, kmax=6, noise=None):
        """a few random Fourier modes plus cell-wise noise of random weight
        (Newton corrections carry both smooth and grid-scale content),
        zero off the active cells"""
        g = self.sim.geom
        ng, ny = g.ng, g.ny
        x, yy = np.arange(ng) / ng, np.arange(ny) / ny
        out = np.zeros((3, ny, ng))
        for c in range(3):
            for _ in range(nmodes):
                kx, ky = rng.integers(0, kmax + 1, 2)
                ph = rng.uniform(0, 2 * math.pi)
                amp = rng.normal()
                # cos(a + b) with a along x (phase included) and b along y
                a, b = 2 * math.pi * kx * x + ph, 2 * math.pi * ky * yy
                out[c] += np.outer(amp * np.cos(b), np.cos(a))
                out[c] -= np.outer(amp * np.sin(b), np.sin(a))
        if noise is None:
            noise = float(rng.uniform(0.0, 1.0))
        out += noise * math.sqrt(nmodes) * rng.standard_normal(out.shape)
        out *= np.tile(g.active.reshape(ny, ng).astype(float), (3, 1, 1))
        return out.ravel()

    def make_pair(self, y, gamma, rng):
        """device buffers (u, y, x) for one exact pair at state y, gamma"""
        L, sim = lib(), self.sim
        L.LLG_SetLearnedPrecond(sim.ctx, None)
        self._y.set(y)
        vf = sim.vec()
        L.LLG_Rhs(0.0, self._y.v, vf.v, sim.ctx)          # demag field for PrecondSetup
        jc = C.c_int()
        L.PrecondSetup(0.0, self._y.v, vf.v, 0, C.byref(jc), gamma, sim.ctx)
        self._x.set(self.smooth_field(rng))
        L.LLG_Atimes(self._x.v, self._b.v, gamma, 0.0, self._y.v, None, sim.ctx)
        L.PrecondSolve(0.0, self._y.v, None, self._b.v, self._u.v, gamma, 0.0, 1, sim.ctx)
        xs = float(np.sqrt(np.mean(self._x.get() ** 2)) + 1e-30)
        return 1.0 / xs

    def dptr(self, v):
        return lib().N_VGetDeviceArrayPointer_Hip(v.v)

    def loss_grad(self, gamma, scale):
        loss = C.c_double()
        if lib().LLG_LpcLossGrad(self.h, self.dptr(self._u), self.dptr(self._y),
                                 self.dptr(self._x), gamma, scale, C.byref(loss)) != 0:
            raise RuntimeError("LLG_LpcLossGrad failed: " + lib().LLG_LpcLastError().decode())
        return loss.value

    def forward(self, gamma, scale):
        o = np.zeros(4 * self.sim.ncell, np.float32)
        lib().LLG_LpcForward(self.h, self.dptr(self._u), self.dptr(self._y), gamma, scale,
                             o.ctypes.data_as(C.POINTER(C.c_float)))
        return o.reshape(-1, 4)

    def train(self, states, steps=200, lr=1e-3, gamma_range=(1e-3, 1e-1), seed=0,
              log_every=0):
        rng = np.random.default_rng(seed)
        hist = []
        for it in range(steps):
            y = states[int(rng.integers(len(states)))]
            gamma = float(math.exp(rng.uniform(math.log(gamma_range[0]),
                                               math.log(gamma_range[1]))))
            scale = self.make_pair(y, gamma, rng)
            loss = C.c_double()
            if lib().LLG_LpcTrainStep(self.h, self.dptr(self._u), self.dptr(self._y),
                                      self.dptr(self._x), gamma, scale, lr,
                                      C.byref(loss)) != 0:
                raise RuntimeError("LLG_LpcTrainStep failed")
            hist.append(loss.value)
            if log_every and it % log_every == 0:
                print(f"[lpc] step {it} loss {loss.value:.4e}", flush=True)
        return hist

    def set_fused(self, on=True):
        """inference form: the one-kernel MFMA pass (default) or the
        layer-by-layer kernels"""
        lib().LLG_LpcSetFused(self.h, 1 if on else 0)

    def apply(self, y, gamma, u):
        """z = u + active * N(u) at state y (LLG_LpcApply), host arrays"""
        vy, vu, vz = self.sim.vec(y), self.sim.vec(u), self.sim.vec()
        if lib().LLG_LpcApply(self.h, self.dptr(vy), gamma, self.dptr(vu), self.dptr(vz)) != 0:
            raise RuntimeError("LLG_LpcApply failed: " + lib().LLG_LpcLastError().decode())
        return vz.get()

    def enable(self, on=True):
        lib().LLG_SetLearnedPrecond(self.sim.ctx, self.h if on else None)

    def close(self):
        if getattr(self, "h", None):
            lib().LLG_SetLearnedPrecond(self.sim.ctx, None)
            lib().LLG_LpcDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
