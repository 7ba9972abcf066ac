"""Python access to the CPU oracle (oracle/llg_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py.  The product package never imports it.
`build_oracle()` compiles liborc.so with gcc (OpenMP) into oracle/_build/.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liborc.so"

DP = C.POINTER(C.c_double)
UCP = C.POINTER(C.c_ubyte)
VP = C.c_void_p

_lib = None


def build_oracle() -> Path:
    r = subprocess.run(["make", "-C", str(HERE), "-B"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stderr[-4000:])
    return LIB


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        build_oracle()
    L = C.CDLL(str(LIB))
    sig = {
        "orc_ctt": (None, [C.c_double, C.c_double, C.c_double, C.c_double, DP]),
        "orc_tensor_at": (None, [C.c_double, C.c_int, C.c_int, DP]),
        "orc_calt": (None, [C.c_double, C.c_int, C.c_int, DP]),
        "orc_demag_direct": (None, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, DP, DP, DP]),
        "orc_demag_create": (VP, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int]),
        "orc_demag_destroy": (None, [VP]),
        "orc_demag_apply": (None, [VP, DP, DP, DP]),
        "orc_demag_self": (None, [VP, DP]),
        "orc_create": (VP, [VP, UCP, DP, DP, DP]),
        "orc_destroy": (None, [VP]),
        "orc_counts": (C.c_int, [VP, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "orc_ymsk": (DP, [VP]),
        "orc_active_ids": (C.POINTER(C.c_int), [VP]),
        "orc_inactive_ids": (C.POINTER(C.c_int), [VP]),
        "orc_hd": (DP, [VP]),
        "orc_pinv": (DP, [VP]),
        "orc_self_coupling": (None, [VP, DP]),
        "orc_set_happ": (None, [VP, C.c_int, DP]),
        "orc_normalize": (None, [VP, DP]),
        "orc_rhs_arrays": (None, [VP, DP, DP]),
        "orc_jtv_arrays": (None, [VP, DP, DP, DP]),
        "orc_psetup_arrays": (C.c_int, [VP, DP, C.c_int, C.POINTER(C.c_int), C.c_double]),
        "orc_psolve_arrays": (None, [VP, DP, DP]),
        "orc_nv_new": (VP, [C.c_int64]),
        "orc_nv_data": (DP, [VP]),
        "orc_nv_destroy": (None, [VP]),
        "orc_nv_num_reductions": (C.c_long, []),
        "orc_num_threads": (C.c_int, []),
        "orc_set_num_threads": (None, [C.c_int]),
        "orc_fft1d": (None, [C.POINTER(C.c_double), C.c_int, C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def addr(name):
    return C.cast(getattr(load(), name), C.c_void_p).value


def _dp(a):
    return a.ctypes.data_as(DP) if a is not None else None


# ---------------------------------------------------------------- tensor
def tensor_at(thick, dx, dy):
    out = np.zeros(9)
    load().orc_tensor_at(thick, dx, dy, _dp(out))
    return out


def calt(thick, mdx, mdy):
    out = np.zeros(9 * mdx * mdy)
    load().orc_calt(thick, mdx, mdy, _dp(out))
    return out.reshape(9, mdy, mdx)


def demag_direct(nx, ny, thick, strength, y, padded=False, w=None):
    h = np.zeros(3 * nx * ny)
    y = np.ascontiguousarray(y, dtype=np.float64)
    load().orc_demag_direct(nx, ny, thick, strength, 1 if padded else 0, _dp(y), _dp(w), _dp(h))
    return h


class Demag:
    def __init__(self, nx, ny, thick, strength, padded=False):
        self.h = load().orc_demag_create(nx, ny, thick, strength, 1 if padded else 0)
        self.n = nx * ny

    def apply(self, y, w=None):
        h = np.zeros(3 * self.n)
        y = np.ascontiguousarray(y, dtype=np.float64)
        load().orc_demag_apply(self.h, _dp(y), _dp(w), _dp(h))
        return h

    def self_coupling(self):
        v = np.zeros(3)
        load().orc_demag_self(self.h, _dp(v))
        return v

    def __del__(self):
        try:
            load().orc_demag_destroy(self.h)
        except Exception:
            pass


# --------------------------------------------------------------- context
class OracleSim:
    """CPU restatement of f / Jv / PrecondSetup / PrecondSolve."""

    def __init__(self, cfg, geom=None):
        L = load()
        self.cfg = cfg
        a = None if geom is None else geom.active
        w = None if geom is None else geom.w
        msk = None if geom is None else geom.msk
        nsk = None if geom is None else geom.nsk
        self._keep = (a, w, msk, nsk)
        self.ctx = L.orc_create(C.cast(C.pointer(cfg), VP),
                                a.ctypes.data_as(UCP) if a is not None else None,
                                _dp(w), _dp(msk), _dp(nsk))
        self.ncell = cfg.ng * cfg.ny
        self.neq = 3 * self.ncell

    def close(self):
        if getattr(self, "ctx", None):
            load().orc_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def rhs(self, y):
        y = np.array(y, dtype=np.float64, copy=True)
        f = np.zeros(self.neq)
        load().orc_rhs_arrays(self.ctx, _dp(y), _dp(f))
        return f, y

    def jtv(self, y, v):
        y = np.ascontiguousarray(y, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.zeros(self.neq)
        load().orc_jtv_arrays(self.ctx, _dp(y), _dp(v), _dp(out))
        return out

    def psetup(self, y, gamma, jok=0):
        y = np.ascontiguousarray(y, dtype=np.float64)
        jc = C.c_int(0)
        load().orc_psetup_arrays(self.ctx, _dp(y), jok, C.byref(jc), gamma)
        return jc.value

    def psolve(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros(self.neq)
        load().orc_psolve_arrays(self.ctx, _dp(r), _dp(z))
        return z

    def hdemag(self):
        return np.ctypeslib.as_array(load().orc_hd(self.ctx), shape=(self.neq,)).copy()

    def pinv(self):
        return np.ctypeslib.as_array(load().orc_pinv(self.ctx), shape=(9 * self.ncell,)).copy()

    def mask(self):
        return np.ctypeslib.as_array(load().orc_ymsk(self.ctx), shape=(self.neq,)).copy()

    def ids(self):
        nc, na, ni = C.c_int(), C.c_int(), C.c_int()
        load().orc_counts(self.ctx, C.byref(nc), C.byref(na), C.byref(ni))
        a = np.ctypeslib.as_array(load().orc_active_ids(self.ctx), shape=(na.value,)).copy() \
            if na.value else np.zeros(0, np.int32)
        h = np.ctypeslib.as_array(load().orc_inactive_ids(self.ctx), shape=(ni.value,)).copy() \
            if ni.value else np.zeros(0, np.int32)
        return a, h

    def self_coupling(self):
        v = np.zeros(3)
        load().orc_self_coupling(self.ctx, _dp(v))
        return v

    def set_applied_field(self, happ, which=5):
        h = np.ascontiguousarray(happ, dtype=np.float64)
        load().orc_set_happ(self.ctx, which, _dp(h))


class SerialVector:
    def __init__(self, n, values=None):
        self.n = n
        self.v = load().orc_nv_new(n)
        if values is not None:
            self.data()[:] = values

    def data(self):
        return np.ctypeslib.as_array(load().orc_nv_data(self.v), shape=(self.n,))

    def __del__(self):
        try:
            load().orc_nv_destroy(self.v)
        except Exception:
            pass


def cpu_integrate(product_lib, cfg, geom, y0, t_end, t_out=0.1, rtol=1e-4, atol=1e-4, maxl=5,
                  gstype=None, max_order=5, mxstep=500, threads=None):
    """Run the product's BDF integrator on the oracle's serial vectors and
    CPU callbacks (the CPU baseline).  Returns (y, stats dict, wall seconds)."""
    import time
    L = load()
    if threads:
        L.orc_set_num_threads(threads)
    sim = OracleSim(cfg, geom)
    n = sim.neq
    y = SerialVector(n, y0)
    P = product_lib
    mem = P.LLGCVodeCreate(2, None)
    gs = gstype if gstype else (2 if n < 500000 else 1)
    assert P.LLGCVodeSetMaxOrd(mem, max_order) == 0
    assert P.LLGCVodeInit(mem, addr("orc_cv_rhs"), 0.0, y.v) == 0
    assert P.LLGCVodeSetUserData(mem, sim.ctx) == 0
    assert P.LLGCVodeSStolerances(mem, rtol, atol) == 0
    assert P.LLGCVodeSetLinearSolverSPGMR(mem, 1, maxl, gs) == 0
    assert P.LLGCVodeSetJacTimes(mem, None, addr("orc_cv_jtv")) == 0
    assert P.LLGCVodeSetPreconditioner(mem, addr("orc_cv_psetup"), addr("orc_cv_psolve")) == 0
    assert P.LLGCVodeSetMaxNumSteps(mem, mxstep) == 0
    tret = C.c_double(0.0)
    t0 = time.perf_counter()
    tout = t_out
    rc = 0
    while tout <= t_end + 1e-9:
        rc = P.LLGCVode(mem, tout, y.v, C.byref(tret), 1)
        if rc < 0:
            break
        tout += t_out
    wall = time.perf_counter() - t0
    from importlib import import_module
    native = import_module("llg_amd.native")
    st = native.llgcv_stats()
    P.LLGCVodeGetStats(mem, C.byref(st))
    m = C.c_void_p(mem)
    P.LLGCVodeFree(C.byref(m))
    out = y.data().copy()
    sim.close()
    return out, dict(st.as_dict(), retval=rc, t=tret.value), wall


def cpu_ramp(product_lib, cfg, geom, y0, fields, t_stage, which=5, t_out=0.1, rtol=1e-4,
             atol=1e-4, maxl=5, gstype=None, max_order=5, mxstep=5000):
    """A field ramp on the oracle's callbacks with the product's integrator:
    per stage set the applied field, restart the integrator at the current
    (t, y) and advance t_stage on the t_out cadence (the CPU counterpart of
    LLG_RunRamp).  Returns (final y, per-stage states)."""
    sim = OracleSim(cfg, geom)
    n = sim.neq
    y = SerialVector(n, y0)
    P = product_lib
    mem = P.LLGCVodeCreate(2, None)
    gs = gstype if gstype else (2 if n < 500000 else 1)
    assert P.LLGCVodeSetMaxOrd(mem, max_order) == 0
    assert P.LLGCVodeInit(mem, addr("orc_cv_rhs"), 0.0, y.v) == 0
    assert P.LLGCVodeSetUserData(mem, sim.ctx) == 0
    assert P.LLGCVodeSStolerances(mem, rtol, atol) == 0
    assert P.LLGCVodeSetLinearSolverSPGMR(mem, 1, maxl, gs) == 0
    assert P.LLGCVodeSetJacTimes(mem, None, addr("orc_cv_jtv")) == 0
    assert P.LLGCVodeSetPreconditioner(mem, addr("orc_cv_psetup"), addr("orc_cv_psolve")) == 0
    assert P.LLGCVodeSetMaxNumSteps(mem, mxstep) == 0
    tret = C.c_double(0.0)
    t = 0.0
    states = []
    for h in np.asarray(fields, dtype=np.float64).reshape(-1, 3):
        sim.set_applied_field(h, which)
        assert P.LLGCVodeReInit(mem, t, y.v) == 0
        tout, t_end = t + t_out, t + t_stage
        while tout <= t_end + 1e-9 * max(1.0, abs(t_end)):
            assert P.LLGCVode(mem, tout, y.v, C.byref(tret), 1) >= 0
            t = tret.value
            tout += t_out
        states.append(y.data().copy())
    m = C.c_void_p(mem)
    P.LLGCVodeFree(C.byref(m))
    sim.close()
    return states[-1], states
