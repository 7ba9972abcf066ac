"""The BDF/Newton/SPGMR integrator (product, libllg_amd.so) driven through
the oracle's serial N_Vector on the CPU.

Pins (the integrator restates SUNDIALS CVODE, which is absent here, so its
step/iteration sequence is "parity unpinned"; these fix its numerics):
* macrospin closed form: tan(theta/2) = tan(theta0/2) exp(-alpha H t),
  phi = phi0 + chg H t
* linear ODE y' = A y against expm(A t) y0 (difference-quotient Jv, no
  preconditioner) -- exercises DQ Jv and order changes
* the plain Neumann/LL form on 6x5 cells against a scipy Radau trajectory
  (rtol 1e-11) committed as a fixture
* failure handling: unrecoverable / recoverable RHS errors, mxstep exhaustion
"""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest
from scipy.linalg import expm

import llg_amd as la
from oracle import oracle as orc

GOLD = Path(__file__).resolve().parent / "golden"
RHSFN = C.CFUNCTYPE(C.c_int, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p)


def _integrator(f_addr, ud, y, rtol, atol, maxl=5, prec=True, jtv=None, pset=None, psol=None,
                mxstep=5000):
    P = la.lib()
    mem = P.LLGCVodeCreate(2, None)
    assert P.LLGCVodeInit(mem, f_addr, 0.0, y.v) == 0
    assert P.LLGCVodeSetUserData(mem, ud) == 0
    assert P.LLGCVodeSStolerances(mem, rtol, atol) == 0
    assert P.LLGCVodeSetLinearSolverSPGMR(mem, 1 if prec else 0, maxl, 1) == 0
    assert P.LLGCVodeSetJacTimes(mem, None, jtv) == 0
    if prec:
        assert P.LLGCVodeSetPreconditioner(mem, pset, psol) == 0
    assert P.LLGCVodeSetMaxNumSteps(mem, mxstep) == 0
    return mem


def _free(mem):
    m = C.c_void_p(mem)
    la.lib().LLGCVodeFree(C.byref(m))


def test_macrospin_closed_form():
    a, H = 0.1, 1.0
    cfg = la.preset("plain128", ng=1, ny=1, a=a, k_u=0.0, h_ext=(0.0, 0.0, H))
    sim = orc.OracleSim(cfg, la.Geometry.full(1, 1))
    th0, ph0 = 2.5, 0.3
    y = orc.SerialVector(3, [np.sin(th0) * np.cos(ph0), np.sin(th0) * np.sin(ph0), np.cos(th0)])
    mem = _integrator(orc.addr("orc_cv_rhs"), sim.ctx, y, 1e-10, 1e-12,
                      jtv=orc.addr("orc_cv_jtv"), pset=orc.addr("orc_cv_psetup"),
                      psol=orc.addr("orc_cv_psolve"))
    alpha, chg = a / (1 + a * a), 1 / (1 + a * a)
    tret = C.c_double()
    for t in (0.5, 2.0, 5.0, 10.0):
        assert la.lib().LLGCVode(mem, t, y.v, C.byref(tret), 1) == 0
        m = y.data()
        th = 2 * np.arctan(np.tan(th0 / 2) * np.exp(-alpha * H * t))
        ph = ph0 + chg * H * t
        expect = [np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)]
        assert np.max(np.abs(m - expect)) < 2e-7, (t, m, expect)
    st = la.native.llgcv_stats()
    la.lib().LLGCVodeGetStats(mem, C.byref(st))
    assert st.qlast >= 3          # variable order climbs past BDF2
    _free(mem)


def test_linear_ode_against_expm():
    rng = np.random.default_rng(0)
    n = 6
    Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
    A = Q @ np.diag([-0.5, -2.0, -10.0, -50.0, -1.0, -0.1]) @ Q.T
    A[0, 1] += 1.5
    A[1, 0] -= 1.5                 # rotation component
    y0 = rng.standard_normal(n)
    L = orc.load()

    @RHSFN
    def f(t, yv, ydv, ud):
        yy = np.ctypeslib.as_array(L.orc_nv_data(yv), shape=(n,))
        yd = np.ctypeslib.as_array(L.orc_nv_data(ydv), shape=(n,))
        yd[:] = A @ yy
        return 0

    y = orc.SerialVector(n, y0)
    addr = C.cast(f, C.c_void_p).value
    mem = _integrator(addr, None, y, 1e-9, 1e-11, prec=False)
    tret = C.c_double()
    for t in (0.1, 1.0, 3.0):
        assert la.lib().LLGCVode(mem, t, y.v, C.byref(tret), 1) == 0
        assert np.max(np.abs(y.data() - expm(A * t) @ y0)) < 1e-6
    _free(mem)


def test_plain_form_against_radau_fixture():
    d = np.load(GOLD / "plain_traj_6x5.npz")
    ng, ny = int(d["ng"]), int(d["ny"])
    cfg = la.preset("plain128", ng=ng, ny=ny)
    sim = orc.OracleSim(cfg, la.Geometry.full(ng, ny))
    y = orc.SerialVector(3 * ng * ny, d["y0"])
    mem = _integrator(orc.addr("orc_cv_rhs"), sim.ctx, y, 1e-8, 1e-10,
                      jtv=orc.addr("orc_cv_jtv"), pset=orc.addr("orc_cv_psetup"),
                      psol=orc.addr("orc_cv_psolve"))
    tret = C.c_double()
    for t, yref in zip(d["t"], d["y"]):
        assert la.lib().LLGCVode(mem, float(t), y.v, C.byref(tret), 1) == 0
        assert np.max(np.abs(y.data() - yref)) < 2e-6
    _free(mem)


def test_rhs_failures_and_mxstep():
    L = orc.load()
    calls = {"n": 0}

    def make(code_after):
        @RHSFN
        def f(t, yv, ydv, ud):
            calls["n"] += 1
            yd = np.ctypeslib.as_array(L.orc_nv_data(ydv), shape=(2,))
            yd[:] = [-1.0, 1.0]
            return code_after if calls["n"] > 1 else 0
        return f

    # unrecoverable failure after the first call -> LLGCV_RHSFUNC_FAIL or a conv failure
    f = make(-1)
    y = orc.SerialVector(2, [1.0, 0.0])
    mem = _integrator(C.cast(f, C.c_void_p).value, None, y, 1e-6, 1e-8, prec=False)
    rc = la.lib().LLGCVode(mem, 1.0, y.v, C.byref(C.c_double()), 1)
    assert rc < 0
    _free(mem)
    # recoverable failure on the very first call
    calls["n"] = 0

    @RHSFN
    def g(t, yv, ydv, ud):
        return 1

    y = orc.SerialVector(2, [1.0, 0.0])
    mem = _integrator(C.cast(g, C.c_void_p).value, None, y, 1e-6, 1e-8, prec=False)
    assert la.lib().LLGCVode(mem, 1.0, y.v, C.byref(C.c_double()), 1) == -9   # FIRST_RHSFUNC_ERR
    _free(mem)
    # too much work
    A = np.array([[-1.0, 30.0], [-30.0, -1.0]])

    @RHSFN
    def h(t, yv, ydv, ud):
        yy = np.ctypeslib.as_array(L.orc_nv_data(yv), shape=(2,))
        yd = np.ctypeslib.as_array(L.orc_nv_data(ydv), shape=(2,))
        yd[:] = A @ yy
        return 0

    y = orc.SerialVector(2, [1.0, 0.0])
    mem = _integrator(C.cast(h, C.c_void_p).value, None, y, 1e-8, 1e-10, prec=False, mxstep=20)
    tret = C.c_double()
    assert la.lib().LLGCVode(mem, 10.0, y.v, C.byref(tret), 1) == -1     # TOO_MUCH_WORK
    assert 0.0 < tret.value < 10.0
    _free(mem)


def _fused_ops_type():
    """ctypes mirror of llgcv_fused_ops: every member is a function pointer,
    in the header's order."""
    import re
    text = (Path(__file__).resolve().parent.parent / "include" / "llg" / "llg_cvode.h").read_text()
    body = text[text.index("typedef struct llgcv_fused_ops {"):text.index("} llgcv_fused_ops;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = re.findall(r"\(\*(\w+)\)\s*\(|^\s*CVLs\w+\s+(\w+);", body, flags=re.M)
    names = [a or b for a, b in names]
    return type("llgcv_fused_ops", (C.Structure,), {"_fields_": [(n, C.c_void_p) for n in names]})


RESFN = C.CFUNCTYPE(C.c_int, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                    C.c_void_p, C.c_void_p)


@pytest.mark.parametrize("declared", [True, False])
def test_rhs_residual_only_with_fy_free_callbacks(declared):
    """The fused Newton residual never stores f, so the fy a preconditioner
    setup receives would be stale: the integrator must take that path only
    while the registered jtimes / psetup are the ones declared fy-free.  The
    hook here declines every call (returns 2), so the trajectory is the
    plain one either way; only whether it is consulted differs."""
    cfg = la.preset("plain128", ng=6, ny=5)
    sim = orc.OracleSim(cfg, la.Geometry.full(6, 5))
    rng = np.random.default_rng(5)
    th, ph = 0.1 * rng.random(30), 2 * np.pi * rng.random(30)
    y = orc.SerialVector(90, np.concatenate([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph),
                                             np.cos(th)]))
    jtv, pset = orc.addr("orc_cv_jtv"), orc.addr("orc_cv_psetup")
    mem = _integrator(orc.addr("orc_cv_rhs"), sim.ctx, y, 1e-6, 1e-8, jtv=jtv, pset=pset,
                      psol=orc.addr("orc_cv_psolve"))
    calls = [0]

    def hook(t, yv, c, zn1, ycor, res, ud):
        calls[0] += 1
        return 2

    cb = RESFN(hook)
    fo = _fused_ops_type()()
    fo.rhs_residual = C.cast(cb, C.c_void_p).value
    fo.fy_free_jtimes = jtv
    fo.fy_free_psetup = pset if declared else None
    try:
        assert la.lib().LLGCVodeSetFusedOps(mem, C.byref(fo)) == 0
        tret = C.c_double()
        assert la.lib().LLGCVode(mem, 1.0, y.v, C.byref(tret), 1) == 0
    finally:
        _free(mem)
        sim.close()
    assert (calls[0] > 0) == declared
