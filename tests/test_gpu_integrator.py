"""The integration loop on the GPU (HIP plug-ins + HIP N_Vector + the BDF
integrator) against the same integrator driving the CPU oracle's
callbacks and serial vectors.

Trajectory bar: at rtol = atol = 1e-6 the GPU state stays within 1e-5
max |dm| of the oracle's; |m| drift is printed.  The integrator's one-pass
kernels (Nordsieck predict/restore, error weights, SPGMR scale/divide,
product + Gram-Schmidt multi-dot, combination + norm) must reproduce the
standard N_Vector op sequence bit for bit, as must the one-launch Krylov
operator chain (scale/divide, J v, block-Jacobi solve and product) and the
device-resident SPGMR solve (scalars, Givens rotations and the convergence
test on the device, one host wait per solve), so the fused and unfused runs
are compared for equality.
"""
import numpy as np
import pytest

import llg_amd as la
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _case(name="square_hole", ng=48, ny=32, **kw):
    cfg = la.preset(name, ng=ng, ny=ny, **kw)
    geom = la.Geometry.square_hole(ng, ny, halfside_frac=0.2)
    return cfg, geom, geom.ic_split_x()


def _compact_case():
    cfg = la.preset("i6", ng=64, ny=64, demag=4.0)
    geom = la.Geometry.ellipse(64, 64, 0.3, 0.3)
    return cfg, geom, geom.ic_two_domain()


@pytest.mark.parametrize("gstype,case", [(1, "square_hole"), (2, "square_hole"),
                                         (2, "i6_compact")])
def test_fused_integrator_kernels_are_bitwise_neutral(gstype, case):
    """fused = 1: standard op sequences; 3: the one-pass integrator kernels;
    7: also the one-launch Krylov operator chain; 15: also the whole SPGMR
    solve queued on the device with one host wait (classical GS only);
    263: 7 with the Newton residual formed in the RHS pass; 775: 263 with
    each step's completion left to the next predictor's pass; 8967: 775
    with the completed step's copy of acor into zn[qmax] left to that pass
    too.  All must agree bit for bit, counters included."""
    cfg, geom, y0 = _case() if case == "square_hole" else _compact_case()
    sim = la.Simulation(cfg, geom)
    res = {}
    for fused in (1, 3, 7, 15, 263, 775, 8967):
        run = sim.integrator(y0, rtol=1e-5, atol=1e-5, gstype=gstype, fused=fused)
        out = run.advance(1.0)
        res[fused] = (run.state(), out)
        run.close()
    ya, sa = res[1]
    for fused in (3, 7, 15, 263, 775, 8967):
        yb, sb = res[fused]
        print(f"[gpu] {case} gs={gstype} fused 1: nst={sa['nst']} nli={sa['nli']} "
              f"syncs={sa['nsyncs']}; fused {fused}: nst={sb['nst']} nli={sb['nli']} "
              f"syncs={sb['nsyncs']}")
        for k in ("nst", "nfe", "nni", "nli", "netf", "ncfn", "nsetups"):
            assert sa[k] == sb[k], (fused, k)
        assert np.array_equal(ya, yb), fused


LAG_CASES = [("square_hole_48x32", lambda: _case()),
             ("square_hole_padded_128x64", lambda: _case(ng=128, ny=64)),
             ("i1_periodic_64x64", lambda: (la.preset("i1", ng=64, ny=64, demag=2.0),
                                            la.Geometry.full(64, 64),
                                            la.Geometry.full(64, 64).ic_stripes()))]


@pytest.mark.parametrize("maxl", [8, 5, 1])
@pytest.mark.parametrize("name,make", LAG_CASES, ids=[c[0] for c in LAG_CASES])
def test_device_driven_lagged_spgmr_is_bitwise_neutral(name, make, maxl):
    """fused 1815 (791 plus bit 10): every lagged SPGMR solve queued whole,
    the Hessenberg column, Pythagorean estimate, Givens rotation,
    convergence test and next coefficients computed on the device, one
    host wait per solve.  The device repeats the host loop's operations,
    so trajectory and counters equal the host-stepped lagged solve's (791)
    bit for bit; the host waits drop.  fused 3863 (1815 plus bit 11): the
    solution and the first Newton update of each iteration queued behind
    the solve (coefficients from the device's qr_sol), so no host round
    trip sits between the solve and the update; the same bits.  fused 7959
    (3863 plus bit 12): on a step that will test an order change, the
    test's two norms queued behind the first Newton update and used when
    that iteration converged; the same bits and waits.  fused 16151 (7959
    plus bit 13): a completed step's copy of acor into zn[qmax] made by the
    next predictor's pass; the same bits and waits.  maxl 8, the widest
    lagged solve, runs every pass and step kernel instantiation (and the
    widest solution + update); maxl 1 makes every column the last: its
    Pythagorean estimate or projection norm, and the Krylov cap."""
    cfg, geom, y0 = make()
    sim = la.Simulation(cfg, geom)
    res = {}
    for fused in (791, 1815, 3863, 7959, 16151):
        run = sim.integrator(y0, rtol=1e-5, atol=1e-5, gstype=2, fused=fused, maxl=maxl)
        s0 = la.lib().NVHip_NumSyncs()      # a process-wide count
        out = run.advance(1.0)
        res[fused] = (run.state(), out, la.lib().NVHip_NumSyncs() - s0)
        run.close()
    (ya, sa, wa), (yb, sb, wb), (yc, sc, wc) = res[791], res[1815], res[3863]
    yd, sd, wd = res[7959]
    ye, se, we = res[16151]
    print(f"[gpu] {name} maxl={maxl}: host-stepped nst={sa['nst']} nli={sa['nli']} "
          f"waits={wa}; device-stepped nst={sb['nst']} nli={sb['nli']} waits={wb}; "
          f"with the update queued behind the solve waits={wc}", flush=True)
    for k in ("nst", "nfe", "nni", "nli", "netf", "ncfn", "nsetups"):
        assert sa[k] == sb[k] == sc[k] == sd[k] == se[k], k
    assert np.array_equal(ya, yb)
    assert np.array_equal(ya, yc)
    assert np.array_equal(ya, yd)
    assert np.array_equal(ya, ye)
    # maxl 1: both forms wait for pass 0 and, when needed, the projection norm
    assert wb < wa if maxl > 1 else wb <= wa
    # bit 11 keeps the waits (the solve's state, then the update's norm) but
    # the update is already queued when the solve's state arrives
    assert wc == wb
    # bit 12 waits for the queued norms instead of the norms it would queue
    assert wd == wc
    assert we == wd
    sim.close()


def test_widest_lagged_solve_is_bitwise_neutral():
    """Krylov dimension 8 on a workload whose solves go deeper than 5
    iterations (the 128x64 square hole at rtol = atol = 1e-3 over three
    units: the run differs from the same run at dimension 5, so some solve
    used more than 5 columns): the pass and step kernels of columns 5..7,
    the solve's second part and the widest solution + Newton update run,
    and the device-stepped forms (1815, 16151) equal the host-stepped
    lagged solve (791) bit for bit."""
    cfg, geom, y0 = _case(ng=128, ny=64)
    sim = la.Simulation(cfg, geom)
    res = {}
    for fused, maxl in ((791, 8), (1815, 8), (16151, 8), (16151, 5)):
        run = sim.integrator(y0, rtol=1e-3, atol=1e-3, gstype=2, fused=fused, maxl=maxl,
                             mxstep=100000)
        out = run.advance(3.0)
        res[(fused, maxl)] = (run.state(), out)
        run.close()
    ya, sa = res[(791, 8)]
    print(f"[gpu] widest lagged solve: nst={sa['nst']} nni={sa['nni']} nli={sa['nli']} "
          f"ncfn={sa['ncfn']}; at dimension 5 nli={res[(16151, 5)][1]['nli']}", flush=True)
    for key in ((1815, 8), (16151, 8)):
        yb, sb = res[key]
        for k in ("nst", "nfe", "nni", "nli", "netf", "ncfn", "nsetups"):
            assert sa[k] == sb[k], (key, k)
        assert np.array_equal(ya, yb), key
    y5, s5 = res[(16151, 5)]
    assert s5["nli"] != sa["nli"] or not np.array_equal(y5, ya)
    assert sa["nli"] > 4 * sa["nni"]
    sim.close()


@pytest.mark.parametrize("fused", [7, 23, 55, 215, 279, 791, 1815, 16151])
@pytest.mark.parametrize("name,kw", [("square_hole", {}), ("i6", {"demag": 4.0}),
                                     ("plain128", {})])
def test_trajectory_matches_oracle(name, kw, fused):
    """fused 7: the standard SPGMR loop with the one-pass kernels (bit-equal
    to the plain op sequence); 23: SPGMR with lagged normalisation
    (LLG_Arnoldi; the compact i6 grid falls back to the standard loop);
    55: the same pipelined, each pass queued behind a gate (polling waits);
    215: the lagged solve with the Newton residual and the SPGMR solution
    formed where they are read (never written); 279: 23 with the Newton
    residual formed in the RHS pass; 791: 279 with each step's completion
    left to the next predictor's pass; 1815: 791 with each SPGMR solve
    queued whole and stepped on the device; 16151: the default (1815 with
    the solution and Newton update, the order-change norms and the acor
    copy queued behind the solve or left to the next predictor)"""
    ng, ny = (48, 32) if name != "plain128" else (32, 32)
    cfg = la.preset(name, ng=ng, ny=ny, **kw)
    if name == "plain128":
        geom = la.Geometry.full(ng, ny)
        y0 = geom.ic_tilted_z()
        t_end = 2.0
    else:
        geom = la.Geometry.square_hole(ng, ny, halfside_frac=0.2)
        y0 = geom.ic_split_x() if name == "square_hole" else geom.ic_two_domain()
        t_end = 1.0
    rtol = atol = 1e-6
    sim = la.Simulation(cfg, geom)
    L = la.lib()
    L.NVHip_SetSyncMode(1 if fused == 55 else 0)
    try:
        run = sim.integrator(y0, rtol=rtol, atol=atol, t_out=0.1, fused=fused)
        out = run.advance(t_end)
    finally:
        L.NVHip_SetSyncMode(0)
    y_gpu = run.state()
    y_cpu, st_cpu, _ = orc.cpu_integrate(la.lib(), cfg, geom, y0, t_end, t_out=0.1,
                                         rtol=rtol, atol=atol, mxstep=5000)
    dm = np.max(np.abs(y_gpu - y_cpu))
    act = geom.active.astype(bool)
    m = y_gpu.reshape(3, -1)[:, act]
    drift = np.max(np.abs(np.sqrt((m ** 2).sum(0)) - 1.0))
    print(f"[gpu] {name} fused {fused}: max|dm| {dm:.2e} (nst gpu {out['nst']} "
          f"cpu {st_cpu['nst']}, nli gpu {out['nli']} cpu {st_cpu['nli']}), "
          f"| |m|-1 | max {drift:.2e}")
    assert dm < 1e-5
