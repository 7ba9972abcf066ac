#!/usr/bin/env python3
"""Throughput of the 2D LLG solver hot path on one MI355X.

Workload: 1536x512-cell square-hole film (centred 444x444 hole, 25.07 %),
two-domain +-x initial state with seeded 0.05 rad noise, FFT demag of
strength 4 and thickness 1 on the zero-padded 3072x1024 grid, FP64,
BDF + Newton + SPGMR(5) + analytic Jv + 3x3 block-Jacobi, rtol = atol = 1e-4.
A step advances the simulation by 1.0 time unit (ten CVode outputs at the
T1 = 0.1 cadence); the default 100 steps cover t in [0, 100].

Prints one JSON line: sim-time/wall-s, SPGMR iterations per step, the
demag column kernel's achieved bandwidth against MI355X HBM peak, and the
CPU oracle timed on a bounded sample of the same workload.

--workload ref512 runs instead the setting the reference published its
"1536x512" FFT timing on (src/irregular/2d_fft/2d_i1/scaling_logs/
fft_1536x512.txt): 512x512 cells, periodic 512x512 demag of strength 2,
the i1 physics (exchange 4, linear anisotropy 4, DMI 0.3 along x, exact
normalisation), head-on stripes, rtol = atol = 1e-4, Krylov 5, t in
[0, 1000]; 29,644 ms there = 33.73 sim-time/wall-s.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for p in (str(ROOT / "gpu-micromagnetics-irregular_amd"), str(ROOT)):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
TRAFFIC_PROFILE = ROOT / "profiles" / "square_hole_1536x512_hbm_traffic.md"


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary of this workload (None when not measured)."""
    if kernel != "demag_col" or workload != "square_hole" or not TRAFFIC_PROFILE.exists():
        return None
    for line in TRAFFIC_PROFILE.read_text().splitlines():
        if line.startswith("demag_col_traffic_bytes:"):
            return int(line.split(":")[1])
    return None
REF512_PUBLISHED = 1000.0 / 29.644070   # sim-time/wall-s of the reference's own log
KERNELS = {"rhs": 1, "jtv": 2, "build_p": 3, "apply_p": 4, "demag_row_fwd": 5,
           "demag_col": 6, "demag_row_inv": 7, "demag": 8}


def algorithmic_bytes(kernel, ng, ny, ncell, padded=True, col_variant=0):
    """Unique bytes read + written by one launch (FP64, byte mask).  The
    blocked and wave-per-transform column kernels (variants 2-6) read the
    zero-padded grid's tensor planes for ky <= Py/2 only (the planes are
    mirror-symmetric in ky)."""
    px, py = (2 * ng, 2 * ny) if padded else (ng, ny)
    hx = px // 2 + 1
    spec = 3 * ny * hx * 16                     # intermediate spectrum (rows stored)
    if padded and col_variant in (2, 3, 4, 5, 6):
        tensor = 4 * hx * (py // 2 + 1) * 8      # xx, yy, zz, xy; half of each column
    else:
        tensor = (4 if padded else 5) * hx * py * 8  # real tensor planes
    state = 3 * ncell * 8
    return {
        "demag_row_fwd": state + spec,
        "demag_col": 2 * spec + tensor,
        "demag_row_inv": spec + state,
        "demag": 2 * state + 4 * spec + tensor,
        "rhs": 3 * state + ncell,                # y, hd, yd + mask byte
        "jtv": 3 * state + ncell,                # y, v, Jv + mask byte
        "apply_p": 3 * ncell * 8 * 3 + 2 * state + ncell,
        "build_p": 2 * state + 9 * ncell * 8,
    }[kernel]


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def m_norm_drift(y, active):
    """max over the active cells of ||m| - 1| (the integration does not
    renormalise m on this workload: the drift is the integration error's)"""
    m = y.reshape(3, -1)[:, active != 0]
    return float(f"{np.max(np.abs(np.sqrt((m * m).sum(0)) - 1.0)):.3e}") if m.size else 0.0


def cpu_baseline(la, cfg, geom, y0, budget_s, threads):
    """The CPU oracle (OpenMP C restatement of the callbacks + serial
    N_Vector + the same BDF integrator) on the same workload from t = 0,
    stopped after ~budget_s seconds of integration on `threads` threads,
    then again on one thread for half that budget."""
    from oracle import oracle as orc
    L = orc.load()
    t_init = time.perf_counter()
    sim = orc.OracleSim(cfg, geom)
    init_s = time.perf_counter() - t_init
    n = sim.neq
    P = la.lib()

    def run(nthr, budget):
        L.orc_set_num_threads(nthr)
        used = L.orc_num_threads()
        y = orc.SerialVector(n, y0)
        mem = P.LLGCVodeCreate(2, None)
        P.LLGCVodeSetMaxOrd(mem, 5)
        P.LLGCVodeInit(mem, orc.addr("orc_cv_rhs"), 0.0, y.v)
        P.LLGCVodeSetUserData(mem, sim.ctx)
        P.LLGCVodeSStolerances(mem, 1e-4, 1e-4)
        P.LLGCVodeSetLinearSolverSPGMR(mem, 1, 5, 2 if n < 500000 else 1)
        P.LLGCVodeSetJacTimes(mem, None, orc.addr("orc_cv_jtv"))
        P.LLGCVodeSetPreconditioner(mem, orc.addr("orc_cv_psetup"), orc.addr("orc_cv_psolve"))
        tret = C.c_double(0.0)
        t0 = time.perf_counter()
        tout, dt = 0.01, 0.01
        while True:
            rc = P.LLGCVode(mem, tout, y.v, C.byref(tret), 1)
            if rc < 0 or time.perf_counter() - t0 > budget:
                break
            tout += dt
        wall = time.perf_counter() - t0
        st = la.native.llgcv_stats()
        P.LLGCVodeGetStats(mem, C.byref(st))
        m = C.c_void_p(mem)
        P.LLGCVodeFree(C.byref(m))
        return used, tret.value, wall, st

    nthreads, t_end, wall, st = run(threads or 0, budget_s)
    one, t1, wall1, st1 = run(1, budget_s / 2)
    sim.close()
    return {"value": t_end / wall, "unit": "sim-time/wall-s", "cores": nthreads,
            "kind": "port", "t_end": t_end,
            "sample": (f"same workload from t=0 to t={t_end:.3f} ({st.nst} BDF steps, "
                       f"{st.nfe} f evals) in {wall:.1f} s on {nthreads} OpenMP threads; "
                       f"oracle setup (tensor + plans) {init_s:.1f} s excluded"),
            "single_thread": {"value": t1 / wall1, "cores": one,
                              "sample": f"t=0 to t={t1:.4f} ({st1.nst} BDF steps) in "
                                        f"{wall1:.1f} s on 1 thread"},
            "host": {"nproc": os.cpu_count(), "model": cpu_model()}}


def ensemble_bench(la, args):
    """64 independent instances (grids 256x256..1024x512, ellipse/ring
    geometries with active fractions 0.2-0.8, in-plane h_ext up to 0.1,
    periodic FFT demag 2.0), each with its own BDF step control, one host
    thread and stream per instance.  value = instance-simulated time per
    wall second, summed over instances; the sequential comparison runs a
    few instances one after another on the same GPU."""
    la.lib().NVHip_SetSyncMode(2)   # many host threads: poll and yield
    members = la.ensemble_members(n=args.instances, seed=4)
    ncell = sum(c.ng * c.ny for c, _, _ in members)
    print(f"[bench] ensemble of {len(members)} instances, {ncell / 1e6:.2f} M cells",
          file=sys.stderr, flush=True)
    kw = dict(rtol=1e-4, atol=1e-4, maxl=5, gstype=args.gstype, max_order=5, t_out=0.1,
              mxstep=100000)
    step = 1.0
    if args.warmup > 0:
        w = la.Ensemble(members, nthreads=0, **kw)
        w.advance(args.warmup * step)
        w.close()
        print(f"[bench] warmup {args.warmup} steps done", file=sys.stderr, flush=True)
    ens = la.Ensemble(members, nthreads=0, **kw)
    out = ens.advance(args.steps * step)
    ens.close()
    value = out["inst_time"] / out["wall_s"]
    print(f"[bench] ensemble advanced {out['inst_time']:.1f} instance-time units in "
          f"{out['wall_s']:.3f} s: nst={out['nst']} nli={out['nli']}", file=sys.stderr, flush=True)
    # the same run with f evaluated per shape group (one launch per pass for
    # all same-shape instances whose f calls meet); same trajectories
    batched = {}
    for mode, name in ((1, "all_shapes"), (2, "small_shapes")):
        bat = la.Ensemble(members, nthreads=0, **kw)
        groups = bat.set_batched(mode)
        outb = bat.advance(args.steps * step)
        stb = bat.batch_stats()
        bat.close()
        batched[name] = {
            "value": round(outb["inst_time"] / outb["wall_s"], 4),
            "ratio_to_independent": round(outb["inst_time"] / outb["wall_s"] / value, 3),
            "shape_groups": groups, "f_calls_batched": stb["calls"],
            "group_evaluations": stb["launches"], "nst": outb["nst"],
            "same_step_counts": outb["nst"] == out["nst"]}
        print(f"[bench] batched f ({name}): {batched[name]}", file=sys.stderr, flush=True)
    # the same instances one at a time (a sample of 4, a short interval each)
    seq_t, seq_wall = 0.0, 0.0
    for cfg, geom, y0 in members[:4]:
        sim = la.Simulation(cfg, geom)
        run = sim.integrator(y0, **kw)
        r = run.advance(0.5 * step)
        seq_t += r["t"]
        seq_wall += r["wall_s"]
        run.close()
        sim.close()
    la.lib().NVHip_SetSyncMode(args.sync_mode)
    seq = seq_t / seq_wall if seq_wall > 0 else None
    rec = {
        "metric": "instance-sim-time/wall-s (64-instance ensemble)",
        "value": round(value, 4), "unit": "instance-sim-time/wall-s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * out["wall_s"] / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded ensemble: grids, geometries, fields and two-domain states)",
        "config": {"workload": f"{len(members)} independent instances, grids "
                               "256x256/512x256/768x512/1024x512, ellipse/ring 0.2-0.8 active, "
                               "in-plane h_ext <= 0.1, periodic FFT demag 2.0, i4 physics, "
                               "rtol=atol=1e-4, own step control each",
                   "cells": ncell, "nst": out["nst"], "nli": out["nli"], "nfe": out["nfe"],
                   "failed": out["failed"],
                   "sequential_sample_inst_time_per_wall_s": round(seq, 4) if seq else None,
                   "batched_f": batched},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(rec), flush=True)


def training_states(geom, make_ic, integrator, seeds=(2,), times=(1.0, 2.0, 3.0, 4.0, 5.0)):
    """states of trajectories from other seeds than the timed one (the
    learned correction never sees the benchmarked trajectory)"""
    states = []
    for sd in seeds:
        y = make_ic(sd)
        states.append(y)
        tr = integrator(y)
        for t in times:
            tr.advance(t)
            states.append(tr.state())
        tr.close()
    return states


def train_lpc(la, sim, states, steps):
    """train the learned correction on exact Newton-system pairs of the
    FP64 operator at the given states; returns (weights, info)"""
    t0 = time.perf_counter()
    lpc = la.LearnedPrecond(sim)
    lpc.init_random(seed=0, std=1e-3)
    hist = lpc.train(states, steps=steps // 2, lr=2e-3, gamma_range=(1e-3, 2e-1), seed=1)
    hist += lpc.train(states, steps=steps - steps // 2, lr=5e-4, gamma_range=(1e-3, 2e-1), seed=2)
    w = lpc.weights()
    rng = np.random.default_rng(123)
    e_bj, e_l = [], []
    for i in range(4):            # held-out pairs: relative error of the preconditioned solution
        y = states[(2 * i + 1) % len(states)]
        g = float(np.exp(rng.uniform(np.log(1e-3), np.log(2e-1))))
        sc = lpc.make_pair(y, g, rng)
        o = lpc.forward(g, sc)[:, :3].astype(np.float64)
        u = lpc._u.get().reshape(3, -1).T * sc
        x = lpc._x.get().reshape(3, -1).T * sc
        e_bj.append(np.linalg.norm(u - x) / np.linalg.norm(x))
        e_l.append(np.linalg.norm(u + o - x) / np.linalg.norm(x))
    lpc.close()
    info = {"train_steps": steps, "train_s": round(time.perf_counter() - t0, 1),
            "train_states": len(states), "loss_first": float(np.mean(hist[:10])),
            "loss_last": float(np.mean(hist[-10:])),
            "heldout_rel_err_block_jacobi": round(float(np.mean(e_bj)), 4),
            "heldout_rel_err_learned": round(float(np.mean(e_l)), 4)}
    return w, info


def lpc_apply_us(la, sim, y0, reps=30):
    """one learned-correction application (fused kernel), HIP events"""
    L = la.lib()
    lpc = la.LearnedPrecond(sim)
    lpc.init_random(seed=1, std=1e-3)
    vy, vu, vz = sim.vec(y0), sim.vec(np.random.default_rng(0).standard_normal(sim.neq)), sim.vec()
    e0, e1 = L.LLG_EventCreate(), L.LLG_EventCreate()
    for _ in range(3):
        L.LLG_LpcApply(lpc.h, lpc.dptr(vy), 0.03, lpc.dptr(vu), lpc.dptr(vz))
    L.LLG_EventRecord(e0, sim.stream)
    for _ in range(reps):
        L.LLG_LpcApply(lpc.h, lpc.dptr(vy), 0.03, lpc.dptr(vu), lpc.dptr(vz))
    L.LLG_EventRecord(e1, sim.stream)
    L.LLG_DeviceSynchronize()
    us = 1e3 * L.LLG_EventElapsedMs(e0, e1) / reps
    L.LLG_EventDestroy(e0)
    L.LLG_EventDestroy(e1)
    lpc.close()
    return round(us, 2)


def north_star_ab(la, cfg, geom, y0, y_plain, out_plain, value_plain, t_end, ropts, args,
                  make_ic, kernels):
    """The forms beside the plain FP64 block-Jacobi run, on the same workload
    and window: FP32 spectral demag, the learned correction, and both; each
    with its wall rate and ratio to the plain run, Krylov counts, and the
    max |dm| of its final state -- and of the plain run's -- against an FP64
    reference of the same trajectory at rtol = atol = 1e-6 (the accuracy the
    plain method reaches at rtol = 1e-4 is the bar)."""
    L = la.lib()

    def make_sim(fp32):
        c = type(cfg).from_buffer_copy(cfg)
        c.demag_fp32 = fp32
        sim = la.Simulation(c, geom)
        dm = L.LLG_GetDemag(sim.ctx)
        if dm and kernels:
            L.Demag_SetFftKernels(dm, *kernels)
        return sim

    sim64 = make_sim(0)
    t0 = time.perf_counter()
    states = training_states(geom, make_ic, lambda y: sim64.integrator(y, **ropts))
    w, info = train_lpc(la, sim64, states, args.lpc_steps)
    info["state_s"] = round(time.perf_counter() - t0 - info["train_s"], 1)
    print(f"[bench] learned correction trained: {info}", file=sys.stderr, flush=True)
    ref = sim64.integrator(y0, **dict(ropts, rtol=1e-6, atol=1e-6))
    ref.advance(t_end)
    y_ref = ref.state()
    ref.close()
    act = np.tile(geom.active.astype(bool), 3)
    dm_plain = float(np.max(np.abs(y_plain - y_ref)[act]))
    forms = {}
    for name, fp32, learned in (("fp32_demag", 1, False), ("learned_pc", 0, True),
                                ("learned_pc_fp32_demag", 1, True)):
        sim = sim64 if fp32 == 0 else make_sim(1)
        lpc = None
        if learned:
            lpc = la.LearnedPrecond(sim)
            lpc.set_weights(w)
            lpc.enable(True)
        if args.warmup > 0:
            wr = sim.integrator(y0, **ropts)
            wr.advance(1.0)
            wr.close()
        run = sim.integrator(y0, **ropts)
        o = run.advance(t_end)
        y = run.state()
        run.close()
        if lpc is not None:
            lpc.close()
        if sim is not sim64:
            sim.close()
        v = t_end / o["wall_s"]
        forms[name] = {"value": round(v, 4), "ratio_to_plain_fp64": round(v / value_plain, 3),
                       "nst": o["nst"], "nli": o["nli"], "nni": o["nni"], "ncfn": o["ncfn"],
                       "max_dm_vs_rtol1e-6_reference": float(f"{np.max(np.abs(y - y_ref)[act]):.3e}"),
                       "max_dm_vs_plain": float(f"{np.max(np.abs(y - y_plain)[act]):.3e}"),
                       "m_norm_drift": m_norm_drift(y, geom.active)}
        print(f"[bench] form {name}: {forms[name]}", file=sys.stderr, flush=True)
    apply_us = lpc_apply_us(la, sim64, y0)
    sim64.close()
    return {"window": f"t in [0, {t_end:g}], same initial state, rtol=atol=1e-4",
            "plain_fp64": {"value": round(value_plain, 4), "nst": out_plain["nst"],
                           "nli": out_plain["nli"], "nni": out_plain["nni"],
                           "ncfn": out_plain["ncfn"],
                           "max_dm_vs_rtol1e-6_reference": float(f"{dm_plain:.3e}"),
                           "m_norm_drift": m_norm_drift(y_plain, geom.active),
                           "m_norm_drift_rtol1e-6_reference": m_norm_drift(y_ref, geom.active)},
            "forms": forms, "learned_pc": info, "learned_pc_apply_us": apply_us}


def gpu_same_window(la, sim, y0, t_end, ropts, dt=0.01):
    """the GPU over the CPU baseline's window, on its output cadence"""
    run = sim.integrator(y0, **dict(ropts, t_out=dt))
    o = run.advance(t_end)
    run.close()
    return {"t_end": round(t_end, 4), "t_out": dt, "value": round(o["t"] / o["wall_s"], 4),
            "nst": o["nst"], "unit": "sim-time/wall-s"}


def hysteresis_bench(la, args):
    """The hysteresis sweep on the 4096x4096 polycrystal (400 grains, 10 %
    dead grains as holes, compact execution, zero-padded 8192x8192 FFT demag
    of strength 1, the i3 constants): h_ext along +x ramped from -0.2 to
    +0.2 in `--stages` quasi-static stages of `--t-stage` each (LLG_RunStage:
    the field set in the RHS, Jv and preconditioner term sets, the
    integrator re-initialised at the current state).  The plain FP64 run
    and the FP32-spectral-demag form run side by side, stage by stage, from
    the same random initial state; value = simulated time / summed stage
    wall time of the FP64 run, and each stage reports the FP32 form's max
    |dm| against the FP64 state at the stage's end.  Host copies of the
    states for that comparison sit outside the timed stages."""
    ng = ny = 4096
    t0 = time.perf_counter()
    geom = la.Geometry.polycrystal(ng, ny, num_grains=400, dead_frac=0.10, seed=5)
    y0 = geom.ic_random_sphere(seed=3)
    fields = np.zeros((args.stages, 3))
    fields[:, 0] = np.linspace(-0.2, 0.2, args.stages)
    which = la.TERMS_RHS | la.TERMS_JV | la.TERMS_PC
    dead = int((geom.active == 0).sum())
    print(f"[bench] hysteresis: 4096x4096 polycrystal, {dead} hole cells; geometry "
          f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    forms = {}
    for name, fp32 in (("fp64", 0), ("fp32_demag", 1)):
        cfg = la.preset("polycrystal", ng=ng, ny=ny, demag=1.0, compact=True, padded=True,
                        happ=tuple(fields[0]))
        cfg.demag_fp32 = fp32
        sim = la.Simulation(cfg, geom)
        run = sim.integrator(y0, rtol=1e-4, atol=1e-4, maxl=5, max_order=5,
                             gstype=args.gstype, t_out=args.t_stage, mxstep=1000000)
        forms[name] = {"sim": sim, "run": run, "wall": 0.0, "nst": 0, "nli": 0, "m": []}
    per_stage = []
    for k in range(args.stages):
        ys = {}
        for name, f in forms.items():
            m, r = f["run"].stage(fields[k], args.t_stage, which=which)
            f["wall"] += r["wall_s"]
            f["nst"] += r["nst"]
            f["nli"] += r["nli"]
            f["m"].append(m.tolist())
            ys[name] = f["run"].state()
        dm = float(np.max(np.abs(ys["fp32_demag"] - ys["fp64"])))
        per_stage.append({"hx": round(float(fields[k, 0]), 4), "mx": round(forms["fp64"]["m"][-1][0], 5),
                          "max_dm_fp32": dm})
        print(f"[bench] stage {k + 1}/{args.stages} hx={fields[k, 0]:+.3f}: m_x "
              f"{forms['fp64']['m'][-1][0]:+.4f}, fp32 max|dm| {dm:.2e}", file=sys.stderr,
              flush=True)
    span = args.stages * args.t_stage
    out = {}
    for name, f in forms.items():
        out[name] = {"value": round(span / f["wall"], 4), "wall_s": round(f["wall"], 3),
                     "nst": f["nst"], "nli": f["nli"]}
        f["run"].close()
        f["sim"].close()
    value = out["fp64"]["value"]
    rec = {
        "metric": "sim-time/wall-s at rtol=1e-4 (4096x4096 polycrystal hysteresis sweep)",
        "value": value, "unit": "sim-time/wall-s", "n_gpus": 1, "steps": args.stages,
        "warmup": 0, "ms_per_step": round(1e3 * out["fp64"]["wall_s"] / args.stages, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded polycrystal, random initial m)",
        "config": {"workload": "4096x4096 polycrystal, 400 grains, 10% dead (holes), compact, "
                               "zero-padded 8192x8192 FFT demag 1.0, h_ext along +x from -0.2 to "
                               f"+0.2 in {args.stages} stages of {args.t_stage} time units, "
                               "rtol=atol=1e-4, Krylov 5, block-Jacobi",
                   "stages": args.stages, "t_stage": args.t_stage,
                   "forms": out, "per_stage": per_stage,
                   "fp32_over_fp64": round(out["fp32_demag"]["value"] / value, 3)},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(rec), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed steps (1.0 sim time each)")
    ap.add_argument("--warmup", type=int, default=2, help="untimed warmup steps (separate run)")
    ap.add_argument("--ng", type=int, default=1536)
    ap.add_argument("--ny", type=int, default=512)
    ap.add_argument("--periodic", action="store_true", help="reference periodic demag instead of padded")
    ap.add_argument("--profile-kernel", default="demag_col", choices=sorted(KERNELS))
    ap.add_argument("--profile-every", type=int, default=8,
                    help="time one launch in this many of the roofline kernel (HIP events "
                         "on its stream, inside the timed region)")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--gstype", type=int, default=2,
                    help="Gram-Schmidt: 2 classical (default: one host wait per Krylov "
                         "iteration), 1 modified (the reference's choice at this size), 0 auto")
    ap.add_argument("--sync-mode", type=int, default=1,
                    help="reduction wait: 1 poll the pinned result word (default), 0 stream sync")
    ap.add_argument("--fp32-demag", action="store_true", help="FP32 spectral demag form")
    ap.add_argument("--workload", default="square_hole",
                    choices=["square_hole", "ref512", "polycrystal", "ensemble", "hysteresis"],
                    help="square_hole: 1536x512 cells, padded demag (default); ref512: the "
                         "reference's published 512x512 periodic i1 run (use --steps 1000); "
                         "polycrystal: 1024x1024 Voronoi polycrystal (72 grains, axes uniform "
                         "on the sphere, 16 %% dead), compact execution, windowed demag 1.0, "
                         "random initial m; ensemble: 64 independent instances run concurrently")
    ap.add_argument("--no-ab", action="store_true",
                    help="skip the FP32-demag / learned-preconditioner forms measured beside "
                         "the plain run (square_hole and polycrystal workloads)")
    ap.add_argument("--instances", type=int, default=64, help="ensemble size")
    ap.add_argument("--stages", type=int, default=40, help="hysteresis: field stages")
    ap.add_argument("--t-stage", type=float, default=0.05,
                    help="hysteresis: simulated time per stage")
    ap.add_argument("--learned-pc", action="store_true",
                    help="train the learned block-Jacobi correction (on states of a "
                         "trajectory from another seed, untimed) and use it in the timed run")
    ap.add_argument("--lpc-steps", type=int, default=400, help="training steps")
    ap.add_argument("--spgmr", default="lagged",
                    choices=["pipelined", "lagged", "lagged-host", "host", "device"],
                    help="lagged: SPGMR with lagged normalisation, one pass (projection, "
                         "operator chain and every Gram-Schmidt sum) per iteration, the solve "
                         "queued whole with each iteration's scalar step on the device and "
                         "one host wait per solve (default); lagged-host: the same solve "
                         "stepped by the host, one wait per iteration (the same bits); "
                         "pipelined: lagged-host with each pass queued behind a gate before "
                         "the previous pass's sums are back; host: the standard loop with "
                         "one-pass kernels, bit-equal to the plain op sequence; device: the "
                         "standard loop queued whole on the device, one host wait per solve")
    ap.add_argument("--fused-extra", type=int, default=0,
                    help="bits added to the integrator's fused-kernel mask (64: the first "
                         "SPGMR pass forms the Newton residual where it reads it; 128: the "
                         "first Newton update forms the SPGMR solution where it reads it; "
                         "every --spgmr mask already holds 256, the Newton residual formed "
                         "in the RHS pass, and 512, each step's completion left to the next "
                         "predictor's pass)")
    ap.add_argument("--red-blocks", type=int, default=0,
                    help="partial workgroups per reduction (NVHip_SetReduceConfig; 0: the "
                         "library default)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the launch-shape autotuner at setup (untimed)")
    ap.add_argument("--fft-kernels", default=None,
                    help="demag FFT kernel families ROW,COL kept after the autotuner "
                         "(Demag_SetFftKernels: 0 / 1 register-resident variants, 2 / 3 / 4 "
                         "variant 0 on 2 / 4 / 8 columns per workgroup (same bits as 0), "
                         "5 / 6 variant 1 with one transform per wave (same bits as 1), -1 "
                         "Stockham), or 'auto' for the autotuner's choice.  The families "
                         "round differently, and over 100 time units the step count the "
                         "metric depends on follows the rounding (4,109 steps with 0,5; "
                         "4,202 with 0,3), so the FP64 default pins 0,5, the fastest "
                         "column kernel (the FP32-spectra default: auto)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1 "
                                   "(the multi-GPU decomposition is design-only)"}))
        sys.exit(2)

    import llg_amd as la
    la.lib()   # raises when the native library is missing
    la.lib().NVHip_SetSyncMode(args.sync_mode)
    if args.workload == "ensemble":
        return ensemble_bench(la, args)
    if args.workload == "hysteresis":
        return hysteresis_bench(la, args)

    ref = args.workload == "ref512"
    poly = args.workload == "polycrystal"
    if poly:
        ng = ny = 1024
        padded = False
        geom = la.Geometry.polycrystal(ng, ny, num_grains=72, dead_frac=0.16, seed=20251104,
                                       z_bias=1.0)
        cfg = la.preset("polycrystal", ng=ng, ny=ny, demag=1.0, compact=True)
        make_ic = lambda sd: geom.ic_random_sphere(seed=sd)   # noqa: E731
        y0 = make_ic(1)
    elif ref:
        ng = ny = 512
        padded = False
        geom = la.Geometry.full(ng, ny)
        cfg = la.preset("i1", ng=ng, ny=ny, demag=2.0)
        make_ic = lambda sd: geom.ic_stripes(left=0.25, right=0.75, eps=0.01,   # noqa: E731
                                             seed=12345 if sd == 1 else 777 + sd)
        y0 = make_ic(1)
    else:
        ng, ny = args.ng, args.ny
        padded = not args.periodic
        geom = la.Geometry.square_hole(ng, ny, halfside_frac=222.0 / 512.0)
        cfg = la.preset("square_hole", ng=ng, ny=ny, padded=padded)
        make_ic = lambda sd: geom.ic_split_x(max_angle=0.05, seed=sd)   # noqa: E731
        y0 = make_ic(1)
    cfg.demag_fp32 = 1 if args.fp32_demag else 0
    print(f"[bench] workload {args.workload} {ng}x{ny} cells, hole {int((geom.active == 0).sum())} "
          f"cells, {'padded' if padded else 'periodic'} demag; building context",
          file=sys.stderr, flush=True)
    t_setup = time.perf_counter()
    sim = la.Simulation(cfg, geom)
    tune = None
    if not args.no_autotune:
        ch = (C.c_int * 8)()
        if la.lib().LLG_Autotune(sim.ctx, 5, ch) == 0:
            tune = {"stencil_block": [ch[0], ch[1]], "demag_row_block": ch[2],
                    "demag_col_block": ch[3], "demag_row_pad": ch[4], "demag_col_pad": ch[5],
                    "demag_row_kernel": ch[6], "demag_col_kernel": ch[7]}
    if args.fft_kernels is None:
        args.fft_kernels = "auto" if args.fp32_demag else "0,5"
    if args.fft_kernels != "auto":
        rk, ck = (int(v) for v in args.fft_kernels.split(","))
        dm = la.lib().LLG_GetDemag(sim.ctx)
        if dm and la.lib().Demag_SetFftKernels(dm, rk, ck) != 0:
            raise SystemExit(f"--fft-kernels {rk},{ck}: no such kernel for this grid")
        if tune is not None:
            tune["demag_row_kernel"], tune["demag_col_kernel"] = rk, ck
            tune["fft_kernels"] = "pinned"
    setup_s = time.perf_counter() - t_setup
    step = 1.0
    ropts = dict(rtol=1e-4, atol=1e-4, maxl=5, gstype=args.gstype, max_order=5, t_out=0.1,
                 mxstep=100000)
    ropts["fused"] = {"pipelined": 823, "lagged": 16151, "lagged-host": 791, "host": 775,
                      "device": 783}[args.spgmr] | \
        args.fused_extra
    if args.red_blocks:
        la.lib().NVHip_SetReduceConfig(args.red_blocks, 1)

    lpc_info = None
    lpc = None
    if args.learned_pc:
        states = training_states(geom, make_ic, lambda y: sim.integrator(y, **ropts))
        wts, lpc_info = train_lpc(la, sim, states, args.lpc_steps)
        lpc = la.LearnedPrecond(sim)
        lpc.set_weights(wts)
        lpc.enable(True)
        print(f"[bench] learned correction trained: {lpc_info}", file=sys.stderr, flush=True)

    if args.warmup > 0:
        w = sim.integrator(y0, **ropts)
        w.advance(args.warmup * step)
        w.close()
        print(f"[bench] warmup {args.warmup} steps done", file=sys.stderr, flush=True)

    run = sim.integrator(y0, **ropts)
    L = la.lib()
    L.LLG_ProfileStart(KERNELS[args.profile_kernel], 200000)
    # every 8th launch of the kernel carries the timing events: an event pair
    # delays the next dispatch (about 12 us per demag on this workload)
    L.LLG_ProfileSetEvery(args.profile_every)
    out = run.advance(args.steps * step)
    tot_ms, nl, dropped = C.c_double(), C.c_long(), C.c_long()
    L.LLG_ProfileStop(C.byref(tot_ms), C.byref(nl), C.byref(dropped))
    y_final = run.state()
    run.close()
    wall = out["wall_s"]
    value = args.steps * step / wall
    print(f"[bench] timed {args.steps} steps in {wall:.3f} s: nst={out['nst']} nfe={out['nfe']} "
          f"nli={out['nli']} nni={out['nni']} netf={out['netf']} ncfn={out['ncfn']}",
          file=sys.stderr, flush=True)

    ncell = ng * ny
    cv = C.c_int(0)
    dmh = L.LLG_GetDemag(sim.ctx)
    if dmh:
        L.Demag_GetFftKernels(dmh, None, C.byref(cv))
    nbytes = algorithmic_bytes(args.profile_kernel, ng, ny, ncell, padded, cv.value)
    avg_s = (tot_ms.value / 1e3) / max(nl.value, 1)
    achieved = nbytes / avg_s / 1e9 if nl.value else 0.0
    roof = {"bound": "hbm", "kernel": args.profile_kernel, "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(args.profile_kernel,
                                        "square_hole" if (not ref and padded and ng == 1536
                                                          and ny == 512 and not args.fp32_demag)
                                        else "other"),
            "bytes_per_launch": nbytes, "avg_launch_us": round(avg_s * 1e6, 2),
            "launches": nl.value, "timed_every": args.profile_every}

    ab = None
    if not args.no_ab and not ref and not args.learned_pc and not args.fp32_demag:
        print("[bench] forms beside the plain run", file=sys.stderr, flush=True)
        kern = None if args.fft_kernels == "auto" else tuple(int(v) for v in
                                                             args.fft_kernels.split(","))
        ab = north_star_ab(la, cfg, geom, y0, y_final, out, value, args.steps * step, ropts,
                           args, make_ic, kern)

    cpu = None
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(la, cfg, geom, y0, args.cpu_budget,
                           args.cpu_threads or min(16, os.cpu_count() or 1))
        # the GPU over the same window and output cadence as the CPU sample
        cpu["gpu_same_window"] = gpu_same_window(la, sim, y0, cpu["t_end"], ropts)
        print(f"[bench] gpu over the cpu window: {cpu['gpu_same_window']}", file=sys.stderr,
              flush=True)

    if poly:
        workload = ("polycrystal 1024x1024 cells, 72 grains, axes uniform on the sphere, 16% "
                    "dead grains as holes, compact execution, windowed periodic FFT demag 1.0, "
                    "rtol=atol=1e-4, Krylov 5, block-Jacobi")
        data = "synthetic (seeded Voronoi geometry and random initial magnetisation)"
    elif ref:
        workload = ("reference reading of 1536x512: 512x512 cells, periodic 512x512 FFT demag "
                    "2.0, i1 physics, head-on stripes, rtol=atol=1e-4, Krylov 5, block-Jacobi")
        data = "synthetic (seeded head-on stripes, the reference's i1 initial state)"
    else:
        workload = (f"square-hole {ng}x{ny} cells, "
                    f"{'zero-padded %dx%d' % (2 * ng, 2 * ny) if padded else 'periodic'} "
                    "FFT demag 4.0, rtol=atol=1e-4, Krylov 5, block-Jacobi")
        data = "synthetic (seeded two-domain state on the square-hole geometry)"
    rec = {
        "metric": ("sim-time/wall-s at rtol=1e-4 (1024x1024 polycrystal, compact)" if poly
                   else "sim-time/wall-s at rtol=1e-4 (1536x512+demag)"),
        "value": round(value, 4),
        "unit": "sim-time/wall-s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(value / REF512_PUBLISHED, 3) if ref else None,
        "dtype": "f64" if not args.fp32_demag else "f64 (demag spectra f32)",
        "data": data,
        "config": {"workload": workload,
                   "t_end": args.steps * step,
                   "m_norm_drift": m_norm_drift(y_final, geom.active),
                   "spgmr_iters_per_step": round(out["nli"] / max(out["nst"], 1), 3),
                   "nli_per_nni": round(out["nli"] / max(out["nni"], 1), 3),
                   "nst": out["nst"], "nfe": out["nfe"], "nni": out["nni"], "nli": out["nli"],
                   "nsetups": out["nsetups"], "netf": out["netf"], "ncfn": out["ncfn"],
                   "host_syncs": out["nsyncs"], "setup_s": round(setup_s, 2),
                   "gs": {0: "auto", 1: "modified", 2: "classical"}[args.gstype],
                   "sync_mode": args.sync_mode,
                   "preconditioner": "block-Jacobi + learned correction" if args.learned_pc
                   else "block-Jacobi", "learned_pc": lpc_info, "autotune": tune,
                   "spgmr": {"pipelined": "lagged normalisation, gated pipeline",
                             "lagged": "lagged normalisation, device-stepped",
                             "lagged-host": "lagged normalisation, host-stepped",
                             "host": "host-driven",
                             "device": "device-resident"}[args.spgmr]},
        "roofline": roof,
        "cpu_baseline": cpu,
        "forms": ab,
    }
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
