"""Concurrent ensemble of independent instances (own grid, geometry,
field, step control) on one GPU.  Each instance must compute exactly what
it computes alone: final states bitwise equal and identical integrator
statistics against single-instance runs.  No reference counterpart (the
reference runs one instance per process)."""
import numpy as np
import pytest

import llg_amd as la

pytestmark = pytest.mark.gpu


def test_ensemble_members_match_single_runs():
    members = la.ensemble_members(n=5, seed=11, grids=((64, 48), (48, 32), (96, 64)),
                                  frac=(0.3, 0.7))
    ens = la.Ensemble(members, nthreads=0, rtol=1e-4, atol=1e-4)
    out = ens.advance(0.5)
    print(f"[gpu] ensemble: {out}")
    assert out["failed"] == 0 and abs(out["inst_time"] - 5 * 0.5) < 1e-9
    for i, (cfg, geom, y0) in enumerate(members):
        sim = la.Simulation(cfg, geom)
        run = sim.integrator(y0, rtol=1e-4, atol=1e-4)
        single = run.advance(0.5)
        y1 = run.state()
        r = ens.result(i)
        assert r["nst"] == single["nst"] and r["nli"] == single["nli"], i
        assert np.array_equal(ens.state(i), y1), i
        run.close()
        sim.close()
    ens.close()


def test_ensemble_with_fewer_threads_than_instances():
    members = la.ensemble_members(n=4, seed=3, grids=((32, 32), (64, 32)))
    ens = la.Ensemble(members, nthreads=2, rtol=1e-4, atol=1e-4)
    ens.advance(0.2)
    ens.advance(0.4)
    ref = la.Ensemble(members, nthreads=4, rtol=1e-4, atol=1e-4)
    ref.advance(0.4)
    for i in range(4):
        assert np.array_equal(ens.state(i), ref.state(i))
    ens.close()
    ref.close()
