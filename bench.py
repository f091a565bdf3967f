#!/usr/bin/env python3
"""Throughput of one VEBISPH time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload cylinder_1M] [--precision mixed|fp64]

A "step" is one full div_free step (binning, mirrors, neighbour lists, kernel weights, shifting,
prediction with the log-conformation update, PPE assembly and BiCGSTAB solve, pressure update,
advection) over the whole particle set, with the state resident in HBM.  The workload is the
flow past a fixed cylinder in an 8x4 periodic channel with ~1.0M particles (synthetic, seeded).
Prints one JSON line; `value` is particle-steps/s of the requested precision.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E


def cpu_baseline(case, state, time0, itime0, threads):
    """The reference itself (oracle/_ref, compiled from its Fortran sources) on the host cores, one
    step of the same workload from the same state; falls back to the C restatement (one thread)."""
    try:
        from oracle import refwrap
        if refwrap.available():
            R = refwrap.RefSolver(case, time=time0, itime=itime0, state=state)
            R.set_threads(threads)
            t = time.perf_counter()
            R.step()
            dt = time.perf_counter() - t
            it = R.scalars()["LS_iters"]
            return {"value": case.npfb / dt, "unit": "particle-steps/s", "cores": threads,
                    "kind": "reference", "sample": f"1 step of the same workload from the GPU state after "
                    f"warmup (reference Fortran, OpenMP {threads} threads, {it} BiCGSTAB iterations, {dt:.1f} s)"}
    except Exception as e:  # noqa: BLE001 - the baseline is reported, never fatal
        print(f"[bench] reference baseline unavailable: {e}", file=sys.stderr)
    try:
        from oracle import oraclewrap
        O = oraclewrap.OracleSolver(case, time=time0, itime=itime0, state=state)
        t = time.perf_counter()
        O.step()
        dt = time.perf_counter() - t
        return {"value": case.npfb / dt, "unit": "particle-steps/s", "cores": 1, "kind": "port",
                "sample": f"1 step of the same workload (C restatement, 1 thread, {dt:.1f} s)"}
    except Exception as e:  # noqa: BLE001
        print(f"[bench] port baseline unavailable: {e}", file=sys.stderr)
        return None


def run(case, precision, steps, warmup):
    from vebisph_amd import solver, _lib
    prec = _lib.PREC_MIXED if precision == "mixed" else _lib.PREC_FP64
    s = solver.Solver(case, precision=prec)
    for _ in range(warmup):
        s.div_free()
    sc0 = s.scalars()                      # synchronises
    state = s.get_state()
    iters = []
    t0 = time.perf_counter()
    for _ in range(steps):
        s.div_free()                       # returns after the step's final synchronisation
        iters.append(s.scalars()["LS_iters"])
    t1 = time.perf_counter()
    us, by = s.profile_spmv(reps=20)
    return s, (t1 - t0), iters, (us, by), state, sc0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="cylinder_1M")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (see DESIGN.md on how the "
              "step would shard); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    from vebisph_amd import cases
    case = cases.make(a.workload)
    print(f"[bench] workload {case.name}: {case.npfb} particles ({case.nbS} boundary), "
          f"precision {a.precision}, warmup {a.warmup}, steps {a.steps}", file=sys.stderr, flush=True)
    s, elapsed, iters, (us, by), state, sc0 = run(case, a.precision, a.steps, a.warmup)
    value = case.npfb * a.steps / elapsed
    achieved = by / (us * 1e-6) / 1e9
    print(f"[bench] {value:.4g} particle-steps/s, {1e3 * elapsed / a.steps:.1f} ms/step, "
          f"BiCGSTAB iterations {iters}; SpMV {us:.1f} us, {achieved:.0f} GB/s", file=sys.stderr, flush=True)
    cpu = None
    if not a.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(case, state, sc0["time"], sc0["itime"], threads)
    line = {
        "metric": "particle-steps/s, FP32+FP64-refined vs FP64, 1M particles; adjoint/forward cost",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * elapsed / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": case.name, "particles": case.npfb, "boundary_particles": case.nbS,
                   "precision": a.precision, "Wi": case.Wi, "beta": case.beta,
                   "ppe_iters_per_step": iters},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_bi_spmv_p (BiCGSTAB ELL SpMV)", "alg_bytes_per_launch": by,
                     "avg_us": us},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
