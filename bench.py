#!/usr/bin/env python3
"""Throughput of one VEBISPH time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload cylinder_1M]
                    [--precision mixed|fp64] [--no-compare] [--no-cpu-baseline]

A "step" is one full div_free step (binning, mirrors, neighbour lists, kernel weights, shifting,
prediction with the log-conformation update, PPE assembly and BiCGSTAB solve, pressure update,
advection) over the whole particle set, with the state resident in HBM.  The workload is the flow
past a fixed cylinder in an 8x4 periodic channel with ~1.0M particles (synthetic, seeded lattice).

`value` is particle-steps/s of the requested precision (mixed by default: FP32 matvecs inside the
FP64 BiCGSTAB); the same workload is then run in plain FP64 and reported beside it, with the ratio.
The roofline entry is the BiCGSTAB SpMV kernel, timed with HIP events on the solver's stream.
Prints one JSON line on stdout; progress goes to stderr.
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
SPMV_KERNEL = {"mixed": "k_bi_spmv_p<1, true, true> (BiCGSTAB SpMV, FP32 values read cacheably; "
                        "<1, true, false> where the value plane exceeds the Infinity Cache)",
               "fp64": "k_bi_spmv_p<0, true, false> (BiCGSTAB SpMV, FP64 values)"}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "spmv_traffic_cylinder_1M.json")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic(workload, precision):
    """HBM bytes per SpMV launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE runs."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        if t.get("workload") == workload and precision in t:
            return t[precision]["traffic_bytes"]
    except (OSError, ValueError, KeyError):
        pass
    return None


def host_cpu():
    """The host's CPU: model name, logical CPUs (nproc), physical cores (distinct core/socket pairs
    of lscpu -p), the CPUs this process may run on, and the thread count the CPU baseline uses: the
    physical cores available to the process, capped by OMP_NUM_THREADS where the job sets it (the
    host's CPU share per GPU)."""
    import subprocess
    info = {"model": None, "nproc": os.cpu_count(), "physical_cores": None,
            "affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                info["model"] = line.split(":", 1)[1].strip()
        par = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=10).stdout
        cores = {ln for ln in par.splitlines() if ln and not ln.startswith("#")}
        info["physical_cores"] = len(cores) or None
    except (OSError, subprocess.SubprocessError):
        pass
    avail = info["affinity"] or info["nproc"] or 1
    threads = min(info["physical_cores"] or avail, avail)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        threads = min(threads, int(omp))
        info["omp_num_threads"] = int(omp)
    info["threads"] = max(1, threads)
    return info


def cpu_baseline(case, state, time0, itime0, threads, cpu_info, budget_s=20.0):
    """The reference itself (oracle/_ref, compiled from its Fortran sources) on the host cores,
    stepping the same workload from the same state: one untimed warm-up step, then steps timed for
    about `budget_s` seconds; falls back to the C restatement (one thread)."""
    def sample(S, kind, cores, label):
        t = time.perf_counter()
        S.step()                                   # warm-up: page-in, thread start-up, first touch
        warm = time.perf_counter() - t
        dts, its = [], []
        while True:
            t = time.perf_counter()
            S.step()
            dts.append(time.perf_counter() - t)
            its.append(S.scalars()["LS_iters"])
            if sum(dts) + dts[-1] > budget_s or len(dts) >= 10:
                break
        tot = sum(dts)
        return {"value": case.npfb * len(dts) / tot, "unit": "particle-steps/s", "cores": cores,
                "kind": kind, "host": cpu_info,
                "sample": f"{len(dts)} timed step(s) of the same workload from the GPU state after "
                f"warmup, after one untimed warm-up step ({warm:.1f} s) ({label}, BiCGSTAB iterations "
                f"{its}, {tot:.1f} s)"}
    try:
        from oracle import refwrap
        if refwrap.available():
            R = refwrap.RefSolver(case, time=time0, itime=itime0, state=state)
            R.set_threads(threads)
            return sample(R, "reference", threads, f"reference Fortran, OpenMP {threads} threads")
    except Exception as e:  # noqa: BLE001 - the baseline is reported, never fatal
        log(f"reference baseline unavailable: {e}")
    try:
        from oracle import oraclewrap
        O = oraclewrap.OracleSolver(case, time=time0, itime=itime0, state=state)
        return sample(O, "port", 1, "C restatement, 1 thread")
    except Exception as e:  # noqa: BLE001
        log(f"port baseline unavailable: {e}")
        return None


def force_diff(force, other):
    """max |F - F_other| / |F_other| over circles and components (None without both)."""
    if not force or not other or not other.get("force_on_cylinder"):
        return None
    import numpy as np
    a, b = np.array(force), np.array(other["force_on_cylinder"])
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def run(case, precision, steps, warmup):
    """Warmup, then exactly `steps` timed steps; the SpMV roofline figures after the timed region."""
    from vebisph_amd import _lib, solver
    prec = _lib.PREC_MIXED if precision == "mixed" else _lib.PREC_FP64
    s = solver.Solver(case, precision=prec)
    for _ in range(warmup):
        s.div_free()
    sc0 = s.scalars()                      # synchronises
    state = s.get_state()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.div_free()                       # each step ends with a stream synchronisation
    t1 = time.perf_counter()
    sc = s.scalars()
    force = s.circle_forces()[0].tolist() if case.nb_circles else None
    true_res = ppe_true_residual(s, case)
    us, by = s.profile_spmv(reps=20)
    adj = adjoint_cost(s, case)
    elapsed = t1 - t0
    res = {"value": case.npfb * steps / elapsed, "ms_per_step": 1e3 * elapsed / steps,
           "ls_iters_last_step": sc["LS_iters"], "ls_residual_last_step": sc["LS_residual"],
           "ls_true_residual_last_step": true_res,
           "spmv_us": us, "spmv_bytes": by, "spmv_GBs": by / (us * 1e-6) / 1e9, "force": force,
           "adjoint_ppe": adj}
    log(f"{precision}: {res['value']:.4g} particle-steps/s, {res['ms_per_step']:.1f} ms/step, "
        f"BiCGSTAB {sc['LS_iters']} iterations (residual {sc['LS_residual']:.3g}, FP64 residual "
        f"{true_res:.3g}); SpMV {us:.1f} us, {res['spmv_GBs']:.0f} GB/s")
    log(f"{precision}: PPE solve {adj['forward_ms']:.1f} ms ({adj['forward_iters']} iterations, true "
        f"residual {adj['forward_true_residual']:.3g}), adjoint PPE solve {adj['adjoint_ms']:.1f} ms "
        f"({adj['adjoint_iters']} iterations, true residual {adj['adjoint_true_residual']:.3g}): "
        f"{adj['ratio']:.2f}x {adj['ratio_kind']}")
    return s, res, state, sc0


def ppe_true_residual(s, case):
    """|B - A p| / |B| in FP64 on the last step's assembled system and its solved pressure (in a
    closed domain with the constant component removed, the part the DC-projected iteration
    controls)."""
    import numpy as np
    _, _, _, rhs = s.ppe_ell(1)
    r = rhs - s.ppe_operator(s.get_state()["p"])
    if not case.i_open_domain:
        r = r - r.mean()
    return float(np.linalg.norm(r) / max(np.linalg.norm(rhs), 1e-300))


def adjoint_cost(s, case):
    """The PPE solve of one more step (assembly + BiCGSTAB, phase 10) against the adjoint PPE solve
    A^T lambda = g on the matrix it assembled (vb_ppe_adjoint_solve, FP64 BiCGSTAB; g a seeded
    mean-zero vector; the transposed map built beforehand, untimed; host copies of g and lambda
    included), each with its status and its true FP64 residual on the assembled system: |B - A p| /
    |B| for the forward solve (in a closed domain with the constant component removed, the part the
    DC-projected iteration controls; also given raw) and |g - A^T lambda| / |g| for the adjoint.  Both
    solves stop at the same tolerance (1e-8) and iteration limit (1000); where either stops at the
    limit the ratio compares equal iteration counts, so it is a per-iteration cost, and is labelled
    so.  Run after the timed steps."""
    import numpy as np
    for k in range(1, 10):
        s.phase(k)
    s.scalars()
    t = time.perf_counter()
    s.phase(10)
    sc = s.scalars()
    tf = time.perf_counter() - t
    fstatus = sc["status"]
    _, _, _, rhs = s.ppe_ell(1)
    p = s.get_state()["p"]
    rf = rhs - s.ppe_operator(p)
    nb = max(np.linalg.norm(rhs), 1e-300)
    fwd_true_raw = float(np.linalg.norm(rf) / nb)
    fwd_true = float(np.linalg.norm(rf - (0.0 if case.i_open_domain else rf.mean())) / nb)
    g = np.random.default_rng(5).standard_normal(case.npfb)
    if not case.i_open_domain:
        g -= g.mean()
    s.ppe_adjoint_solve(g, itermax=2)
    t = time.perf_counter()
    lam, it, res, ast = s.ppe_adjoint_solve(g)
    ta = time.perf_counter() - t
    adj_true = float(np.linalg.norm(g - s.ppe_operator(lam, transpose=True)) / np.linalg.norm(g))
    converged = fstatus == 0 and ast == 0
    return {"forward_ms": 1e3 * tf, "forward_iters": sc["LS_iters"], "forward_status": fstatus,
            "forward_residual": sc["LS_residual"], "forward_true_residual": fwd_true,
            "forward_true_residual_raw": fwd_true_raw,
            "adjoint_ms": 1e3 * ta, "adjoint_iters": it, "adjoint_status": ast, "adjoint_residual": res,
            "adjoint_true_residual": adj_true, "converged": converged, "ratio": ta / tf,
            "ratio_kind": "per solve (both converged)" if converged else
                          "per iteration (a solve stopped at the iteration limit)",
            "scope": "the PPE solve of one step and its adjoint solve"}


def adjoint_run_cost(workload, nsteps, every, precision="fp64"):
    """The adjoint's cost at an adjoint inversion's size: a plain forward run of nsteps
    steps (vb_div_free), the same run recorded for the adjoint (vb_adjoint_record: checkpoints every
    `every` steps, every step's pressure kept), and the reverse sweep (vb_adjoint_backward) for the
    velocity-mismatch loss 1/2 |u - u_target|^2 with u_target = 0.9 u_final.  The start is the fluid at
    rest with a pre-stretched conformation (vebisph_amd.adjoint.prestretched_state).  Reports the
    sweep's time over the plain forward's ("ratio"), and record + sweep over the plain forward."""
    import numpy as np
    from vebisph_amd import _lib, adjoint, cases, solver
    case = cases.make(workload)
    prec = _lib.PREC_MIXED if precision == "mixed" else _lib.PREC_FP64
    start = adjoint.prestretched_state(case)
    s = solver.Solver(case, precision=prec, state=start)
    s.scalars()
    t = time.perf_counter()
    for _ in range(nsteps):
        s.div_free()
    s.scalars()
    t_fwd = time.perf_counter() - t
    s.close()
    s = solver.Solver(case, precision=prec, state=start)
    t = time.perf_counter()
    fit, fst = s.adjoint_record(nsteps, every)
    t_rec = time.perf_counter() - t
    fin = s.get_state()
    nb = case.nbS
    bars = {"u": np.zeros(case.npfb), "v": np.zeros(case.npfb)}
    bars["u"][nb:] = 0.1 * fin["u"][nb:]
    bars["v"][nb:] = 0.1 * fin["v"][nb:]
    t = time.perf_counter()
    res = s.adjoint_backward(bars)
    t_bwd = time.perf_counter() - t
    s.close()
    out = {"workload": case.name, "particles": case.npfb, "steps": nsteps, "ckpt_every": every,
           "precision": precision, "forward_s": t_fwd, "record_s": t_rec, "backward_s": t_bwd,
           "ratio": t_bwd / t_fwd, "ratio_with_record": (t_rec + t_bwd) / t_fwd, "target_ratio": 4.0,
           "met": t_bwd / t_fwd <= 4.0,
           "forward_ppe_iters_mean": float(np.mean(fit)), "adjoint_ppe_iters_mean": float(np.mean(res["adj_iters"])),
           "adjoint_status": res["status"], "record_status": fst,
           "grads_Wi_beta_Ra": [float(g) for g in res["grads"]]}
    log(f"adjoint at {case.name}: forward {t_fwd:.2f} s, record {t_rec:.2f} s, backward {t_bwd:.2f} s over "
        f"{nsteps} steps: backward/forward {out['ratio']:.2f}x (with the record {out['ratio_with_record']:.2f}x); "
        f"PPE iterations forward {out['forward_ppe_iters_mean']:.0f}, adjoint {out['adjoint_ppe_iters_mean']:.0f}")
    return out


def drift(case, nsteps, every):
    """Mixed precision against FP64 over a run, beside the FP64 path's own spread: mixed, FP64, and
    FP64 from the start with every fluid coordinate moved by one ulp (seeded) step in lockstep; every
    `every` steps, max |a - FP64| / max |FP64| of each field for a = mixed ("mixed") and a = the
    moved FP64 run ("spread"), positions in units of dx, and the force on each circle (the C3
    configuration's drag, vb_circle_forces)."""
    import numpy as np
    from vebisph_amd import _lib, solver
    moved = solver.initial_state(case)
    rng = np.random.default_rng(7)
    for key in ("x", "y"):
        v = np.array(moved[key], dtype=np.float64)
        up = rng.random(v.size - case.nbS) < 0.5
        v[case.nbS:] = np.where(up, np.nextafter(v[case.nbS:], np.inf), np.nextafter(v[case.nbS:], -np.inf))
        moved[key] = v
    runs = {"mixed": solver.Solver(case, precision=_lib.PREC_MIXED),
            "spread": solver.Solver(case, precision=_lib.PREC_FP64, state=moved)}
    f = solver.Solver(case, precision=_lib.PREC_FP64)
    out = []
    rel = lambda a, b: float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
    for k in range(1, nsteps + 1):
        f.div_free()
        for r in runs.values():
            r.div_free()
        if k % every and k != nsteps:
            continue
        b = f.get_state()
        rec = {"step": k, "iters_fp64": f.scalars()["LS_iters"]}
        ff = f.circle_forces()[0] if case.nb_circles else None
        for name, r in runs.items():
            a = r.get_state()
            d = {key: rel(a[key], b[key]) for key in ("u", "v", "p", "tau", "trace")}
            d["pos_dx"] = float(max(np.max(np.abs(a["x"] - b["x"])), np.max(np.abs(a["y"] - b["y"]))) / case.dx)
            if case.nb_circles:
                d["force_rel"] = rel(r.circle_forces()[0], ff)
            rec[name] = d
        if case.nb_circles:
            rec["force_fp64"] = ff.tolist()
        out.append(rec)
        log(f"drift step {k}: " + "; ".join(
            f"{name} " + ", ".join(f"{key} {v:.1e}" for key, v in rec[name].items()) for name in runs))
    f.close()
    for r in runs.values():
        r.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="cylinder_1M")
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--no-compare", action="store_true", help="skip the run at the other precision")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--adjoint-workload", default="poiseuille_250k",
                    help="the adjoint cost's workload ('' skips it)")
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--adjoint-every", type=int, default=20)
    ap.add_argument("--adjoint-convection-workload", default="convection_1M",
                    help="the convection inversion's adjoint cost ('' skips it): mixed-precision forward")
    ap.add_argument("--adjoint-convection-steps", type=int, default=20)
    ap.add_argument("--adjoint-convection-every", type=int, default=10)
    ap.add_argument("--drift-steps", type=int, default=0,
                    help="also step mixed, FP64 and FP64 from a start one ulp away side by side this many "
                         "steps and report their differences from FP64")
    ap.add_argument("--drift-every", type=int, default=50)
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: the step has no multi-GPU path in this library (the box "
              "has one MI355X); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    from vebisph_amd import cases
    case = cases.make(a.workload)
    log(f"workload {case.name}: {case.npfb} particles ({case.nbS} boundary), warmup {a.warmup}, "
        f"steps {a.steps}")

    s, main_res, state, sc0 = run(case, a.precision, a.steps, a.warmup)
    copy_bytes = 1 << 30
    copy_us = s.profile_copy(copy_bytes, reps=10)
    copy_gbs = 2 * copy_bytes / (copy_us * 1e-6) / 1e9
    log(f"plain copy of {copy_bytes >> 20} MiB: {copy_us:.1f} us, {copy_gbs:.0f} GB/s")
    s.close()

    other = None
    if not a.no_compare:
        oprec = "fp64" if a.precision == "mixed" else "mixed"
        so, ores, _, _ = run(case, oprec, a.steps, a.warmup)
        so.close()
        other = {"precision": oprec, "value": ores["value"], "ms_per_step": ores["ms_per_step"],
                 "ls_iters_last_step": ores["ls_iters_last_step"],
                 "ls_residual_last_step": ores["ls_residual_last_step"],
                 "ls_true_residual_last_step": ores["ls_true_residual_last_step"],
                 "spmv_us": ores["spmv_us"], "spmv_GBs": ores["spmv_GBs"], "force_on_cylinder": ores["force"],
                 "adjoint_ppe": ores["adjoint_ppe"]}

    adj_run = None
    if a.adjoint_workload:
        adj_run = adjoint_run_cost(a.adjoint_workload, a.adjoint_steps, a.adjoint_every)
    adj_conv = None
    if a.adjoint_convection_workload:
        adj_conv = adjoint_run_cost(a.adjoint_convection_workload, a.adjoint_convection_steps,
                                    a.adjoint_convection_every, precision="mixed")

    drift_rec = drift(case, a.drift_steps, a.drift_every) if a.drift_steps > 0 else None

    cpu = None
    if not a.no_cpu_baseline:
        info = host_cpu()
        log(f"host CPU: {info}")
        cpu = cpu_baseline(case, state, sc0["time"], sc0["itime"], info["threads"], info)
        if cpu:
            log(f"CPU baseline ({cpu['kind']}, {cpu['cores']} cores): {cpu['value']:.4g} particle-steps/s")

    achieved = main_res["spmv_GBs"]
    line = {
        "metric": "particle-steps/s, FP32+FP64-refined vs FP64, 1M particles; adjoint/forward cost",
        "value": main_res["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": main_res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": case.name, "particles": case.npfb, "boundary_particles": case.nbS,
                   "precision": a.precision, "Wi": case.Wi, "beta": case.beta, "Re": case.Re,
                   "ppe_iters_last_step": main_res["ls_iters_last_step"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(case.name, a.precision),
                     "kernel": SPMV_KERNEL[a.precision],
                     "alg_bytes_per_launch": main_res["spmv_bytes"], "avg_us": main_res["spmv_us"],
                     "copy_GBs": copy_gbs},
        "compare": other,
        # force of the fluid on the cylinder after the same steps in both precisions (vb_circle_forces)
        "force_on_cylinder": main_res["force"],
        "force_rel_diff_vs_other_precision": force_diff(main_res["force"], other),
        "ratio_vs_other_precision": (main_res["value"] / other["value"]) if other else None,
        # the mixed/FP64 throughput target and whether this run met it
        "target_ratio": 2.0,
        "target_ratio_met": (main_res["value"] / other["value"] >= 2.0) if other and a.precision == "mixed" else None,
        # the last timed step's PPE solve: iterations, recurrence residual, FP64 residual |B - A p| / |B|
        "ppe_last_step": {a.precision: {"iters": main_res["ls_iters_last_step"],
                                        "residual": main_res["ls_residual_last_step"],
                                        "true_residual": main_res["ls_true_residual_last_step"]},
                          **({other["precision"]: {"iters": other["ls_iters_last_step"],
                                                   "residual": other["ls_residual_last_step"],
                                                   "true_residual": other["ls_true_residual_last_step"]}}
                             if other else {})},
        # adjoint/forward cost of the PPE solve at this precision's forward solve
        "adjoint_ppe": main_res["adjoint_ppe"],
        # adjoint/forward cost of a whole run at the adjoint inversion's size (checkpointed sweep)
        "adjoint": adj_run,
        # the same at the convection inversion's size, 1M particles, mixed-precision forward
        "adjoint_convection": adj_conv,
        # mixed against FP64 over a longer run (--drift-steps; not run by default)
        **({"drift": drift_rec} if drift_rec is not None else {}),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
