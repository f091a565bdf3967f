"""The benchmark configurations at their own sizes against the reference itself: the flow past a
cylinder at 1,003,417 particles (FP64 and mixed precision; fields and the force on the cylinder) and
the Boussinesq convection box at 80,000 particles (400 x 200); and the Poiseuille configuration over
its whole 1000-step run against the oracle.

The reference (oracle/_ref, its Fortran compiled from its own sources, OpenMP on the host cores)
steps the same start state; its one-ulp spread -- the reference against itself from fluid positions
moved by one ulp, two such runs -- is the scale.  At 1M every BiCGSTAB solve, the reference's
included, stops at the 1000-iteration limit with a relative residual of 1e-6..1e-3 that depends on
rounding, so p (and the velocity update through grad p) agrees between any two orders of summation
only to that level.  The FP64 fields are held to ten times the spread (1e-10 where the spread is
below that); the mixed path's, whose kernel weights are rounded to FP32, to ten times the spread or
1e-5 (its precision target against FP64; positions 1e-9), whichever is larger.  The relative errors
are printed beside a 1e-5 mark for u, p, theta and tau.

The reference's predictor runs on one thread (see _reference_run).  The comparison is taken after
phase 11 of the last step (the pressure update; step 1 at 1M, where the reference takes ~25 s per
step, step 2 at 80k), where the force on the cylinder is defined: the oracle, started from the
reference's state before that step, rebuilds its lists and weights (phases 1-9, identical to the
reference's) and evaluates the force of each run's rows with the same code."""
import os

import numpy as np
import pytest

from tests.conftest import have_ref
from vebisph_amd import _lib, cases, solver

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "u", "v", "theta", "p", "tau", "trace")


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


def _threads():
    return max(1, min(16, os.cpu_count() or 1))


def _ulp_start(case, seed):
    """The start state with each fluid coordinate moved by one ulp up or down (walls stay: one ulp
    there can change the mirror set at a periodic corner)."""
    rng = np.random.default_rng(seed)
    s = solver.initial_state(case)
    nb = case.nbS
    for k in ("x", "y"):
        a = np.array(s[k], dtype=np.float64)
        up = rng.random(a.size - nb) < 0.5
        a[nb:] = np.where(up, np.nextafter(a[nb:], np.inf), np.nextafter(a[nb:], -np.inf))
        s[k] = a
    return s


def _reference_run(case, state, nsteps):
    """The reference: nsteps - 1 whole steps, then phases 1-11 of the last.  Returns the state before
    the last step with its scalars, the rows after its phase 11 (mirror rows included) and its
    scalars.  The predictor (phase 8) runs on one thread: its kick draws random numbers inside a
    parallel loop (predictor_mod.F90:369-371), so on more threads the draws reach the particles in an
    order that changes from run to run; on one they go in index order, as the GPU's kick assigns
    them."""
    from oracle import refwrap
    R = refwrap.RefSolver(case, state=state)

    def phases(last):
        for k in range(1, last + 1):
            R.set_threads(1 if k == 8 else _threads())
            R.phase(k)
    for _ in range(nsteps - 1):
        phases(12)
    s0 = R.get_state()
    sc0 = R.scalars()
    phases(11)
    sc = R.scalars()
    return s0, sc0, R.get_state(sc["np"]), sc


def _gpu_run(case, precision, nsteps):
    G = solver.Solver(case, precision=precision)
    for _ in range(nsteps - 1):
        G.div_free()
    for k in range(1, 12):
        G.phase(k)
    sc = G.scalars()
    rows = G.get_state(sc["np"])
    force = G.circle_forces()[0] if case.nb_circles else None
    G.close()
    return rows, sc, force


def _floor(name, key):
    """The bar below which no spread is needed: FP64 1e-10; the mixed path (FP32 kernel weights and
    matvec values) 1e-5 on the fields, its precision target against the FP64 reference, and 1e-9 on
    positions."""
    if name == "fp64":
        return 1e-10
    return 1e-9 if key in ("x", "y") else 1e-5


def _compare(case, ref, spreads, runs, npfb):
    """Prints each field's error against the reference beside the spread, then asserts the bar: ten
    times the spread, or the run's floor where that is larger."""
    fails = []
    for key in FIELDS:
        spread = max(_rel(s[key][:npfb], ref[key][:npfb]) for s in spreads)
        line = [f"{key:6s} spread {spread:.2e}"]
        for name, rows in runs.items():
            err = _rel(rows[key][:npfb], ref[key][:npfb])
            mark = "" if key not in ("u", "p", "theta", "tau") else (" (<=1e-5)" if err <= 1e-5 else " (>1e-5)")
            line.append(f"{name} {err:.2e}{mark}")
            if err > max(10.0 * spread, _floor(name, key)):
                fails.append((name, key, err, spread))
        print("  " + ", ".join(line), flush=True)
    assert not fails, fails


@pytest.mark.skipif(not have_ref(), reason="oracle/_ref (the reference build) is absent")
def test_cylinder_1M_against_reference(gpu_lib):
    from oracle import oraclewrap as ow
    case = cases.make("cylinder_1M")
    n = case.npfb
    print(f"case {case.name} npfb={n}, reference on {_threads()} threads", flush=True)
    nsteps = 1                      # the reference takes ~25 s per step at this size on 16 cores
    s0, sc0, ref, rsc = _reference_run(case, solver.initial_state(case), nsteps)
    sp = [_reference_run(case, _ulp_start(case, seed), nsteps)[2:] for seed in (1, 2)]
    spreads = [a for a, _ in sp]
    print(f"reference: LS_iters {rsc['LS_iters']}, residual {rsc['LS_residual']:.2e}, np {rsc['np']}; "
          f"one-ulp runs: residuals {', '.join('%.2e' % b['LS_residual'] for _, b in sp)}", flush=True)
    g64, sc64, f64 = _gpu_run(case, _lib.PREC_FP64, nsteps)
    gmx, scmx, fmx = _gpu_run(case, _lib.PREC_MIXED, nsteps)
    print(f"GPU: FP64 LS_iters {sc64['LS_iters']} residual {sc64['LS_residual']:.2e}; "
          f"mixed LS_iters {scmx['LS_iters']} residual {scmx['LS_residual']:.2e}", flush=True)
    assert sc64["np"] == rsc["np"] and scmx["np"] == rsc["np"]
    _compare(case, ref, spreads, {"fp64": g64, "mixed": gmx}, n)
    # the force on the cylinder of each run's rows, on the step's lists rebuilt by the oracle
    O = ow.OracleSolver(case, time=sc0["time"], itime=sc0["itime"], state=s0)
    for k in range(1, 10):
        O.phase(k)
    assert O.scalars()["np"] == rsc["np"]
    forces = {}
    for name, rows in (("reference", ref), ("spread1", spreads[0]), ("spread2", spreads[1]),
                       ("fp64", g64), ("mixed", gmx)):
        O.set_rows(rows, rsc["np"])
        forces[name] = O.circle_forces()[0][0]
    fref = forces["reference"]
    fspread = max(_rel(forces[s], fref) for s in ("spread1", "spread2"))
    e64, emx = _rel(forces["fp64"], fref), _rel(forces["mixed"], fref)
    print(f"force on the cylinder: reference {fref}, GPU FP64 {forces['fp64']}, mixed {forces['mixed']}; "
          f"relative error FP64 {e64:.2e}, mixed {emx:.2e}, reference spread {fspread:.2e}", flush=True)
    # the GPU's own force kernel on its own rows agrees with the oracle's evaluation
    assert _rel(f64, forces["fp64"]) <= 1e-10 and _rel(fmx, forces["mixed"]) <= 1e-10
    assert e64 <= max(10.0 * fspread, 1e-10) and emx <= max(10.0 * fspread, 1e-5)


@pytest.mark.skipif(not have_ref(), reason="oracle/_ref (the reference build) is absent")
def test_convection_80k_against_reference(gpu_lib):
    case = cases.make("convection_80k")
    n = case.npfb
    print(f"case {case.name} npfb={n}, reference on {_threads()} threads", flush=True)
    nsteps = 2
    _, _, ref, rsc = _reference_run(case, solver.initial_state(case), nsteps)
    spreads = [_reference_run(case, _ulp_start(case, seed), nsteps)[2] for seed in (1, 2)]
    print(f"reference: LS_iters {rsc['LS_iters']}, residual {rsc['LS_residual']:.2e}", flush=True)
    g64, sc64, _ = _gpu_run(case, _lib.PREC_FP64, nsteps)
    gmx, scmx, _ = _gpu_run(case, _lib.PREC_MIXED, nsteps)
    print(f"GPU: FP64 LS_iters {sc64['LS_iters']} residual {sc64['LS_residual']:.2e}; "
          f"mixed LS_iters {scmx['LS_iters']} residual {scmx['LS_residual']:.2e}", flush=True)
    assert sc64["np"] == rsc["np"] and scmx["np"] == rsc["np"]
    _compare(case, ref, spreads, {"fp64": g64, "mixed": gmx}, n)


def _oracle_run(args):
    """Worker: the oracle over nsteps from the case's start state, or from fluid positions moved by
    one ulp (seed); returns the final state and the scalars."""
    name, nsteps, seed = args
    from oracle import oraclewrap as ow
    case = cases.make(name)
    O = ow.OracleSolver(case, state=None if seed is None else _ulp_start(case, seed))
    for _ in range(nsteps):
        O.step()
    return O.get_state(), O.scalars()


def test_poiseuille_2500_whole_run_against_oracle(gpu_lib):
    """The Poiseuille configuration over its whole run, 1000 steps (2.6k particles, Oldroyd-B, Wi = 1,
    beta = 0.1, Re = 0.1): the GPU in FP64 and in mixed precision against the oracle, with the oracle's
    own one-ulp spread as the scale (both oracle runs in CPU worker processes, started before this
    process uses the GPU)."""
    import concurrent.futures as cf
    import multiprocessing as mp
    name, nsteps = "poiseuille_2500", 1000
    with cf.ProcessPoolExecutor(max_workers=2, mp_context=mp.get_context("spawn")) as ex:
        fut = [ex.submit(_oracle_run, (name, nsteps, seed)) for seed in (None, 1)]
        case = cases.make(name)
        runs, scs = {}, {}
        for prec, label in ((_lib.PREC_FP64, "fp64"), (_lib.PREC_MIXED, "mixed")):
            G = solver.Solver(case, precision=prec)
            worst = 0
            for _ in range(nsteps):
                G.div_free()
                st = G.scalars()["status"]
                assert st in (0, 6), st
                worst = max(worst, G.scalars()["LS_iters"])
            runs[label], scs[label] = G.get_state(), G.scalars()
            print(f"GPU {label}: time {scs[label]['time']:.6f}, n_inbin {scs[label]['n_inbin']}, "
                  f"most BiCGSTAB iterations {worst}", flush=True)
            G.close()
        (ref, rsc), (twin, _) = fut[0].result(), fut[1].result()
    print(f"oracle: time {rsc['time']:.6f}, LS_iters (last) {rsc['LS_iters']}", flush=True)
    for label in runs:
        assert scs[label]["n_inbin"] == rsc["n_inbin"] == 0
        assert abs(scs[label]["time"] - rsc["time"]) <= 1e-10 * rsc["time"]      # dt: CFL or viscous limit
    umax = np.max(np.abs(ref["u"][case.nbS:]))
    print(f"centreline-scale velocity after {nsteps} steps: {umax:.6f}", flush=True)
    _compare(case, ref, [twin], runs, case.npfb)
