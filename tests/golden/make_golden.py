"""Writes the golden fixtures of the oracle: the reference itself (oracle/_ref/libvbref.so, built from
the reference's Fortran sources by oracle/Makefile, one OpenMP thread) stepped phase by phase on
small seeded cases.

    python tests/golden/make_golden.py [nnf]      (nnf: only the viscosity-model fixture)

For every phase it records the SHA-256 of each array the phase produces (mirror relations and
offsets, neighbour lists, kernel weights and KGC, free-surface flags and quantities, shifting
velocity, particle state, pressure gradient) and the step scalars; the final state is stored whole.  tests/test_golden.py checks the C
restatement against these files, so the oracle stays pinned where the reference is absent.  It also
runs the reference's case generator (oracle/_ref/gen2D) on its eight cases and records each run
directory as the readers parse it (generator_cases.json, for tests/test_generator.py), and the
reference's nnf_models subroutines on a shear-rate sweep (nnf_reference.npz).
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from vebisph_amd import cases  # noqa: E402

def _generated(k, const_mod=2):
    """Generator case k (vebisph_amd.generator) as a run reads it, for the given build."""
    import tempfile
    from vebisph_amd import generator, refio
    d = tempfile.mkdtemp()
    generator.write_run_dir(k, d)
    c = refio.read_case_dir(d).case
    c.const_mod = const_mod
    return c


CASES = {
    "poiseuille_2500": (lambda: cases.make("poiseuille_2500"), 2),
    "convection_1250": (lambda: cases.convection(nx=50, ny=25, name="convection_1250"), 2),
    # a coarser cylinder (nx=80) puts more than nplink=75 neighbours on particles next to the circle,
    # which the reference writes past the end of ij_link (heap corruption); the GPU path returns
    # VB_ERR_CAPACITY there instead
    "cylinder_8k": (lambda: cases.cylinder(nx=128, ny=64, name="cylinder_8k"), 2),
    # open domains: free surface, ramp rows (jitter 0.2 dx), Jacobi BiCGSTAB, surface shifting
    "fs_convection_11k": (lambda: cases.free_surface_convection(jitter=0.2), 2),
    "dam_break_3k": (lambda: cases.dam_break(jitter=0.2, shift_coeff=1), 2),
    # the reference built as its Newtonian program (-Dconst_mod=1) on the generator's Newtonian cases
    "gen1_dam_break_newtonian": (lambda: _generated(1, const_mod=1), 2),
    "gen2_taylor_green_newtonian": (lambda: _generated(2, const_mod=1), 2),
}
FINAL_STATE = ("poiseuille_2500", "convection_1250")     # whole final state kept (small files)


def digest(a):
    """SHA-256 of the values; -0.0 is hashed as +0.0 (the restatement and the reference may differ
    in the sign of an exact zero, e.g. tau = A - I on untouched particles, which no later operation
    of the step can observe)."""
    a = np.asarray(a)
    a = a.astype(np.int32) if a.dtype.kind in "iub" else a.astype(np.float64) + 0.0
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def phase_record(S, k):
    """Digests of what phase k (1..12 of div_free) produces; S is a RefSolver or OracleSolver."""
    sc = S.scalars()
    rec = {}
    if k == 1:
        rec["dt"] = sc["dt"]
    if k == 3:
        rec["np"] = sc["np"]
        rec["mirrors"] = [digest(a) for a in S.mirrors()]
    if k == 4:
        rec["neighbours"] = [digest(a) for a in S.neighbours()]
    if k == 5:
        w = sc["maxneighbours"]
        rec["maxneighbours"] = w
        rec["weights"] = [digest(a) for a in S.weights(w)]
    if k == 6:
        rec["surface"] = {key: digest(v) for key, v in S.surface().items()}
    if k == 7:
        rec["ushift"] = digest(S.ushift())
    if k in (8, 9, 10, 11):
        st = S.get_state(sc["np"])
        rec["state"] = {key: digest(st[key]) for key in ("x", "y", "u", "v", "theta", "p", "trace", "tau")}
    if k == 10:
        rec["LS_iters"] = sc["LS_iters"]
        rec["LS_residual"] = sc["LS_residual"]
    if k == 11:
        rec["grad_p"] = digest(S.grad_p())
    if k == 12:
        rec["time"] = sc["time"]
    return rec


GEN2D = os.path.join(ROOT, "oracle", "_ref", "gen2D")


def run_dir_record(path):
    """Digests of a run directory as the readers parse it (getdata's view), and INDAT's text."""
    from vebisph_amd import refio
    rd = refio.read_case_dir(path)
    c = rd.case
    rec = {"INDAT": open(os.path.join(path, "INDAT")).read(), "npfb": c.npfb, "nbS": c.nbS,
           "scalars": [rd.tmax, rd.dt_out, rd.dt, rd.alpha_evss, c.h, c.dx, c.dv, c.Ra, c.Pr, c.El, c.beta,
                       c.body_force[0], c.body_force[1], c.dt_coef_cfl, c.dt_coef_visc, c.ve_nonlinearity,
                       c.i_open_domain, c.i_shift_coeff]}
    for key in ("x", "y", "u", "v", "theta", "b_node", "b_type", "b_vel", "b_thermal", "b_periodic_parent",
                "c_centre", "c_radius", "c_omega", "c_vel", "c_thermal"):
        a = getattr(c, key)
        rec[key] = None if a is None else digest(a)
    return rec


def generator_fixtures():
    """The reference's own generator (oracle/_ref/gen2D, datclass.F90) on each of its eight cases."""
    import subprocess
    import tempfile
    out = {}
    for k in range(1, 9):
        d = tempfile.mkdtemp()
        subprocess.run([GEN2D], input=f"{k}\n", text=True, cwd=d, capture_output=True, check=True)
        out[str(k)] = run_dir_record(d)
        print("generator case", k, out[str(k)]["npfb"])
    with open(os.path.join(HERE, "generator_cases.json"), "w") as f:
        json.dump(out, f, indent=0)


NNF_MU = (1.7, 0.02)          # viscosities of the nnf_models sweep


def nnf_sweep(mu):
    """Shear rates that exercise every term of nnf_models_mod.F90:8-110: 0, the Bingham threshold
    tau_y/(mu alpha) and the doubles either side of it, a log sweep over 1e-8..1e8 (the Cross
    transition near 1/kk = 1e3, the Carreau one near 1/kk = 7) and a linear sweep over 0..20."""
    thr = 1.0e-2 / (mu * 1.0e2)
    return np.concatenate([[0.0, thr, np.nextafter(thr, 0.0), np.nextafter(thr, np.inf),
                            np.nextafter(np.nextafter(thr, 0.0), 0.0)],
                           np.logspace(-8.0, 8.0, 161), np.linspace(0.0, 20.0, 81)[1:]])


def nnf_fixtures():
    """The reference's five nnf_models subroutines on the sweep, for each viscosity (nnf_reference.npz)."""
    from oracle import refwrap
    out = {}
    for q, mu in enumerate(NNF_MU):
        sr = nnf_sweep(mu)
        out[f"sr{q}"] = sr
        for model in range(5):
            out[f"mu{q}_model{model}"] = refwrap.nnf(model, sr, mu)
    out["mu"] = np.array(NNF_MU)
    np.savez_compressed(os.path.join(HERE, "nnf_reference.npz"), **out)
    print("nnf fixtures", {k: v.shape for k, v in out.items()})


def main():
    from oracle import refwrap
    if not refwrap.available():
        sys.exit("oracle/_ref/libvbref.so is not built (make -C oracle ref)")
    nnf_fixtures()
    if sys.argv[1:] == ["nnf"]:
        return
    index = {}
    for name, (mk, steps) in CASES.items():
        case = mk()
        R = refwrap.RefSolver(case)
        R.set_threads(1)
        recs = []
        for _ in range(steps):
            step = []
            for k in range(1, 13):
                R.phase(k)
                step.append(phase_record(R, k))
            recs.append(step)
        final = R.get_state()
        if name in FINAL_STATE:
            np.savez_compressed(os.path.join(HERE, f"{name}_final.npz"),
                                **{k: np.asarray(v) for k, v in final.items()})
        index[name] = {"npfb": case.npfb, "steps": steps, "phases": recs,
                       "final": {k: digest(v) for k, v in final.items()}}
        print(name, case.npfb, "LS_iters", [r[9]["LS_iters"] for r in recs])
    with open(os.path.join(HERE, "reference_phases.json"), "w") as f:
        json.dump(index, f, indent=0)
    if not os.path.exists(GEN2D):
        sys.exit("oracle/_ref/gen2D is not built (make -C oracle gen)")
    generator_fixtures()


if __name__ == "__main__":
    main()
