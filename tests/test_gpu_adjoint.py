"""The adjoint of a run of steps (vb_adjoint_record / vb_adjoint_backward; the reference has no
adjoint, so the derivative is pinned by its definition): central finite differences of the GPU's own
forward steps.

* Directional derivatives of a random linear functional of the final state along random directions
  of the start state (positions, velocities, temperature, stress, trace) and of (Wi, beta, Ra), after
  one and three steps of a Poiseuille and a Boussinesq convection case from a developed state (random
  smooth velocities, stress with distinct eigenvalues, so the log-conformation map is smooth).
* dL/dWi, dL/dbeta, dL/dRa of the velocity-mismatch loss of the adjoint Poiseuille inversion, reduced
  to a 100 x 25 lattice and 20 steps with checkpoints every 5, against central differences with
  eps = 1e-6 (relative), through the torch.autograd.Function wrapper.
* The adjoint convection inversion, reduced, with a mixed-precision forward and a Nusselt-number term
  in the loss (formed in torch, so Ra's explicit part goes through autograd).
* The checkpoint interval changes nothing: the segments are recomputed bit for bit.
* Central differences of a run that starts from the isotropic conformation A = I do not resolve a
  derivative (see prestretched), so the inversions start from rest with a pre-stretched conformation.

The forward solves run to a relative residual of 1e-12 (1e-13 is at the floor FP64 BiCGSTAB
reaches on these systems: one of the runs stalled just above it), so that the finite differences
(which see the solver's stopping error divided by 2 eps) resolve the derivative; the adjoint
differentiates the exact solve.  Lists and mirror sets stay fixed under these perturbations (the pair test's kernel
vanishes with its first two derivatives at the cutoff)."""
import math

import numpy as np
import pytest

from vebisph_amd import _lib, adjoint, cases, solver

pytestmark = pytest.mark.gpu

TOL = 1e-12
ITMAX = 3000
FIELDS = solver.Solver.ADJ_FIELDS


def _state_arrays(st):
    out = {"x": st["x"], "y": st["y"], "u": st["u"], "v": st["v"], "theta": st["theta"],
           "tau11": st["tau"][:, 0], "tau12": st["tau"][:, 1], "tau21": st["tau"][:, 2],
           "tau22": st["tau"][:, 3], "trace": st["trace"]}
    return {k: np.array(v, dtype=np.float64) for k, v in out.items()}


def _to_state(a, case):
    return dict(x=a["x"], y=a["y"], u=a["u"], v=a["v"], theta=a["theta"],
                tau=np.stack([a["tau11"], a["tau12"], a["tau21"], a["tau22"]], axis=1), trace=a["trace"],
                p=np.zeros(case.npfb), inbin=np.zeros(case.npfb, np.int32))


def _developed(case, seed):
    """The case's start state with smooth random fluid velocities and a symmetric stress whose
    eigenvalues are apart (walls keep their own velocity and temperature)."""
    rng = np.random.default_rng(seed)
    a = _state_arrays(solver.initial_state(case))
    nb, n = case.nbS, case.npfb
    k1, k2 = rng.uniform(1, 3, 2)
    L = case.b_node[:, 0].max() - case.b_node[:, 0].min()
    H = case.b_node[:, 1].max() - case.b_node[:, 1].min()
    xs, ys = a["x"][nb:], a["y"][nb:]
    a["u"][nb:] = 0.05 * np.sin(2 * np.pi * k1 * xs / L) * np.sin(np.pi * ys / H)
    a["v"][nb:] = 0.03 * np.cos(2 * np.pi * k2 * xs / L) * np.sin(np.pi * ys / H)
    t12 = 0.05 + 0.02 * rng.standard_normal(n)
    a["tau11"] = 0.3 + 0.05 * rng.standard_normal(n)
    a["tau22"] = 0.05 + 0.02 * rng.standard_normal(n)
    a["tau12"] = t12; a["tau21"] = t12.copy()
    a["trace"] = 2.0 + a["tau11"] + a["tau22"]
    return a


def _final(case, params, state_arrays, nsteps):
    st = adjoint.run_forward(case, *params, nsteps, tol=TOL, itermax=ITMAX, state=_to_state(state_arrays, case))
    return _state_arrays(st)


def _functional(case, seed):
    """Random weights of a linear functional of the final state (fluid rows; walls do not move)."""
    rng = np.random.default_rng(seed)
    n = case.npfb
    w = {k: rng.standard_normal(n) for k in FIELDS}
    for k in ("x", "y", "u", "v", "theta"):
        w[k][:case.nbS] = 0.0
    return w


def _apply(w, a):
    return sum(float(np.dot(w[k], a[k])) for k in FIELDS)


CASES = {
    "poiseuille": lambda: cases.make("poiseuille_2500"),
    "convection": lambda: cases.convection(nx=100, ny=50, name="convection_5k"),
}


@pytest.mark.parametrize("name,nsteps", [("poiseuille", 1), ("poiseuille", 3), ("convection", 1),
                                         ("convection", 3)])
def test_adjoint_directional_derivatives(gpu_lib, name, nsteps):
    case = CASES[name]()
    params = (case.Wi, case.beta, case.Ra)
    s0 = _developed(case, 7)
    w = _functional(case, 8)
    print(f"case {case.name} npfb={case.npfb} steps {nsteps}", flush=True)
    G = solver.Solver(adjoint.case_with(case, *params), state=_to_state(s0, case))
    it, st = G.adjoint_record(nsteps, 2, TOL, ITMAX)
    res = G.adjoint_backward(w, TOL, ITMAX, initial=True)
    print(f"forward PPE iterations {list(it)}, adjoint {list(res['adj_iters'])}", flush=True)
    assert st == 0 and res["status"] == 0
    rng = np.random.default_rng(9)
    n, nb = case.npfb, case.nbS
    # directions: positions of the fluid rows (scale dx), fluid velocities, temperature, stress, trace
    dirs = {}
    d = {k: np.zeros(n) for k in FIELDS}; d["x"][nb:] = rng.standard_normal(n - nb); d["y"][nb:] = rng.standard_normal(n - nb)
    dirs["positions"] = (d, 1e-6 * case.dx)
    d = {k: np.zeros(n) for k in FIELDS}; d["u"][nb:] = rng.standard_normal(n - nb); d["v"][nb:] = rng.standard_normal(n - nb)
    dirs["velocities"] = (d, 1e-6)
    if name == "convection":
        d = {k: np.zeros(n) for k in FIELDS}; d["theta"][nb:] = rng.standard_normal(n - nb)
        dirs["temperature"] = (d, 1e-6)
    d = {k: np.zeros(n) for k in FIELDS}
    t = rng.standard_normal(n); d["tau11"] = rng.standard_normal(n); d["tau22"] = rng.standard_normal(n)
    d["tau12"] = t; d["tau21"] = t.copy(); d["trace"] = d["tau11"] + d["tau22"]
    dirs["stress"] = (d, 1e-6)
    for key, (d, eps) in dirs.items():
        plus = {k: s0[k] + eps * d[k] for k in FIELDS}
        minus = {k: s0[k] - eps * d[k] for k in FIELDS}
        fd = (_apply(w, _final(case, params, plus, nsteps)) - _apply(w, _final(case, params, minus, nsteps))) / (2 * eps)
        adj = sum(float(np.dot(res["initial"][k], d[k])) for k in FIELDS)
        err = abs(adj - fd) / max(abs(fd), 1e-300)
        print(f"  {key}: adjoint {adj:.10e}  central FD {fd:.10e}  rel {err:.2e}", flush=True)
        assert err <= 1e-5, key
    # parameters (step 1e-4 relative: the smallest of these derivatives, dL/dWi ~ 5e-4 against a
    # functional of ~1e2, is within the solves' stopping error / 2 eps at smaller steps; the
    # truncation error at 1e-4 is ~1e-8 relative)
    for q, pname in enumerate(("Wi", "beta", "Ra")):
        h = 1e-4 * params[q]
        pp = list(params); pm = list(params)
        pp[q] += h; pm[q] -= h
        fd = (_apply(w, _final(case, pp, s0, nsteps)) - _apply(w, _final(case, pm, s0, nsteps))) / (2 * h)
        adj = res["grads"][q]
        err = abs(adj - fd) / max(abs(fd), 1e-300)
        print(f"  d/d{pname}: adjoint {adj:.10e}  central FD {fd:.10e}  rel {err:.2e}", flush=True)
        assert err <= 1e-4, pname


def _reduced_c4():
    """The adjoint Poiseuille inversion's geometry ([0,4] x [0,1], walls at y = 0, 1, periodic in x,
    body force (8, 0), Re = 0.1), on a 100 x 25 lattice."""
    return cases.poiseuille(nx=100, ny=25, L=4.0, H=1.0, seed=4, Wi=1.0, beta=0.1, Re=0.1, steps=20,
                            name="poiseuille_c4_2500")


def prestretched(case, amp=1e-2):
    """The inversions' start state (vebisph_amd.adjoint.prestretched_state): measured on the reduced
    Poiseuille case from A = I instead, central differences of dL/dWi change with their step (6.9e-8,
    -1.0e-4, 1.7e-4, 5.4e-5 at eps = 1e-4 .. 1e-7) -- no derivative to compare with."""
    return adjoint.prestretched_state(case, amp)


def test_adjoint_gradients_reduced_poiseuille_inversion(gpu_lib):
    """dL/dWi, dL/dbeta, dL/dRa of L = 1/2 |u - u_target|^2 (fluid velocities) after 20 steps from rest
    (pre-stretched conformation, see prestretched) at Wi = 1, beta = 0.1, the target made by the solver
    at Wi = 1.5, beta = 0.2; checkpoints every 5 steps; the gradient through torch.autograd; central
    differences with eps = 1e-6 (relative)."""
    import torch
    case = _reduced_c4()
    nsteps, every = 20, 5
    nb = case.nbS
    start = prestretched(case)
    tgt = adjoint.run_forward(case, 1.5, 0.2, case.Ra, nsteps, tol=TOL, itermax=ITMAX, state=start)
    ut, vt = tgt["u"][nb:], tgt["v"][nb:]

    def loss_np(p):
        st = adjoint.run_forward(case, *p, nsteps, tol=TOL, itermax=ITMAX, state=start)
        return 0.5 * (np.sum((st["u"][nb:] - ut) ** 2) + np.sum((st["v"][nb:] - vt) ** 2))

    Run = adjoint.make_run_function()
    p0 = (1.0, 0.1, case.Ra)
    params = torch.tensor(p0, dtype=torch.float64, requires_grad=True)
    final = Run.apply(params, case, nsteps, every, TOL, ITMAX, _lib.PREC_FP64, start)
    loss = 0.5 * (((final[2, nb:] - torch.from_numpy(ut)) ** 2).sum() + ((final[3, nb:] - torch.from_numpy(vt)) ** 2).sum())
    loss.backward()
    grads = params.grad.numpy()
    print(f"loss {loss.item():.10e} (forward loss {loss_np(p0):.10e})", flush=True)
    for q, name in enumerate(("Wi", "beta", "Ra")):
        h = 1e-6 * p0[q]
        pp = list(p0); pm = list(p0)
        pp[q] += h; pm[q] -= h
        fd = (loss_np(pp) - loss_np(pm)) / (2 * h)
        err = abs(grads[q] - fd) / max(abs(fd), 1e-300)
        print(f"dL/d{name}: adjoint {grads[q]:.10e}  central FD {fd:.10e}  rel {err:.2e}", flush=True)
        assert err <= 1e-4, name
    # the run's adjoint PPE solves met their tolerance (the backward records their iteration counts)
    assert max(final.grad_fn.adj_iters) < ITMAX


def test_adjoint_reduced_convection_with_nusselt(gpu_lib):
    """The adjoint convection inversion, reduced: a 2 x 1 box (80 x 40 lattice) at Ra = 5e4, Wi = 1,
    beta = 0.5 with seeded temperature noise, pre-stretched conformation, 10 steps with checkpoints
    every 5; L = 1/2 |v(mid-plane) - v_target|^2 + (Nu - Nu_target)^2 with the target from the solver at
    Ra = 6e4.  Nu = 1 + sqrt(Ra Pr) <v theta> H / dT is formed in torch from the run's final v, theta,
    so its explicit dependence on Ra goes through autograd and the rest through the C ABI's adjoint.
    Checked against central differences (eps = 1e-6 relative) of the FP64 run, solved to 1e-11; then
    the same gradient through the mixed-precision run (FP32 kernel weights and matvec values) against
    that FP64 adjoint.  (Central differences of the mixed run itself do not resolve a derivative at
    this step: its weights are rounded to FP32, a 6e-8 quantisation of the forward map.)"""
    import torch
    case = cases.convection(nx=80, ny=40, seed=5, Ra=5e4, Wi=1.0, beta=0.5, steps=10, name="convection_c5_3k")
    nsteps, every, tol = 10, 5, 1e-11
    nb = case.nbS
    start = prestretched(case)
    y = case.y[nb:]
    mid = np.abs(y - 0.5) < 0.5 * case.dx + 1e-12        # the particles of the mid-plane row
    H, dT = 1.0, 2.0

    def nusselt_np(st, ra):
        return 1.0 + math.sqrt(ra * case.Pr) * np.mean(st["v"][nb:] * st["theta"][nb:]) * H / dT

    tgt = adjoint.run_forward(case, 1.0, 0.5, 6e4, nsteps, tol=tol, itermax=ITMAX, state=start)
    vt, nut = tgt["v"][nb:][mid], nusselt_np(tgt, 6e4)

    def loss_np(p):
        st = adjoint.run_forward(case, *p, nsteps, tol=tol, itermax=ITMAX, state=start)
        return 0.5 * np.sum((st["v"][nb:][mid] - vt) ** 2) + (nusselt_np(st, p[2]) - nut) ** 2

    Run = adjoint.make_run_function()
    p0 = (1.0, 0.5, 5e4)

    def gradient(prec):
        params = torch.tensor(p0, dtype=torch.float64, requires_grad=True)
        final = Run.apply(params, case, nsteps, every, tol, ITMAX, prec, start)
        v, th = final[3, nb:], final[4, nb:]
        nu = 1.0 + torch.sqrt(params[2] * case.Pr) * torch.mean(v * th) * H / dT
        loss = 0.5 * ((v[torch.from_numpy(mid)] - torch.from_numpy(vt)) ** 2).sum() + (nu - nut) ** 2
        loss.backward()
        return loss.item(), nu.item(), params.grad.numpy().copy()

    loss, nu, grads = gradient(_lib.PREC_FP64)
    print(f"loss {loss:.10e} (forward {loss_np(p0):.10e}), Nu {nu:.6f} target {nut:.6f}", flush=True)
    for q, name in enumerate(("Wi", "beta", "Ra")):
        h = 1e-6 * p0[q]
        pp = list(p0); pm = list(p0)
        pp[q] += h; pm[q] -= h
        fd = (loss_np(pp) - loss_np(pm)) / (2 * h)
        err = abs(grads[q] - fd) / max(abs(fd), 1e-300)
        print(f"dL/d{name}: adjoint {grads[q]:.10e}  central FD {fd:.10e}  rel {err:.2e}", flush=True)
        assert err <= 1e-4, name
    _, _, gmix = gradient(_lib.PREC_MIXED)
    for q, name in enumerate(("Wi", "beta", "Ra")):
        err = abs(gmix[q] - grads[q]) / abs(grads[q])
        print(f"dL/d{name} through the mixed run: {gmix[q]:.10e}, against the FP64 adjoint {err:.2e}", flush=True)
        assert err <= 1e-4, name


def test_adjoint_checkpoint_interval_changes_nothing(gpu_lib):
    case = CASES["poiseuille"]()
    s0 = _to_state(_developed(case, 3), case)
    w = _functional(case, 4)
    out = []
    for every in (1, 3, 6):
        G = solver.Solver(case, state=s0)
        G.adjoint_record(6, every, 1e-10, ITMAX)
        r = G.adjoint_backward(w, 1e-10, ITMAX, initial=True)
        out.append((r["grads"].copy(), {k: v.copy() for k, v in r["initial"].items()}))
        # the sweep leaves the context at the recorded final state, and runs again the same way
        r2 = G.adjoint_backward(w, 1e-10, ITMAX)
        assert np.array_equal(r2["grads"], r["grads"])
    for g, ini in out[1:]:
        assert np.array_equal(g, out[0][0])
        for k in FIELDS:
            assert np.array_equal(ini[k], out[0][1][k]), k


def test_adjoint_refuses_what_it_does_not_differentiate(gpu_lib):
    for case in (cases.cylinder(nx=160, ny=80, name="cylinder_13k"), cases.dam_break()):
        G = solver.Solver(case)
        with pytest.raises(solver.VbError) as e:
            G.adjoint_record(2, 1)
        assert e.value.status == 8
