"""Gradients of a run of VEBISPH steps with respect to the rheology and flow parameters, as a
torch.autograd.Function over the C ABI (vb_adjoint_record forward, vb_adjoint_backward backward).

    params = torch.tensor([Wi, beta, Ra], dtype=torch.float64, requires_grad=True)
    final = Run.apply(params, case, nsteps, ckpt_every)        # (5, npfb): x, y, u, v, theta
    loss = 0.5 * ((final[2] - u_target) ** 2).sum()
    loss.backward()                                            # params.grad = dL/d(Wi, beta, Ra)

The parameters are independent: El = Wi sqrt(Pr/Ra) with Pr fixed (so Ra moves the viscous, thermal
and polymeric coefficients and the viscous time-step limit, div_free.F90:142).  Tensors are FP64 on
the CPU: the library keeps its own device state (torch carries a second HIP runtime, so no device
tensor is shared).  There is no CPU fallback: without the HIP library the forward raises.
"""
from __future__ import annotations

import copy
import math

import numpy as np

from . import _lib, solver

FINAL_FIELDS = ("x", "y", "u", "v", "theta")


def case_with(case, wi, beta, ra):
    """A copy of `case` with Wi, beta and Ra set (El = Wi sqrt(Pr / Ra))."""
    c = copy.copy(case)
    c.Ra = float(ra)
    c.beta = float(beta)
    c.El = float(wi) * math.sqrt(case.Pr / float(ra))
    return c


def prestretched_state(case, amp=1e-2):
    """The start state of the inversions: the fluid at rest with a smooth, non-isotropic conformation
    (tau11 - tau22 >= amp / 2, tau12 = tau21 of amplitude amp).  From the isotropic A = I the reference's
    log-conformation map divides its rotation rate by the 1e-10 regularisation
    (predictor_mod.F90:203-207), so the ulp-level rounding of A11 - A22 -- which moves with Wi through
    exp(-dt/Wi) -- reaches the shear stress amplified ~1e9 times, and the run has no derivative in Wi
    that a finite difference resolves; a pre-stretched start keeps the eigenvalues apart."""
    st = solver.initial_state(case)
    L = float(case.b_node[:, 0].max() - case.b_node[:, 0].min())
    H = float(case.b_node[:, 1].max() - case.b_node[:, 1].min())
    x, y = np.asarray(case.x), np.asarray(case.y)
    t11 = 2.0 * amp + amp * np.sin(2 * np.pi * x / L) * np.cos(np.pi * y / H)
    t22 = 0.5 * amp + 0.2 * amp * np.cos(2 * np.pi * x / L)
    t12 = amp * np.cos(2 * np.pi * x / L) * np.sin(np.pi * y / H)
    st["tau"] = np.stack([t11, t12, t12.copy(), t22], axis=1)
    st["trace"] = 2.0 + t11 + t22
    return st


def run_forward(case, wi, beta, ra, nsteps, tol=1e-8, itermax=1000, precision=_lib.PREC_FP64, state=None):
    """The final state (dict) of nsteps plain steps at the given parameters (no record)."""
    s = solver.Solver(case_with(case, wi, beta, ra), precision=precision, state=state)
    for _ in range(nsteps):
        s.set_time_step(); s.put_in_bin(); s.create_mirror_particles(); s.particle_allocation()
        s.kernel_calculation(); s.locate_free_surf_part(); s.new_fick_shift(); s.prediction_step()
        s.mirror_velocities(); s.ppe_solve_in_house_simple_dirichlet(tol, itermax)
        s.calc_pressure_gradient(); s.final_advection()
    out = s.get_state()
    s.close()
    return out


def _torch():
    import torch
    return torch


def make_run_function():
    torch = _torch()

    class Run(torch.autograd.Function):
        """forward(params = [Wi, beta, Ra], case, nsteps, ckpt_every, tol, itermax, precision, state)
        -> (5, npfb) final x, y, u, v, theta; backward: dL/dparams through vb_adjoint_backward."""

        @staticmethod
        def forward(ctx, params, case, nsteps, ckpt_every, tol=1e-8, itermax=1000, precision=_lib.PREC_FP64,
                    state=None):
            wi, beta, ra = (float(v) for v in params.detach().cpu().numpy())
            s = solver.Solver(case_with(case, wi, beta, ra), precision=precision, state=state)
            it, _ = s.adjoint_record(nsteps, ckpt_every, tol, itermax)
            st = s.get_state()
            ctx.solver, ctx.tol, ctx.itermax = s, tol, itermax
            ctx.ppe_iters = it
            return torch.from_numpy(np.stack([st[k] for k in FINAL_FIELDS]))

        @staticmethod
        def backward(ctx, gout):
            g = gout.detach().cpu().numpy()
            res = ctx.solver.adjoint_backward({k: g[q] for q, k in enumerate(FINAL_FIELDS)}, ctx.tol, ctx.itermax)
            ctx.adj_iters = res["adj_iters"]
            return (torch.tensor(res["grads"], dtype=torch.float64), None, None, None, None, None, None, None)

    return Run
