"""vb_logconf_adjoint: the adjoint (vector-Jacobian product) of the per-particle log-conformation map
vb_logconf (predictor_mod.F90:165-245,253-290,318-331), one piece of an adjoint step.  The reference
has no adjoint, so the derivative is pinned by its definition:

* a torch.autograd.Function over the C-ABI (forward vb_logconf, backward vb_logconf_adjoint; FP64
  CPU tensors staged through the library's device buffers, since torch carries its own HIP runtime)
  passes torch.autograd.gradcheck: central finite differences of the GPU map itself, eps = 1e-6;
* the adjoint against central finite differences of the oracle's map (itself bit-exact against the
  reference), per input, to 1e-6 of each input's cotangent scale, with and without the reference's
  B perturbation ("kick") and with the PTT trace term (eps_ptt > 0).

Rows are generic (|A21| > 0.02 and an eigenvalue gap > 0.1) or diagonal with A11 > A22 + 0.1 (A21 = 0
exactly, where the map is smooth and the derivative pass takes the eigenvector branch), so the map is
smooth at every row.  The reference's non-smooth points (A21 = 0 with A11 <= A22, among them the
degenerate start A = I, where the 1e-10 in the rotation-rate denominator makes the map vary on a 1e-10
scale, and the discriminant clamp) are outside what a finite difference can check; there the adjoint
is the derivative of the branch taken, and the test checks that it is finite."""
import numpy as np
import pytest

from oracle import oraclewrap as ow
from tests.hipmem import DeviceArray as D

pytestmark = pytest.mark.gpu


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    tau = np.zeros((n, 4))
    tau[:, 0] = rng.uniform(-0.2, 0.6, n)
    tau[:, 3] = rng.uniform(-0.2, 0.6, n)
    tau[:, 1] = tau[:, 2] = rng.uniform(-0.3, 0.3, n)
    gap = np.sqrt((tau[:, 0] - tau[:, 3]) ** 2 + 4 * tau[:, 1] * tau[:, 2])
    # |A21| bounded away from 0: the eigenvector (L1 - A22, A21) of calc_eigens cancels as A21 -> 0
    # when A11 < A22, and a finite difference then divides that rounding by 2h
    keep = (gap > 0.1) & (np.abs(tau[:, 2]) > 0.02)
    tau = tau[keep]
    # and diagonal rows with A11 > A22 (tau_xy = tau_yx = 0, gap 0.1..0.5): the map is smooth across
    # A21 = 0 there, and the derivative in A21 is that of the eigenvector branch on either side
    nd = n // 8
    diag = np.zeros((nd, 4))
    diag[:, 3] = rng.uniform(-0.2, 0.3, nd)
    diag[:, 0] = diag[:, 3] + rng.uniform(0.1, 0.5, nd)
    tau = np.concatenate([tau, diag])
    m = tau.shape[0]
    trc = 2.0 + tau[:, 0] + tau[:, 3]
    gv = rng.uniform(-2, 2, (m, 4))
    return tau, trc, gv


def _adjoint(L, tau, trc, gv, dt, wi, eps, kick, tb, trb):
    n = tau.shape[0]
    args = [D(tau), D(trc), D(gv)]
    bars = [D(tb), D(trb)]
    out = [D(np.zeros((n, 4))), D(np.zeros(n)), D(np.zeros((n, 4))), D(np.zeros(n)), D(np.zeros(n))]
    st = L.vb_logconf_adjoint(n, *[a.ptr for a in args], dt, wi, eps, kick, *[b.ptr for b in bars],
                              *[o.ptr for o in out])
    assert st == 0
    return [o.numpy() for o in out]


@pytest.mark.parametrize("kick,eps", [(0, 0.0), (1, 0.0), (0, 0.25)])
def test_logconf_adjoint_against_oracle_fd(gpu_lib, kick, eps):
    L = gpu_lib
    tau, trc, gv = _inputs(4096, 21)
    n = tau.shape[0]
    dt, wi = 1e-3, 0.8
    rng = np.random.default_rng(4)
    tb, trb = rng.standard_normal((n, 4)), rng.standard_normal(n)
    tau_b, trc_b, gv_b, wi_b, dt_b = _adjoint(L, tau, trc, gv, dt, wi, eps, kick, tb, trb)

    def vjp_fd(f_plus, f_minus, h):
        (tp, rp), (tm, rm) = f_plus, f_minus
        return (np.sum(tb * (tp - tm), axis=1) + trb * (rp - rm)) / (2 * h)

    f = lambda ta, tr, g, d, w: ow.logconf(ta, tr, g, d, w, eps, kick=bool(kick))
    got = {}
    # steps sized to each input: grad_vel, Wi and the trace act through dt = 1e-3 (small cotangents,
    # an almost linear dependence), so a 1e-6 step would leave the map's few-ulp rounding / 2h at
    # ~1e-6 of their scale; with these steps truncation and rounding both stay below 1e-8 of it
    h, hg, hd = 1e-6, 1e-4, 1e-7
    for q in range(4):
        e = np.zeros(4); e[q] = h
        eg = np.zeros(4); eg[q] = hg
        got[f"tau{q}"] = (tau_b[:, q], vjp_fd(f(tau + e, trc, gv, dt, wi), f(tau - e, trc, gv, dt, wi), h))
        got[f"gv{q}"] = (gv_b[:, q], vjp_fd(f(tau, trc, gv + eg, dt, wi), f(tau, trc, gv - eg, dt, wi), hg))
    got["trace"] = (trc_b, vjp_fd(f(tau, trc + hg, gv, dt, wi), f(tau, trc - hg, gv, dt, wi), hg))
    got["wi"] = (wi_b, vjp_fd(f(tau, trc, gv, dt, wi + hg), f(tau, trc, gv, dt, wi - hg), hg))
    got["dt"] = (dt_b, vjp_fd(f(tau, trc, gv, dt + hd, wi), f(tau, trc, gv, dt - hd, wi), hd))
    for key, (adj, fd) in got.items():
        scale = max(np.abs(fd).max(), 1e-300)
        err = np.abs(adj - fd).max() / scale
        print(f"kick {kick} eps {eps} {key}: |adjoint - central FD| / scale = {err:.2e} (scale {scale:.3e})",
              flush=True)
        if key == "trace" and eps == 0.0:
            assert not np.any(adj)                    # the trace enters only through eps_ptt
        else:
            assert err <= 1e-6, key


def test_logconf_adjoint_gradcheck_and_degenerate_rows(gpu_lib):
    import torch
    L = gpu_lib
    dt, eps = 1e-3, 0.0

    class LogConf(torch.autograd.Function):
        """vb_logconf forward, vb_logconf_adjoint backward (Wi a differentiable 0-d tensor)."""

        @staticmethod
        def forward(ctx, tau, trc, gv, wi):
            n = tau.shape[0]
            d_in = [D(t.detach().numpy()) for t in (tau, trc, gv)]
            to, tro = D(np.zeros((n, 4))), D(np.zeros(n))
            assert L.vb_logconf(n, *[a.ptr for a in d_in], dt, float(wi), eps, 0, to.ptr, tro.ptr) == 0
            ctx.save_for_backward(tau, trc, gv, wi)
            return torch.from_numpy(to.numpy()), torch.from_numpy(tro.numpy())

        @staticmethod
        def backward(ctx, tb, trb):
            tau, trc, gv, wi = ctx.saved_tensors
            g = [t.detach().numpy() for t in (tau, trc, gv)]
            out = _adjoint(L, *g, dt, float(wi), eps, 0, tb.detach().numpy(), trb.detach().numpy())
            return (torch.from_numpy(out[0]), torch.from_numpy(out[1]), torch.from_numpy(out[2]),
                    torch.tensor(out[3].sum(), dtype=torch.float64).reshape(wi.shape))

    tau, trc, gv = _inputs(24, 8)
    dev = lambda a: torch.tensor(a, dtype=torch.float64, requires_grad=True)
    args = (dev(tau), dev(trc), dev(gv), dev(np.array(0.8)))
    # each Jacobian entry to 1e-6 + 1e-5 |J| (entries are O(1); the finite difference's own error is
    # ~1e-7 here)
    assert torch.autograd.gradcheck(LogConf.apply, args, eps=1e-6, atol=1e-6, rtol=1e-5)

    # the reference's degenerate start, A = I (A21 = 0: the diagonal branch), and a near-isotropic
    # state on the clamp: finite cotangents
    n = 64
    tau0 = np.zeros((n, 4)); tau0[32:, 0] = tau0[32:, 3] = 0.1
    gv0 = np.random.default_rng(2).uniform(-2, 2, (n, 4))
    res = _adjoint(L, tau0, 2.0 + tau0[:, 0] + tau0[:, 3], gv0, dt, 0.8, eps, 0, np.ones((n, 4)),
                   np.ones(n))
    assert all(np.all(np.isfinite(r)) for r in res)
