"""vb_ppe_operator: the assembled PPE as an operator, y = A x and y = A^T x (the operator an adjoint
solve needs; the reference has no adjoint, so A^T is pinned by its definition).  A is the matrix the
GPU's own vb_ppe_solve assembled (exported with vb_export_ppe and itself pinned against the oracle by
tests/test_gpu_parity.py); the reference values are scipy's products with that matrix and its
transpose.  Tolerance: elementwise 1e-13 of sum_j |A_ij| |x_j| (the summation orders differ), and the
adjoint identity <A x, y> = <x, A^T y> to 1e-12 of its terms' scale.

Cases cover every source of entries: walls and periodic mirrors (Poiseuille), circle mirrors
(cylinder), the open domain's Dirichlet rows (1e30 diagonals) and ppe_smooth ramp rows, which make A
unsymmetric (free-surface convection), and wall corners (dam break)."""
import numpy as np
import pytest
import scipy.sparse as sp

from vebisph_amd import _lib, cases, solver

pytestmark = pytest.mark.gpu

CASES = {
    "poiseuille": lambda: cases.make("poiseuille_2500"),
    "cylinder": lambda: cases.cylinder(nx=160, ny=80, name="cylinder_13k"),
    "fs_convection": lambda: cases.free_surface_convection(jitter=0.2),
    "dam_break": lambda: cases.dam_break(jitter=0.2, shift_coeff=1),
}


def _matrix(G):
    w = G.scalars()["maxneighbours"]
    diag, col, val, rhs = G.ppe_ell(w)
    cnt = G.neighbours()[0]
    a, ija = solver.ell_to_nr(diag, col, val, cnt)
    n = diag.shape[0]
    lens = ija[1:n + 1] - ija[:n]
    rows = np.concatenate([np.arange(n), np.repeat(np.arange(n), lens)])
    cols = np.concatenate([np.arange(n), ija[ija[0] - 1:ija[n] - 1] - 1])
    vals = np.concatenate([a[:n], a[ija[0] - 1:ija[n] - 1]])
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))     # duplicate slots summed


def _close(got, want, scale):
    err = np.abs(got - want)
    bad = err > 1e-13 * scale + 1e-300
    assert not bad.any(), f"{bad.sum()} rows off, worst {np.max(err / np.maximum(scale, 1e-300)):.3e}"


@pytest.mark.parametrize("name", list(CASES))
def test_ppe_operator_and_transpose(gpu_lib, name):
    case = CASES[name]()
    print(f"case {case.name} npfb={case.npfb}", flush=True)
    G = solver.Solver(case)
    G.div_free()
    for k in range(1, 11):
        G.phase(k)
    A = _matrix(G)
    absA = abs(A)
    rng = np.random.default_rng(11)
    x = rng.standard_normal(case.npfb)
    y = rng.standard_normal(case.npfb)
    ax = G.ppe_operator(x)
    atx = G.ppe_operator(x, transpose=True)
    _close(ax, A @ x, absA @ np.abs(x))
    _close(atx, A.T @ x, absA.T @ np.abs(x))
    # the solvers' own rows: the adjoint solve's A^T row (compact A^T plane, map rows, mirror
    # buckets) and the forward FP64 solve's compact rows, one application each
    _close(G.ppe_operator(x, transpose=True, solver_rows=True), A.T @ x, absA.T @ np.abs(x))
    _close(G.ppe_operator(x, solver_rows=True), A @ x, absA @ np.abs(x))
    aty = G.ppe_operator(y, transpose=True)
    lhs, rhs = float(np.dot(A @ x, y)), float(np.dot(x, aty))
    scale = float(np.abs(y) @ (absA @ np.abs(x)))
    print(f"<Ax,y> {lhs:.17e} <x,A^T y> {rhs:.17e} (scale {scale:.3e}); A symmetric: "
          f"{abs(A - A.T).max() == 0}", flush=True)
    assert abs(lhs - rhs) <= 1e-12 * scale
    # the same bits from a second context built the same way (fixed summation order)
    G2 = solver.Solver(case)
    G2.div_free()
    for k in range(1, 11):
        G2.phase(k)
    assert np.array_equal(G2.ppe_operator(x, transpose=True), atx)
    # the next step's lists no longer match the stored matrix: refused
    for k in range(1, 5):
        G.phase(k)
    with pytest.raises(solver.VbError):
        G.ppe_operator(x, transpose=True)


def test_ppe_operator_refused_for_a_callers_system(gpu_lib):
    """after vb_sparse_pbicgstab_nonull the stored matrix is the caller's (no lists): refused"""
    case = CASES["poiseuille"]()
    G = solver.Solver(case)
    G.div_free()
    w = G.scalars()["maxneighbours"]
    diag, col, val, rhs = G.ppe_ell(w)
    a, ija = solver.ell_to_nr(diag, col, val, G.neighbours()[0])
    G.sparse_pbicgstab_nonull(a, ija, rhs, itermax=5)
    with pytest.raises(solver.VbError):
        G.ppe_operator(np.ones(case.npfb))


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("name", list(CASES))
def test_ppe_adjoint_solve(gpu_lib, name, prec):
    """A^T lambda = g with the reference's BiCGSTAB on the transposed PPE: the residual meets the
    reference criterion on A^T itself (checked with scipy), and in the open domains, where A is
    nonsingular, lambda is scipy's direct solution to the accuracy the criterion allows
    (cond(A) * 1e-8: 1e-3 relative).  Closed domains (Poiseuille; the cylinder, whose circle mirrors
    make A unsymmetric) take g with zero mean (A^T maps into the mean-zero vectors); a g with a
    constant added gives the same lambda, since the solve removes g's mean first.  On the mixed path
    the refinement's inner BiCGSTAB reads the compact rows' A^T values as FP32 (exact there, the
    kernel weights being FP32) and its residual stays FP64: the same bars hold."""
    import scipy.sparse.linalg as spla
    case = CASES[name]()
    G = solver.Solver(case, precision=_lib.PREC_MIXED if prec == "mixed" else _lib.PREC_FP64)
    G.div_free()
    for k in range(1, 11):
        G.phase(k)
    A = _matrix(G)
    g = np.random.default_rng(5).standard_normal(case.npfb)
    if not case.i_open_domain:
        g -= g.mean()
    lam, it, res, st = G.ppe_adjoint_solve(g)
    print(f"{case.name}: adjoint BiCGSTAB {it} iterations, residual {res:.3e}, status {st}", flush=True)
    assert st == 0 and res <= 1e-8 and 1 <= it <= 1000
    r = g - A.T @ lam
    print(f"  scipy's residual {np.linalg.norm(r) / np.linalg.norm(g):.3e}", flush=True)
    assert np.linalg.norm(r) <= 1e-7 * np.linalg.norm(g)
    if case.i_open_domain:
        ref = spla.spsolve(A.T.tocsc(), g)
        err = np.linalg.norm(lam - ref) / np.linalg.norm(ref)
        print(f"  against scipy's direct solve: {err:.2e}", flush=True)
        assert err <= 1e-3
    else:
        lam2, it2, res2, st2 = G.ppe_adjoint_solve(g + 0.37)
        print(f"  g + 0.37: {it2} iterations, residual {res2:.3e}, status {st2}", flush=True)
        assert st2 == 0 and res2 <= 1e-8
        # the reported residual is that of the mean-free system, not the constant's rounding: the
        # refinement's norm of r - mean(r) is taken in a second pass (sum(r^2) - n mean^2 cancels
        # here, where the constant dominates)
        r2 = g - A.T @ lam2
        print(f"  scipy's residual {np.linalg.norm(r2) / np.linalg.norm(g):.3e}", flush=True)
        assert res2 > 0.0
        assert np.linalg.norm(r2) <= 1e-7 * np.linalg.norm(g)


def _grad_matrix(G, n):
    """The gradient operator as scipy matrices (Gx, Gy) from the exported lists, kernel gradients, KGC
    matrices and mirror parents (calc_pressure_gradient, gradient_calculation.F90:16-48)."""
    w = G.scalars()["maxneighbours"]
    cnt, link = G.neighbours(w)
    wg, _, kg = G.weights(w)
    irel, _, dp = G.mirrors()
    rows = np.repeat(np.arange(n), cnt)
    mask = np.arange(w)[None, :] < cnt[:, None]
    j = link[mask] - 1                                               # the exports are 1-based
    par = np.where(j < n, j, irel[np.maximum(j - n, 0)] - 1)
    wx, wy = wg[mask][:, 0], wg[mask][:, 1]
    S = lambda v: (sp.csr_matrix((v, (rows, par)), shape=(n, n))
                   - sp.csr_matrix((v, (rows, rows)), shape=(n, n)))
    Sx, Sy = S(wx), S(wy)
    K = [sp.diags(kg[:, q]) for q in range(4)]
    # the magnitudes of the terms before they are summed (a wall row's mirror and real terms cancel,
    # so |G| itself is no scale for the rounding of the sums)
    T = lambda v: (sp.csr_matrix((np.abs(v), (rows, par)), shape=(n, n))
                   + sp.csr_matrix((np.abs(v), (rows, rows)), shape=(n, n)))
    aK = [abs(k) for k in K]
    mag = (aK[0] @ T(wx) + aK[1] @ T(wy), aK[2] @ T(wx) + aK[3] @ T(wy))
    off = np.where(j < n, 0.0, dp[np.maximum(j - n, 0)])          # the pressure mirrors' offsets
    ox = np.bincount(rows, off * wx, n); oy = np.bincount(rows, off * wy, n)
    offset = np.stack([kg[:, 0] * ox + kg[:, 1] * oy, kg[:, 2] * ox + kg[:, 3] * oy], axis=1)
    return K[0] @ Sx + K[1] @ Sy, K[2] @ Sx + K[3] @ Sy, offset, mag


@pytest.mark.parametrize("name", list(CASES))
def test_grad_operator_and_transpose(gpu_lib, name):
    """vb_grad_operator: G x and G^T g against scipy's products with the matrix built from the
    exports (1e-13 of the sum of |terms| per row, terms taken before the wall rows' mirror and real
    terms cancel), the adjoint identity to 1e-12, and G p plus the mirrors' pressure offsets equal to
    the step's own pressure gradient (phase 11) to 1e-13."""
    case = CASES[name]()
    n = case.npfb
    print(f"case {case.name} npfb={n}", flush=True)
    G = solver.Solver(case)
    G.div_free()
    for k in range(1, 11):
        G.phase(k)
    Gx, Gy, offset, (aGx, aGy) = _grad_matrix(G, n)
    rng = np.random.default_rng(17)
    x = rng.standard_normal(n)
    g = rng.standard_normal((n, 2))
    gxv = G.grad_operator(x)
    _close(gxv[:, 0], Gx @ x, aGx @ np.abs(x))
    _close(gxv[:, 1], Gy @ x, aGy @ np.abs(x))
    gt = G.grad_operator(g, transpose=True)
    _close(gt, Gx.T @ g[:, 0] + Gy.T @ g[:, 1], aGx.T @ np.abs(g[:, 0]) + aGy.T @ np.abs(g[:, 1]))
    lhs, rhs = float(np.sum(gxv * g)), float(np.dot(x, gt))
    scale = float(np.abs(g[:, 0]) @ (aGx @ np.abs(x)) + np.abs(g[:, 1]) @ (aGy @ np.abs(x)))
    print(f"<Gx,g> {lhs:.17e} <x,G^T g> {rhs:.17e} (scale {scale:.3e})", flush=True)
    assert abs(lhs - rhs) <= 1e-12 * scale
    # the step's pressure gradient is G p plus the mirror offsets
    p = G.get_state()["p"]
    gp_op = G.grad_operator(p) + offset
    G.phase(11)
    gp = G.grad_p()
    sc = aGx @ np.abs(p) + aGy @ np.abs(p) + np.abs(offset).sum(axis=1) + 1e-300
    err = np.max(np.abs(gp_op - gp) / sc[:, None])
    print(f"step's grad p against G p + offsets: {err:.2e}", flush=True)
    assert err <= 1e-13
