// gfx950 pressure Poisson equation: RHS, ELL assembly and the BiCGSTAB of the reference
// (ppe_solve_in_house_simple_dirichlet.F90:1-158, linear_solver.F90:231-425).
//
// Matrix: ELL, slot-major like the neighbour list (col/val[k*npfb + i]) plus a diagonal; slot k of
// row i is the k-th neighbour of i with mirrors folded onto their parents, which is the
// reference's Numerical-Recipes row layout slot for slot, including its zero slots (duplicate
// columns accumulate into the last slot, self-mirrors into the diagonal).  The SpMV therefore
// performs the reference's sparse_atime operations in the reference's order.
//
// BiCGSTAB: five kernels per iteration (two of them SpMVs) with every scalar (rho, alpha, omega,
// norms, means) kept on the device; each reduction is a deterministic block tree whose last block
// (acq_rel ticket) folds the partials and computes the next scalar, so no extra launches and no
// host round trip per iteration.  The remove_dc projections and the inverseM scaling are applied on
// the fly inside the SpMV gathers, producing bit-identical values to a separate pass.
#include "vb_ctx.h"

namespace vb {

constexpr double kOnePlusEps = 1.0 + 1.0e-15;   // M(i) + 1e-15 with M = 1, linear_solver.F90:55,318

// ------------------------------------------------------------------------------------------------
// RHS, ppe_solve_in_house_simple_dirichlet.F90:27-53 (DC removal for the closed domain)
// ------------------------------------------------------------------------------------------------
__global__ void k_ppe_rhs(Geom g, State st, const int* cnt, const int* link, const double* wg1,
                          const double* wg2, const double* k11, const double* k12, const double* k21,
                          const double* k22, double* rhs, double* partials, Scal* sc) {
  const int npfb = g.npfb;
  RowRange rr = block_rows(npfb, gridDim.x);
  const double dt = sc->dt;
  double acc[1] = {0.0};
  for (int i = rr.begin + threadIdx.x; i < rr.end; i += kBlock) {
    const int n = cnt[i];
    const double ui = st.u[i], vi = st.v[i];
    const double K11 = k11[i], K12 = k12[i], K21 = k21[i], K22 = k22[i];
    double fb = 0.0;
    for (int k = 0; k < n; k++) {
      int s = k * npfb + i;
      int j = link[s];
      double w1 = wg1[s], w2 = wg2[s];
      double uij = ui - st.u[j], vij = vi - st.v[j];
      double kw1 = K11 * w1 + K12 * w2, kw2 = K21 * w1 + K22 * w2;
      fb = fb - (uij * kw1 + vij * kw2);
    }
    double b = fb / dt;
    rhs[i] = b;
    acc[0] += b;
  }
  block_sum<1>(acc);
  if (publish_partials<1>(acc, partials, &sc->counter)) {
    double s = fold_partials(partials, 0, gridDim.x);
    if (threadIdx.x == 0) { sc->mean = s / (double)npfb; sc->counter = 0; }
  }
}

// Assembly, :60-122.  One thread per row.
__global__ void k_assemble(Geom g, State st, const int* cnt, const int* link, const double* wl,
                           const int* nsurf, double* rhs, double* a_diag, int* a_col, double* a_val,
                           float* a_valf, float* a_diagf, const Scal* sc) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  const int npfb = g.npfb;
  if (i >= npfb) return;
  const int n = cnt[i];
  double bi = rhs[i] - sc->mean;
  double adiag = 0.0;
  bool has_m = false;
  for (int k = 0; k < n; k++) {
    int s = k * npfb + i;
    int j = link[s];
    double t = -wl[s];
    adiag = adiag - t;
    if (j >= npfb) {
      has_m = true;
      int ij = st.irel[j];
      if (ij == i) adiag = adiag + t;
      bi = bi - st.dP[j] * t;
    }
  }
  if (!has_m) {
    for (int k = 0; k < n; k++) {
      int s = k * npfb + i;
      double val = 0.0 + (-wl[s]);
      a_col[s] = link[s];
      a_val[s] = val;
      if (a_valf) a_valf[s] = (float)val;
    }
  } else {
    for (int k = 0; k < n; k++) {
      int s = k * npfb + i;
      int j = link[s];
      int ck = j < npfb ? j : st.irel[j];
      double val = 0.0;
      if (ck != i) {
        bool last = true;
        for (int k2 = k + 1; k2 < n && last; k2++) {
          int j2 = link[k2 * npfb + i];
          if ((j2 < npfb ? j2 : st.irel[j2]) == ck) last = false;
        }
        if (last) {
          double acc = 0.0;
          for (int k2 = 0; k2 <= k; k2++) {
            int j2 = link[k2 * npfb + i];
            if ((j2 < npfb ? j2 : st.irel[j2]) == ck) acc = acc + (-wl[k2 * npfb + i]);
          }
          val = acc;
        }
      }
      a_col[s] = ck;
      a_val[s] = val;
      if (a_valf) a_valf[s] = (float)val;
    }
  }
  if (nsurf[i] != 0) adiag = 1.0e30;
  double d = 0.0 + adiag;
  a_diag[i] = d;
  if (a_diagf) a_diagf[i] = (float)d;
  rhs[i] = bi;
}

// ------------------------------------------------------------------------------------------------
// SPARSE_PBICGSTAB_nonull, linear_solver.F90:275-425, FP64
// ------------------------------------------------------------------------------------------------
struct Vecs {
  double *x, *r, *rb, *v, *p, *pn, *pinv, *s, *s2, *sinv, *t;
};

__global__ void k_bi_init(int n, Vecs w, const double* rhs, double tol, int itermax, double* partials, Scal* sc) {
  RowRange rr = block_rows(n, gridDim.x);
  double acc[1] = {0.0};
  for (int i = rr.begin + threadIdx.x; i < rr.end; i += kBlock) {
    double b = rhs[i];
    w.x[i] = 0.0;            // p(1:npfb) = 0 before the solve (ppe_...:145); A*0 = 0, so r = B
    w.r[i] = b; w.rb[i] = b;
    w.v[i] = 0.0; w.p[i] = 0.0;
    acc[0] += b * b;
  }
  block_sum<1>(acc);
  if (publish_partials<1>(acc, partials, &sc->counter)) {
    double rho = fold_partials(partials, 0, gridDim.x);
    if (threadIdx.x == 0) {
      sc->rho = rho; sc->rho0 = 1.0; sc->alpha = 1.0; sc->omega = 1.0; sc->beta = 0.0;
      sc->n2s0 = 1.0; sc->n2r0 = 1.0; sc->resid = 0.0; sc->tol = tol;
      sc->iter = 1; sc->itermax = itermax; sc->exit_s = 0; sc->status = VB_OK;
      sc->done = 0;
      if (rho == 0.0) { sc->done = 1; sc->status = VB_WARN_BICGSTAB_BREAKDOWN; }
      if (itermax < 1) { sc->done = 1; sc->status = VB_WARN_MAXITER; }
      sc->counter = 0;
    }
  }
}

// :337-349 search direction, and the sum for remove_dc(p)
__global__ void k_bi_p(int n, Vecs w, double* partials, Scal* sc) {
  if (sc->done) return;
  RowRange rr = block_rows(n, gridDim.x);
  const bool first = sc->iter == 1;
  const double beta = sc->beta, omega = sc->omega;
  double acc[1] = {0.0};
  for (int i = rr.begin + threadIdx.x; i < rr.end; i += kBlock) {
    double pn = first ? w.r[i] : w.r[i] + beta * (w.p[i] - omega * w.v[i]);
    w.pn[i] = pn;
    acc[0] += pn;
  }
  block_sum<1>(acc);
  if (publish_partials<1>(acc, partials, &sc->counter)) {
    double s = fold_partials(partials, 0, gridDim.x);
    if (threadIdx.x == 0) { sc->mean_p = s / (double)n; sc->counter = 0; }
  }
}

// :351-356  p = p - mean; p_inv = p/(1+1e-15); v = A p_inv; r_bar . v; alpha
__global__ void k_bi_spmv_p(int n, Vecs w, const double* a_diag, const int* a_col, const double* a_val,
                            const int* cnt, double* partials, Scal* sc) {
  if (sc->done) return;
  RowRange rr = block_rows(n, gridDim.x);
  const double mean = sc->mean_p;
  double acc[1] = {0.0};
  for (int i = rr.begin + threadIdx.x; i < rr.end; i += kBlock) {
    double pi = w.pn[i] - mean;
    double pinv = pi / kOnePlusEps;
    w.p[i] = pi; w.pinv[i] = pinv;
    double bb = a_diag[i] * pinv;
    const int m = cnt[i];
    for (int k = 0; k < m; k++) {
      int s = k * n + i;
      bb = bb + a_val[s] * ((w.pn[a_col[s]] - mean) / kOnePlusEps);
    }
    w.v[i] = bb;
    acc[0] += w.rb[i] * bb;
  }
  block_sum<1>(acc);
  if (publish_partials<1>(acc, partials, &sc->counter)) {
    double temp = fold_partials(partials, 0, gridDim.x);
    if (threadIdx.x == 0) { sc->alpha = sc->rho / (temp + 1.0e-15); sc->counter = 0; }
  }
}

// :358-381  s = r - alpha v; |s|^2; early exit on s; the sum for remove_dc(s)
__global__ void k_bi_s(int n, Vecs w, double* partials, Scal* sc) {
  if (sc->done) return;
  RowRange rr = block_rows(n, gridDim.x);
  const double alpha = sc->alpha;
  double acc[2] = {0.0, 0.0};
  for (int i = rr.begin + threadIdx.x; i < rr.end; i += kBlock) {
    double s = w.r[i] - alpha * w.v[i];
    w.s[i] = s;
    acc[0] += s * s;
    acc[1] += s;
  }
  block_sum<2>(acc);
  if (publish_partials<2>(acc, partials, &sc->counter)) {
    double n2s = fold_partials(partials, 0, gridDim.x);
    double sum = fold_partials(partials, 1, gridDim.x);
    if (threadIdx.x == 0) {
      if (sc->iter == 1) sc->n2s0 = (n2s <= 1.0e-6) ? 1.0 : n2s;
      if (n2s <= sc->tol * sc->tol * sc->n2s0) {
        sc->exit_s = 1;
        sc->resid = sqrt(n2s / sc->n2s0);
        sc->done = 1;
      }
      sc->mean_s = sum / (double)n;
      sc->counter = 0;
    }
  }
}

// :383-389  s = s - mean; s_inv; t = A s_inv; t.s, t.t; omega.  On an early exit: x += alpha p_inv.
__global__ void k_bi_spmv_s(int n, Vecs w, const double* a_diag, const int* a_col, const double* a_val,
                            const int* cnt, double* partials, Scal* sc) {
  RowRange rr = block_rows(n, gridDim.x);
  if (sc->exit_s == 1) {
    const double alpha = sc->alpha;
    for (int i = rr.begin + threadIdx.x; i < rr.end; i += kBlock) w.x[i] = w.x[i] + alpha * w.pinv[i];
    double d[1] = {0.0};
    if (publish_partials<1>(d, partials, &sc->counter)) {
      if (threadIdx.x == 0) { sc->exit_s = 2; sc->counter = 0; }
    }
    return;
  }
  if (sc->done) return;
  const double mean = sc->mean_s;
  double acc[2] = {0.0, 0.0};
  for (int i = rr.begin + threadIdx.x; i < rr.end; i += kBlock) {
    double si = w.s[i] - mean;
    double sinv = si / kOnePlusEps;
    w.s2[i] = si; w.sinv[i] = sinv;
    double bb = a_diag[i] * sinv;
    const int m = cnt[i];
    for (int k = 0; k < m; k++) {
      int s = k * n + i;
      bb = bb + a_val[s] * ((w.s[a_col[s]] - mean) / kOnePlusEps);
    }
    w.t[i] = bb;
    acc[0] += bb * si;
    acc[1] += bb * bb;
  }
  block_sum<2>(acc);
  if (publish_partials<2>(acc, partials, &sc->counter)) {
    double ts = fold_partials(partials, 0, gridDim.x);
    double tt = fold_partials(partials, 1, gridDim.x);
    if (threadIdx.x == 0) { sc->omega = ts / (tt + 1.0e-15); sc->counter = 0; }
  }
}

// :390-418  x, r updates; |r|^2; convergence; next rho and beta
__global__ void k_bi_xr(int n, Vecs w, double* partials, Scal* sc) {
  if (sc->done) return;
  RowRange rr = block_rows(n, gridDim.x);
  const double alpha = sc->alpha, omega = sc->omega;
  double acc[2] = {0.0, 0.0};
  for (int i = rr.begin + threadIdx.x; i < rr.end; i += kBlock) {
    w.x[i] = w.x[i] + alpha * w.pinv[i] + omega * w.sinv[i];
    double r = w.s2[i] - omega * w.t[i];
    w.r[i] = r;
    acc[0] += r * r;
    acc[1] += w.rb[i] * r;
  }
  block_sum<2>(acc);
  if (publish_partials<2>(acc, partials, &sc->counter)) {
    double n2r = fold_partials(partials, 0, gridDim.x);
    double rho = fold_partials(partials, 1, gridDim.x);
    if (threadIdx.x == 0) {
      if (sc->iter == 1) sc->n2r0 = (n2r <= 1.0e-6) ? 1.0 : n2r;
      sc->resid = sqrt(n2r / sc->n2r0);
      if (n2r <= sc->tol * sc->tol * sc->n2r0) {
        sc->done = 1;
      } else {
        sc->iter = sc->iter + 1;
        sc->rho0 = sc->rho;
        sc->rho = rho;
        if (sc->iter > sc->itermax) { sc->done = 1; sc->status = VB_WARN_MAXITER; }
        else if (rho == 0.0) { sc->done = 1; sc->status = VB_WARN_BICGSTAB_BREAKDOWN; }
        else sc->beta = rho / (sc->rho0 + 1.0e-15) * sc->alpha / (sc->omega + 1.0e-15);
      }
      sc->counter = 0;
    }
  }
}

static vb_status solve_fp64(vb_ctx* c, double tol, int itermax) {
  const int n = c->npfb;
  const int nb = grid_for(n);
  Vecs w{c->st.p, c->r, c->rb, c->vv, c->pp, c->pn, c->pinv, c->ss, c->s2, c->sinv, c->tt};
  VB_LAUNCH(c, k_bi_init, nb, n, w, c->rhs, tol, itermax, c->partials, c->scal);
  int launched = 0, batch = 8;
  for (;;) {
    for (int b = 0; b < batch; b++) {
      VB_LAUNCH(c, k_bi_p, nb, n, w, c->partials, c->scal);
      VB_LAUNCH(c, k_bi_spmv_p, nb, n, w, c->a_diag, c->a_col, c->a_val, c->cnt, c->partials, c->scal);
      VB_LAUNCH(c, k_bi_s, nb, n, w, c->partials, c->scal);
      VB_LAUNCH(c, k_bi_spmv_s, nb, n, w, c->a_diag, c->a_col, c->a_val, c->cnt, c->partials, c->scal);
      VB_LAUNCH(c, k_bi_xr, nb, n, w, c->partials, c->scal);
    }
    launched += batch;
    VB_HIP(c, hipMemcpyAsync(c->scal_host, c->scal, sizeof(Scal), hipMemcpyDeviceToHost, c->stream));
    VB_HIP(c, hipStreamSynchronize(c->stream));
    if (c->scal_host->done && c->scal_host->exit_s != 1) break;
    if (launched > itermax + 2) break;
    // launch roughly as many as are likely still needed, capped so the overshoot stays small
    batch = launched < 32 ? 8 : 16;
  }
  return VB_OK;
}

vb_status launch_ppe(vb_ctx* c, double tol, int itermax) {
  const int nb = grid_for(c->npfb);
  VB_LAUNCH(c, k_ppe_rhs, nb, c->g, c->st, c->cnt, c->link, c->wg1, c->wg2, c->k11, c->k12, c->k21,
            c->k22, c->rhs, c->partials, c->scal);
  VB_LAUNCH(c, k_assemble, grid_1d(c->npfb), c->g, c->st, c->cnt, c->link, c->wl, c->nsurf, c->rhs,
            c->a_diag, c->a_col, c->a_val, nullptr, nullptr, c->scal);
  return solve_fp64(c, tol, itermax);
}

}  // namespace vb
