// gfx950 pressure Poisson equation: RHS, ELL assembly and the BiCGSTAB of the reference
// (ppe_solve_in_house_simple_dirichlet.F90:1-158, linear_solver.F90:231-425).
//
// Matrix: ELL, slot-major like the neighbour list (col/val[k*npfb + i]) plus a diagonal; slot k of
// row i is the k-th neighbour of i with mirrors folded onto their parents, which is the
// reference's Numerical-Recipes row layout slot for slot, including its zero slots (duplicate
// columns accumulate into the last slot, self-mirrors into the diagonal).  The SpMV therefore
// performs the reference's sparse_atime operations in the reference's order.
//
// BiCGSTAB: five row kernels per iteration (two of them SpMVs), each followed by a one-workgroup
// finalize that folds the block partials in a fixed order and advances the scalars (rho, alpha,
// omega, norms, means), all kept on the device: no host round trip per iteration.  The remove_dc
// projections and the inverseM scaling are applied on the fly inside the SpMV gathers, producing
// bit-identical values to separate passes.
#include "vb_ctx.h"

namespace vb {

constexpr double kOnePlusEps = 1.0 + 1.0e-15;   // M(i) + 1e-15 with M = 1, linear_solver.F90:55,318

// ------------------------------------------------------------------------------------------------
// RHS, ppe_solve_in_house_simple_dirichlet.F90:27-53 (DC removal for the closed domain)
// ------------------------------------------------------------------------------------------------
__global__ void k_ppe_rhs(Geom g, State st, const int* cnt, const int* link, const double* wg1,
                          const double* wg2, const double* k11, const double* k12, const double* k21,
                          const double* k22, double* rhs, double* partials, double* red, const Scal* sc) {
  const int npfb = g.npfb;
  RowRange rr = block_rows(npfb, gridDim.x);
  const double dt = sc->dt;
  Red<1> R(red, npfb);
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
    R.add(0, i, b);
  }
  R.store(partials);
}

__global__ void k_fin_ppe_rhs(const double* partials, int nblk, const double* red, int n, Scal* sc) {
  double s = fold(partials, 0, nblk, red, n);
  if (threadIdx.x == 0) sc->mean = s / (double)n;
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
// SPARSE_PBICGSTAB_nonull, linear_solver.F90:275-425, FP64.  Each row kernel is followed by a
// one-workgroup finalize that folds its partials and advances the scalars.
// ------------------------------------------------------------------------------------------------
struct Vecs {
  double *x, *r, *rb, *v, *p, *pn, *pinv, *s, *s2, *sinv, *t;
};

__global__ void k_bi_init(int n, Vecs w, const double* rhs, double* partials, double* red) {
  RowRange rr = block_rows(n, gridDim.x);
  Red<1> R(red, n);
  for (int i = rr.begin + threadIdx.x; i < rr.end; i += kBlock) {
    double b = rhs[i];
    w.x[i] = 0.0;            // p(1:npfb) = 0 before the solve (ppe_...:145); A*0 = 0, so r = B
    w.r[i] = b; w.rb[i] = b;
    w.v[i] = 0.0; w.p[i] = 0.0;
    R.add(0, i, b * b);      // rho = vtime(r_bar, r) of the first iteration (:330)
  }
  R.store(partials);
}

__global__ void k_fin_init(const double* partials, int nblk, const double* red, int n, double tol,
                           int itermax, Scal* sc) {
  double rho = fold(partials, 0, nblk, red, n);
  if (threadIdx.x == 0) {
    sc->rho = rho; sc->rho0 = 1.0; sc->alpha = 1.0; sc->omega = 1.0; sc->beta = 0.0;
    sc->n2s0 = 1.0; sc->n2r0 = 1.0; sc->resid = 0.0; sc->tol = tol;
    sc->iter = 1; sc->itermax = itermax; sc->exit_s = 0; sc->status = VB_OK;
    sc->done = 0;
    if (itermax < 1) { sc->done = 1; sc->status = VB_WARN_MAXITER; }
    else if (rho == 0.0) { sc->done = 1; sc->status = VB_WARN_BICGSTAB_BREAKDOWN; }
  }
}

// :337-349 search direction; the sum for remove_dc(p)
__global__ void k_bi_p(int n, Vecs w, double* partials, double* red, const Scal* sc) {
  if (sc->done) return;
  RowRange rr = block_rows(n, gridDim.x);
  const bool first = sc->iter == 1;
  const double beta = sc->beta, omega = sc->omega;
  Red<1> R(red, n);
  for (int i = rr.begin + threadIdx.x; i < rr.end; i += kBlock) {
    double pn = first ? w.r[i] : w.r[i] + beta * (w.p[i] - omega * w.v[i]);
    w.pn[i] = pn;
    R.add(0, i, pn);
  }
  R.store(partials);
}

__global__ void k_fin_p(const double* partials, int nblk, const double* red, int n, Scal* sc) {
  if (sc->done) return;
  double s = fold(partials, 0, nblk, red, n);
  if (threadIdx.x == 0) sc->mean_p = s / (double)n;                      // :243
}

// ELL SpMV y_i = A(i) x_i + sum_k val(i,k) x_col(i,k) (sparse_atime, :264-272) with the vector
// produced on the fly as x = (src - mean) / (1 + 1e-15) (remove_dc then inverseM, bit-identical to
// the separate passes).  Four slots per step keep four gathers in flight; the sum keeps slot order.
__device__ inline double spmv_row(int i, int n, int m, double xi, const double* __restrict__ a_diag,
                                  const int* __restrict__ a_col, const double* __restrict__ a_val,
                                  const double* __restrict__ src, double mean) {
  double bb = a_diag[i] * xi;
  int k = 0;
  for (; k + 4 <= m; k += 4) {
    int s0 = k * n + i;
    int c0 = a_col[s0], c1 = a_col[s0 + n], c2 = a_col[s0 + 2 * n], c3 = a_col[s0 + 3 * n];
    double v0 = a_val[s0], v1 = a_val[s0 + n], v2 = a_val[s0 + 2 * n], v3 = a_val[s0 + 3 * n];
    double x0 = src[c0], x1 = src[c1], x2 = src[c2], x3 = src[c3];
    bb = bb + v0 * ((x0 - mean) / kOnePlusEps);
    bb = bb + v1 * ((x1 - mean) / kOnePlusEps);
    bb = bb + v2 * ((x2 - mean) / kOnePlusEps);
    bb = bb + v3 * ((x3 - mean) / kOnePlusEps);
  }
  for (; k < m; k++) {
    int s0 = k * n + i;
    bb = bb + a_val[s0] * ((src[a_col[s0]] - mean) / kOnePlusEps);
  }
  return bb;
}

// :351-356  p = p - mean; p_inv = p/(1+1e-15); v = A p_inv; the terms of r_bar . v
__global__ void k_bi_spmv_p(int n, Vecs w, const double* a_diag, const int* a_col, const double* a_val,
                            const int* cnt, double* partials, double* red, const Scal* sc) {
  if (sc->done) return;
  RowRange rr = block_rows(n, gridDim.x);
  const double mean = sc->mean_p;
  Red<1> R(red, n);
  for (int i = rr.begin + threadIdx.x; i < rr.end; i += kBlock) {
    double pi = w.pn[i] - mean;
    double pinv = pi / kOnePlusEps;
    w.p[i] = pi; w.pinv[i] = pinv;
    double bb = spmv_row(i, n, cnt[i], pinv, a_diag, a_col, a_val, w.pn, mean);
    w.v[i] = bb;
    R.add(0, i, w.rb[i] * bb);
  }
  R.store(partials);
}

__global__ void k_fin_spmv_p(const double* partials, int nblk, const double* red, int n, Scal* sc) {
  if (sc->done) return;
  double temp = fold(partials, 0, nblk, red, n);
  if (threadIdx.x == 0) sc->alpha = sc->rho / (temp + 1.0e-15);          // :356
}

// :358-364  s = r - alpha v; |s|^2 and sum(s)
__global__ void k_bi_s(int n, Vecs w, double* partials, double* red, const Scal* sc) {
  if (sc->done) return;
  RowRange rr = block_rows(n, gridDim.x);
  const double alpha = sc->alpha;
  Red<2> R(red, n);
  for (int i = rr.begin + threadIdx.x; i < rr.end; i += kBlock) {
    double s = w.r[i] - alpha * w.v[i];
    w.s[i] = s;
    R.add(0, i, s * s);
    R.add(1, i, s);
  }
  R.store(partials);
}

// :366-381 early exit on s
__global__ void k_fin_s(const double* partials, int nblk, const double* red, int n, Scal* sc) {
  if (sc->done) return;
  double n2s = fold(partials, 0, nblk, red, n);
  double sum = fold(partials, 1, nblk, red, n);
  if (threadIdx.x == 0) {
    if (sc->iter == 1) sc->n2s0 = (n2s <= 1.0e-6) ? 1.0 : n2s;
    if (n2s <= sc->tol * sc->tol * sc->n2s0) {
      sc->exit_s = 1;                        // x += alpha p_inv happens in the next row kernel
      sc->resid = sqrt(n2s / sc->n2s0);
      sc->done = 1;
    }
    sc->mean_s = sum / (double)n;
  }
}

// :383-387  s = s - mean; s_inv; t = A s_inv; terms of t.s and t.t.  On an early exit: x += alpha p_inv.
__global__ void k_bi_spmv_s(int n, Vecs w, const double* a_diag, const int* a_col, const double* a_val,
                            const int* cnt, double* partials, double* red, const Scal* sc) {
  RowRange rr = block_rows(n, gridDim.x);
  if (sc->exit_s == 1) {
    const double alpha = sc->alpha;
    for (int i = rr.begin + threadIdx.x; i < rr.end; i += kBlock) w.x[i] = w.x[i] + alpha * w.pinv[i];
    return;
  }
  if (sc->done) return;
  const double mean = sc->mean_s;
  Red<2> R(red, n);
  for (int i = rr.begin + threadIdx.x; i < rr.end; i += kBlock) {
    double si = w.s[i] - mean;
    double sinv = si / kOnePlusEps;
    w.s2[i] = si; w.sinv[i] = sinv;
    double bb = spmv_row(i, n, cnt[i], sinv, a_diag, a_col, a_val, w.s, mean);
    w.t[i] = bb;
    R.add(0, i, bb * si);
    R.add(1, i, bb * bb);
  }
  R.store(partials);
}

__global__ void k_fin_spmv_s(const double* partials, int nblk, const double* red, int n, Scal* sc) {
  if (sc->exit_s == 1) { if (threadIdx.x == 0) sc->exit_s = 2; return; }
  if (sc->done) return;
  double ts = fold(partials, 0, nblk, red, n);
  double tt = fold(partials, 1, nblk, red, n);
  if (threadIdx.x == 0) sc->omega = ts / (tt + 1.0e-15);                 // :389
}

// :390-401  x, r updates; terms of |r|^2 and of the next r_bar . r
__global__ void k_bi_xr(int n, Vecs w, double* partials, double* red, const Scal* sc) {
  if (sc->done) return;
  RowRange rr = block_rows(n, gridDim.x);
  const double alpha = sc->alpha, omega = sc->omega;
  Red<2> R(red, n);
  for (int i = rr.begin + threadIdx.x; i < rr.end; i += kBlock) {
    w.x[i] = w.x[i] + alpha * w.pinv[i] + omega * w.sinv[i];
    double r = w.s2[i] - omega * w.t[i];
    w.r[i] = r;
    R.add(0, i, r * r);
    R.add(1, i, w.rb[i] * r);
  }
  R.store(partials);
}

// :403-419 convergence; next rho (:330) and beta (:342-343)
__global__ void k_fin_xr(const double* partials, int nblk, const double* red, int n, Scal* sc) {
  if (sc->done) return;
  double n2r = fold(partials, 0, nblk, red, n);
  double rho = fold(partials, 1, nblk, red, n);
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
  }
}

#define VB_FIN(c, kern, ...)                                                               \
  do {                                                                                     \
    hipLaunchKernelGGL(kern, dim3(1), dim3(kBlock), 0, (c)->stream, __VA_ARGS__);          \
    VB_HIP(c, hipGetLastError());                                                          \
  } while (0)

static vb_status solve_fp64(vb_ctx* c, double tol, int itermax) {
  const int n = c->npfb;
  const int nb = grid_for(n);
  Vecs w{c->st.p, c->r, c->rb, c->vv, c->pp, c->pn, c->pinv, c->ss, c->s2, c->sinv, c->tt};
  double* P = c->partials;
  double* red = c->red;           // null unless serial reductions were requested
  Scal* sc = c->scal;
  VB_LAUNCH(c, k_bi_init, nb, n, w, c->rhs, P, red);
  VB_FIN(c, k_fin_init, P, nb, red, n, tol, itermax, sc);
  int launched = 0, batch = 8;
  for (;;) {
    for (int b = 0; b < batch; b++) {
      VB_LAUNCH(c, k_bi_p, nb, n, w, P, red, sc);
      VB_FIN(c, k_fin_p, P, nb, red, n, sc);
      VB_LAUNCH(c, k_bi_spmv_p, nb, n, w, c->a_diag, c->a_col, c->a_val, c->cnt, P, red, sc);
      VB_FIN(c, k_fin_spmv_p, P, nb, red, n, sc);
      VB_LAUNCH(c, k_bi_s, nb, n, w, P, red, sc);
      VB_FIN(c, k_fin_s, P, nb, red, n, sc);
      VB_LAUNCH(c, k_bi_spmv_s, nb, n, w, c->a_diag, c->a_col, c->a_val, c->cnt, P, red, sc);
      VB_FIN(c, k_fin_spmv_s, P, nb, red, n, sc);
      VB_LAUNCH(c, k_bi_xr, nb, n, w, P, red, sc);
      VB_FIN(c, k_fin_xr, P, nb, red, n, sc);
    }
    launched += batch;
    VB_HIP(c, hipMemcpyAsync(c->scal_host, c->scal, sizeof(Scal), hipMemcpyDeviceToHost, c->stream));
    VB_HIP(c, hipStreamSynchronize(c->stream));
    if (c->scal_host->done && c->scal_host->exit_s != 1) break;
    if (launched > itermax + 2) break;
    batch = launched < 32 ? 8 : 32;
  }
  return VB_OK;
}

__global__ void k_bi_set_live(Scal* sc, int* nnz_out, const int* cnt, int n) {
  // reopen the solver kernels after a converged solve and count the stored slots
  __shared__ int sh[kBlock];
  int s = 0;
  for (int i = threadIdx.x; i < n; i += kBlock) s += cnt[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) { sc->done = 0; sc->exit_s = 0; *nnz_out = sh[0]; }
}

vb_status profile_spmv(vb_ctx* c, int reps, double* avg_us, double* alg_bytes) {
  const int n = c->npfb, nb = grid_for(n);
  Vecs w{c->st.p, c->r, c->rb, c->vv, c->pp, c->pn, c->pinv, c->ss, c->s2, c->sinv, c->tt};
  int* d_nnz = c->seg_counts;
  hipLaunchKernelGGL(k_bi_set_live, dim3(1), dim3(kBlock), 0, c->stream, c->scal, d_nnz, c->cnt, n);
  VB_HIP(c, hipGetLastError());
  hipEvent_t e0, e1;
  VB_HIP(c, hipEventCreate(&e0));
  VB_HIP(c, hipEventCreate(&e1));
  VB_LAUNCH(c, k_bi_spmv_p, nb, n, w, c->a_diag, c->a_col, c->a_val, c->cnt, c->partials, nullptr, c->scal);
  VB_HIP(c, hipEventRecord(e0, c->stream));
  for (int r = 0; r < reps; r++)
    VB_LAUNCH(c, k_bi_spmv_p, nb, n, w, c->a_diag, c->a_col, c->a_val, c->cnt, c->partials, nullptr, c->scal);
  VB_HIP(c, hipEventRecord(e1, c->stream));
  VB_HIP(c, hipEventSynchronize(e1));
  float ms = 0.f;
  VB_HIP(c, hipEventElapsedTime(&ms, e0, e1));
  int nnz = 0;
  VB_HIP(c, hipMemcpy(&nnz, d_nnz, 4, hipMemcpyDeviceToHost));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *avg_us = 1000.0 * ms / reps;
  // per row: cnt 4, diag 8, gathered p 8 (once), r_bar 8, writes p, p_inv, v 24; per slot: col 4 + val 8
  *alg_bytes = 52.0 * n + 12.0 * nnz;
  return VB_OK;
}

vb_status launch_ppe(vb_ctx* c, double tol, int itermax) {
  const int nb = grid_for(c->npfb);
  VB_LAUNCH(c, k_ppe_rhs, nb, c->g, c->st, c->cnt, c->link, c->wg1, c->wg2, c->k11, c->k12, c->k21,
            c->k22, c->rhs, c->partials, c->red, c->scal);
  VB_FIN(c, k_fin_ppe_rhs, c->partials, nb, c->red, c->npfb, c->scal);
  VB_LAUNCH(c, k_assemble, grid_1d(c->npfb), c->g, c->st, c->cnt, c->link, c->wl, c->nsurf, c->rhs,
            c->a_diag, c->a_col, c->a_val, nullptr, nullptr, c->scal);
  return solve_fp64(c, tol, itermax);
}

}  // namespace vb
