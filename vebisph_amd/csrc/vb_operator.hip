// The assembled PPE operator and its transpose, for the adjoint of the step (the reference has no
// adjoint: this is the operator such a solve needs, A^T lambda = g on the matrix that
// ppe_solve_in_house_simple_dirichlet.F90:55-138 assembles).
//
// The transposed map.  Row c of A^T holds A(i, c) for every row i whose stored row references
// column c.  Between real particles the neighbour relation is symmetric (the pair test
// 0 < |r_ij|^2 <= 9h^2 of neighbour_finding_mod.F90:353-359 is the same both ways), so the real
// entries of column c come from exactly the real neighbours i = link(c, m) of c, and slot m of row c
// of A^T can hold A(i, c) = a_val(i, k) with k the slot of c in i's list ("reverse slot"): A^T's
// real part has the forward ELL shape and gathers the same columns.  Mirror slots of c's own list
// are gaps (value 0, column c).  Entries whose column is the parent of a mirror (a row i near a
// wall referencing a mirror of c; these occur only in the rows with mirror neighbours) go to a
// per-column bucket: counted, scanned, filled, then each bucket sorted by (i, k), so the order of
// every sum is fixed and the result the same run after run.  Self-mirror slots sit in the
// diagonal already (value 0 in the row).
//
// Cost: the reverse-slot search reads ~cnt/2 list entries per slot (~10 ms per build at 1M); it is
// built once per assembled matrix (on first use) and each application is one SpMV-shaped pass.
#include <algorithm>
#include <vector>

#include "vb_ctx.h"

namespace vb {

// reverse slots: tval/tcol[m*npfb + c] for the real part of row c of A^T; nmiss counts real pairs
// whose reverse slot is absent (an asymmetric list, which the pair test rules out)
__global__ void k_transpose_real(int npfb, const int* cnt, const int* link, const double* a_val, double* tval,
                                 int* tcol, int* rslot, int* nmiss) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= npfb) return;
  const int n = cnt[c];
  int miss = 0;
  for (int m = 0; m < n; m++) {
    const size_t s = (size_t)m * npfb + c;
    const int i = link[s];
    double v = 0.0;
    int col = c, rk = -1;
    if (i < npfb) {
      const int ni = cnt[i];
      int k = 0;
      while (k < ni && link[(size_t)k * npfb + i] != c) k++;
      if (k < ni) { v = a_val[(size_t)k * npfb + i]; col = i; rk = k; }
      else miss++;
    }
    tval[s] = v;
    tcol[s] = col;
    rslot[s] = rk;
  }
  if (miss) atomicAdd(nmiss, miss);
}

// Mirror entries of the explicit rows: (i, k) with link(i, k) a mirror whose parent c != i.
__device__ inline int mirror_parent(int npfb, const int* irel, int j) { return j >= npfb ? irel[j] : -1; }

__global__ void k_transpose_mirror_count(int npfb, const int* rows, int nrows, const int* cnt, const int* link,
                                         const int* irel, int* mcount) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= nrows) return;
  const int i = rows[r], n = cnt[i];
  for (int k = 0; k < n; k++) {
    const int c = mirror_parent(npfb, irel, link[(size_t)k * npfb + i]);
    if (c >= 0 && c != i) atomicAdd(&mcount[c], 1);
  }
}

__global__ void k_transpose_mirror_fill(int npfb, const int* rows, int nrows, const int* cnt, const int* link,
                                        const int* irel, const int* moff, int* mfill, int* mkey) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= nrows) return;
  const int i = rows[r], n = cnt[i];
  for (int k = 0; k < n; k++) {
    const int c = mirror_parent(npfb, irel, link[(size_t)k * npfb + i]);
    if (c >= 0 && c != i) {
      const int pos = atomicAdd(&mfill[c], 1);
      mkey[moff[c] + pos] = i * kNplink + k;                // (row, slot), unique per entry
    }
  }
}

// each column's bucket in increasing (row, slot) order, then the values
__global__ void k_transpose_mirror_sort(int npfb, const int* moff, int* mkey, int* mrow, double* mval,
                                        const double* a_val) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= npfb) return;
  const int b = moff[c], e = moff[c + 1];
  for (int p = b + 1; p < e; p++) {
    const int key = mkey[p];
    int q = p - 1;
    while (q >= b && mkey[q] > key) { mkey[q + 1] = mkey[q]; q--; }
    mkey[q + 1] = key;
  }
  for (int p = b; p < e; p++) {
    const int i = mkey[p] / kNplink, k = mkey[p] % kNplink;
    mrow[p] = i;
    mval[p] = a_val[(size_t)k * npfb + i];
  }
}

// y = A^T x: the diagonal, the real part in c's slot order, then c's mirror bucket
__global__ void k_apply_transpose(int npfb, const int* cnt, const double* diag, const double* tval,
                                  const int* tcol, const int* moff, const int* mrow, const double* mval,
                                  const double* x, double* y) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= npfb) return;
  double acc = diag[c] * x[c];
  const int n = cnt[c];
  for (int m = 0; m < n; m++) {
    const size_t s = (size_t)m * npfb + c;
    acc = acc + tval[s] * x[tcol[s]];
  }
  for (int p = moff[c]; p < moff[c + 1]; p++) acc = acc + mval[p] * x[mrow[p]];
  y[c] = acc;
}

// y = A x on the full stored rows in the reference's slot order (sparse_atime, linear_solver.F90:264-272)
__global__ void k_apply_forward(int npfb, const int* cnt, const double* diag, const int* a_col,
                                const double* a_val, const double* x, double* y) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= npfb) return;
  double bb = diag[i] * x[i];
  const int n = cnt[i];
  for (int k = 0; k < n; k++) {
    const size_t s = (size_t)k * npfb + i;
    bb = bb + a_val[s] * x[a_col[s]];
  }
  y[i] = bb;
}

__global__ void k_op_zero(int* a, int n) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) a[i] = 0;
}

// The operator's buffers (per context, allocated on first use, or up front by the adjoint sweep,
// which runs no allocation inside its step loop): the transposed map, its mirror buckets of at least
// bucket_cap entries (grow-only), the adjoint solve's g and lambda.
vb_status alloc_operator(vb_ctx* c, size_t bucket_cap) {
  const int npfb = c->npfb;
  OperatorMap& op = c->op;
  if (!op.tval) {
    const size_t nw = (size_t)npfb * kNplink;
    VB_HIP(c, hipMalloc((void**)&op.tval, nw * sizeof(double)));
    VB_HIP(c, hipMalloc((void**)&op.tcol, nw * sizeof(int)));
    VB_HIP(c, hipMalloc((void**)&op.rslot, nw * sizeof(int)));
    VB_HIP(c, hipMalloc((void**)&op.mcount, ((size_t)npfb + 1) * sizeof(int)));
    VB_HIP(c, hipMalloc((void**)&op.moff, ((size_t)npfb + 1) * sizeof(int)));
    VB_HIP(c, hipMalloc((void**)&op.mfill, ((size_t)npfb + 1) * sizeof(int)));
    VB_HIP(c, hipMalloc((void**)&op.nmiss, sizeof(int)));
    VB_HIP(c, hipMalloc((void**)&op.bv64T, ppe_block_values(npfb) * sizeof(double)));   // as vb_ctx::bv64
    if (c->prm.precision == VB_PREC_MIXED)
      VB_HIP(c, hipMalloc((void**)&op.bv32T, ppe_block_values(npfb) * sizeof(float)));   // as vb_ctx::bv32
    VB_HIP(c, hipMalloc((void**)&op.xtcol, ppe_list_slots(npfb) * sizeof(int)));         // as vb_ctx::xcol
    VB_HIP(c, hipMalloc((void**)&op.xtval, ppe_list_slots(npfb) * sizeof(double)));
  }
  if (!op.g) {
    VB_HIP(c, hipMalloc((void**)&op.g, (size_t)npfb * sizeof(double)));
    VB_HIP(c, hipMalloc((void**)&op.lam, (size_t)npfb * sizeof(double)));
  }
  if (!op.rref) {
    VB_HIP(c, hipMalloc((void**)&op.rref, (size_t)npfb * sizeof(double)));
    VB_HIP(c, hipMalloc((void**)&op.bref, (size_t)npfb * sizeof(double)));
    VB_HIP(c, hipMalloc((void**)&op.dref, (size_t)npfb * sizeof(double)));
    VB_HIP(c, hipMalloc((void**)&op.refs, 4 * sizeof(double)));
  }
  if (bucket_cap > op.cap_m || !op.mkey) {
    if (op.mkey) { VB_HIP(c, hipFree(op.mkey)); VB_HIP(c, hipFree(op.mrow)); VB_HIP(c, hipFree(op.mval)); }
    op.mkey = op.mrow = nullptr; op.mval = nullptr;
    const size_t cap = std::max<size_t>(bucket_cap, 64);
    VB_HIP(c, hipMalloc((void**)&op.mkey, cap * sizeof(int)));
    VB_HIP(c, hipMalloc((void**)&op.mrow, cap * sizeof(int)));
    VB_HIP(c, hipMalloc((void**)&op.mval, cap * sizeof(double)));
    op.cap_m = cap;
  }
  return VB_OK;
}

// The transposed map of the matrix the last vb_ppe_solve assembled (kept until the next assembly).
vb_status build_transpose(vb_ctx* c) {
  const int npfb = c->npfb;
  OperatorMap& op = c->op;
  vb_status s0;
  if (!op.tval && (s0 = alloc_operator(c, 0)) != VB_OK) return s0;
  VB_LAUNCH(c, k_op_zero, grid_1d(npfb + 1), op.mcount, npfb + 1);
  VB_LAUNCH(c, k_op_zero, grid_1d(npfb + 1), op.mfill, npfb + 1);
  VB_LAUNCH(c, k_op_zero, 1, op.nmiss, 1);
  VB_LAUNCH(c, k_transpose_real, grid_1d(npfb), npfb, c->cnt, c->link, c->a_val, op.tval, op.tcol, op.rslot,
            op.nmiss);
  if ((s0 = launch_transpose_blocks(c)) != VB_OK) return s0;
  // the explicit rows listed by the last assembly (the rows with mirror neighbours)
  const int nx = c->scal_host->nexplicit;
  if (nx > 0)
    VB_LAUNCH(c, k_transpose_mirror_count, grid_1d(nx), npfb, c->xrows, nx, c->cnt, c->link, c->st.irel,
              op.mcount);
  vb_status s = launch_exclusive_scan(c, op.mcount, op.moff, npfb);
  if (s != VB_OK) return s;
  int hv[2] = {0, 0};
  VB_HIP(c, hipMemcpyAsync(&hv[0], op.moff + npfb, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  VB_HIP(c, hipMemcpyAsync(&hv[1], op.nmiss, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  VB_HIP(c, hipStreamSynchronize(c->stream));
  if (hv[1] != 0) {
    c->err = "transposed PPE map: a neighbour pair without its reverse entry";
    return VB_ERR_CAPACITY;
  }
  const size_t nm = (size_t)hv[0];
  if (nm > op.cap_m && (s = alloc_operator(c, nm + nm / 4 + 64)) != VB_OK) return s;   // grow-only
  if (nx > 0)
    VB_LAUNCH(c, k_transpose_mirror_fill, grid_1d(nx), npfb, c->xrows, nx, c->cnt, c->link, c->st.irel, op.moff,
              op.mfill, op.mkey);
  VB_LAUNCH(c, k_transpose_mirror_sort, grid_1d(npfb), npfb, op.moff, op.mkey, op.mrow, op.mval, c->a_val);
  op.valid = true;
  return VB_OK;
}

vb_status apply_ppe_operator(vb_ctx* c, bool transpose, const double* x, double* y) {
  const int npfb = c->npfb;
  if (!c->op.step_matrix) {
    c->err = "the PPE operator needs the matrix of a step's vb_ppe_solve (none, or replaced by a caller's system)";
    return VB_ERR_ARG;
  }
  if (transpose) {
    if (!c->op.valid) {
      vb_status s = build_transpose(c);
      if (s != VB_OK) return s;
    }
    VB_LAUNCH(c, k_apply_transpose, grid_1d(npfb), npfb, c->cnt, c->a_diag, c->op.tval, c->op.tcol, c->op.moff,
              c->op.mrow, c->op.mval, x, y);
  } else {
    VB_LAUNCH(c, k_apply_forward, grid_1d(npfb), npfb, c->cnt, c->a_diag, c->a_col, c->a_val, x, y);
  }
  return VB_OK;
}

// The SPH gradient of calc_pressure_gradient (gradient_calculation.F90:16-48) as a linear map of a
// particle field x (npfb): row i gives K_i sum_k (x_parent(j_k) - x_i) grad W_ik, mirrors taking
// their parent's value (the pressure mirrors' offset dP is constant, so it is no part of the map).
// Forward: y = [gx; gy] (2 npfb), the step's slot order.
template <class WT>
__global__ void k_grad_forward(int npfb, const int* cnt, const int* link, const int* irel, const WT* wg1,
                               const WT* wg2, const double* k11, const double* k12, const double* k21,
                               const double* k22, const double* x, double* y) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= npfb) return;
  const int n = cnt[i];
  const double xi = x[i];
  double gx = 0.0, gy = 0.0;
  for (int k = 0; k < n; k++) {
    const size_t s = (size_t)k * npfb + i;
    const int j = link[s];
    const double t = x[j < npfb ? j : irel[j]] - xi;
    gx = gx + t * wg1[s]; gy = gy + t * wg2[s];
  }
  y[i] = k11[i] * gx + k12[i] * gy;
  y[npfb + i] = k21[i] * gx + k22[i] * gy;
}

// Transposed gradient, y = G^T [gx; gy]: with a_i = K11 gx + K21 gy and b_i = K12 gx + K22 gy (the
// weights of row i's two sums), column c collects a_i dW/dx + b_i dW/dy of every slot (i, k) whose
// particle is c or a mirror of c, less its own row's slots (the -x_i term; a slot holding c's own
// mirror cancels and is skipped).  Real pairs through the reverse slots in c's slot order, then c's
// mirror bucket in (row, slot) order, then the own-row term: a fixed order.
__global__ void k_grad_coef(int npfb, const double* k11, const double* k12, const double* k21,
                            const double* k22, const double* g, double* a, double* b) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= npfb) return;
  const double gx = g[i], gy = g[npfb + i];
  a[i] = k11[i] * gx + k21[i] * gy;
  b[i] = k12[i] * gx + k22[i] * gy;
}

template <class WT>
__global__ void k_grad_transpose(int npfb, const int* cnt, const int* link, const int* irel, const WT* wg1,
                                 const WT* wg2, const int* tcol, const int* rslot, const int* moff,
                                 const int* mkey, const double* a, const double* b, double* y) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= npfb) return;
  const int n = cnt[c];
  double acc = 0.0;
  for (int m = 0; m < n; m++) {
    const size_t s = (size_t)m * npfb + c;
    const int k = rslot[s];
    if (k < 0) continue;                                   // a mirror slot of c's own list
    const int i = tcol[s];
    const size_t r = (size_t)k * npfb + i;
    acc = acc + (a[i] * wg1[r] + b[i] * wg2[r]);
  }
  for (int p = moff[c]; p < moff[c + 1]; p++) {
    const int i = mkey[p] / kNplink, k = mkey[p] % kNplink;
    const size_t r = (size_t)k * npfb + i;
    acc = acc + (a[i] * wg1[r] + b[i] * wg2[r]);
  }
  double own = 0.0;
  const double ac = a[c], bc = b[c];
  for (int k = 0; k < n; k++) {
    const size_t s = (size_t)k * npfb + c;
    const int j = link[s];
    if (j >= npfb && irel[j] == c) continue;
    own = own + (ac * wg1[s] + bc * wg2[s]);
  }
  y[c] = acc - own;
}

vb_status apply_grad_operator(vb_ctx* c, bool transpose, const double* x, double* y) {
  const int npfb = c->npfb;
  if (!c->op.step_matrix) {
    c->err = "the gradient operator needs the lists and weights of a step up to its vb_ppe_solve";
    return VB_ERR_ARG;
  }
  if (!transpose)
    return with_weights(c, [&](auto* w1, auto* w2, auto*) -> vb_status {
      VB_LAUNCH(c, k_grad_forward<weight_t<decltype(w1)>>, grid_1d(npfb), npfb, c->cnt, c->link, c->st.irel, w1,
                w2, c->k11, c->k12, c->k21, c->k22, x, y);
      return VB_OK;
    });
  if (!c->op.valid) {
    vb_status s = build_transpose(c);
    if (s != VB_OK) return s;
  }
  OperatorMap& op = c->op;
  if (!op.ga) {
    VB_HIP(c, hipMalloc((void**)&op.ga, (size_t)npfb * sizeof(double)));
    VB_HIP(c, hipMalloc((void**)&op.gb, (size_t)npfb * sizeof(double)));
  }
  VB_LAUNCH(c, k_grad_coef, grid_1d(npfb), npfb, c->k11, c->k12, c->k21, c->k22, x, op.ga, op.gb);
  return with_weights(c, [&](auto* w1, auto* w2, auto*) -> vb_status {
    VB_LAUNCH(c, k_grad_transpose<weight_t<decltype(w1)>>, grid_1d(npfb), npfb, c->cnt, c->link, c->st.irel, w1,
              w2, op.tcol, op.rslot, op.moff, op.mkey, op.ga, op.gb, y);
    return VB_OK;
  });
}

// v -= mean(v) over n rows: per-workgroup partial sums, then every workgroup of the second kernel
// folds them in the same fixed order (same bits) and subtracts.
__global__ void k_sum_partials(const double* v, int n, double* partials) {
  RowRange rr = block_rows(n, gridDim.x);
  Red<1> R(nullptr, n);
  for (int i = rr.begin + threadIdx.x; i < rr.end; i += blockDim.x) R.add(0, i, v[i]);
  R.store(partials);
}

__global__ void k_sub_mean(double* v, int n, const double* partials, int nblk) {
  const double m = fold(partials, 0, nblk, nullptr, n) / (double)n;
  RowRange rr = block_rows(n, gridDim.x);
  for (int i = rr.begin + threadIdx.x; i < rr.end; i += blockDim.x) v[i] = v[i] - m;
}

vb_status remove_mean(vb_ctx* c, double* v, int n) {
  const int nb = grid_for(n);
  VB_LAUNCH(c, k_sum_partials, nb, v, n, c->partials);
  VB_LAUNCH(c, k_sub_mean, nb, v, n, c->partials, nb);
  return VB_OK;
}

void free_operator(vb_ctx* c) {
  void* ptrs[] = {c->op.tval, c->op.tcol, c->op.mcount, c->op.moff, c->op.mfill, c->op.mkey, c->op.mrow,
                  c->op.mval, c->op.nmiss, c->op.g, c->op.lam, c->op.bv64T, c->op.bv32T, c->op.xtcol, c->op.xtval,
                  c->op.rslot, c->op.ga,
                  c->op.rref, c->op.bref, c->op.dref, c->op.refs,
                  c->op.gb, c->op.gin, c->op.gout};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  c->op = OperatorMap{};
}

}  // namespace vb
