// gfx950 pressure Poisson equation: RHS, assembly and the BiCGSTABs of the reference
// (ppe_solve_in_house_simple_dirichlet.F90:1-158, linear_solver.F90:82-425): SPARSE_PBICGSTAB_nonull
// for closed domains, SPARSE_PBICGSTAB (Jacobi, no DC projection) for open ones, whose free-surface
// rows are Dirichlet (A_ii = 1e30) and whose rows near the surface are scaled by ppe_smooth.
//
// Matrix.  The assembly writes the reference's row layout as ELL, slot-major like the neighbour
// list (col/val[k*npfb + i]) plus a diagonal: slot k of row i is the k-th neighbour of i with
// mirrors folded onto their parents, zero slots included (duplicate columns accumulate into the
// last slot, self-mirrors into the diagonal).  The ~98% of rows with no mirror neighbour (and no
// neighbour further than 32767 rows away) are also stored compact, as blocks of 8 consecutive
// columns (layout at kChunk): one 2-byte start per block, FP64 values or FP32 values in the mixed
// path, and four aligned 16-byte gathers of x per block; the row is summed in increasing column
// order (a rounding-level difference from the reference's slot order).  The other rows are copied
// in list order and taken by workgroups of their own.  Serial mode, whose solve reproduces the
// reference bit for bit, uses the full rows.
// (Recomputing w_L from the positions instead of storing it was measured: compute-bound on the FP64
// divisions and square root, ~1.8x slower.)
//
// Mixed precision: the same BiCGSTAB with FP64 vectors, dot products, scalars and solution update,
// and matvecs that read FP32 values on the compact rows and sum in FP64.  The mixed path's kernel
// weights are FP32 (k_weights), so the compact values -w_L are exact in FP32: the FP32 value planes
// are the assembled matrix itself, the SpMV differs from the FP64 rows' only in rounding, and the
// recurrence residual tracks B - A x as in the FP64 solve.
//
// BiCGSTAB: all scalars (rho, alpha, omega, norms, means) stay on the device; the host only polls the
// convergence flag every few iterations.  See the section comment below for the kernel sequence.
#include <vector>

#include "vb_ctx.h"

namespace vb {

constexpr double kOnePlusEps = 1.0 + 1.0e-15;   // M(i) + 1e-15 with M = 1, linear_solver.F90:55,318
// 1 / (1 + 1e-15): with M = 1 the preconditioned directions p_inv = p / (1 + 1e-15) and s_inv are p and
// s times this constant, so the five-kernel closed-domain iteration keeps no p_inv, s_inv vectors: it
// gathers p and s and applies the constant to the SpMV output and the solution update (the same
// quantities up to rounding; 24 bytes less traffic per row and iteration)
constexpr double kInvOnePlusEps = 1.0 / kOnePlusEps;

// SPMV_TRANSPOSE: y = A^T x through the transposed map (vb_operator.hip), for the adjoint solve;
// SPMV_TRANSPOSE_MIXED: the same with the compact rows' A^T values read as FP32 (mixed path)
enum SpmvMode { SPMV_FP64 = 0, SPMV_MIXED = 1, SPMV_TRANSPOSE = 2, SPMV_TRANSPOSE_MIXED = 3 };

// ------------------------------------------------------------------------------------------------
// RHS, ppe_solve_in_house_simple_dirichlet.F90:27-53 (DC removal for the closed domain only)
// ------------------------------------------------------------------------------------------------
template <class WT>
__global__ void k_ppe_rhs(Geom g, State st, const int* cnt, const int* link, const WT* wg1,
                          const WT* wg2, const double* k11, const double* k12, const double* k21,
                          const double* k22, double* rhs, double* partials, double* red, const Scal* sc) {
  const int npfb = g.npfb;
  RowRange rr = block_rows(npfb, gridDim.x);
  const double dt = sc->dt;
  Red<1> R(red, npfb);
  for (int i = rr.begin + threadIdx.x; i < rr.end; i += blockDim.x) {
    const int n = cnt[i];
    const double ui = st.u[i], vi = st.v[i];
    const double K11 = k11[i], K12 = k12[i], K21 = k21[i], K22 = k22[i];
    double fb = 0.0;
    for_slots(link, i, n, npfb, [&](int j) { return make_double2(st.u[j], st.v[j]); },
              [&](int s, int, double2 uj) {
      double w1 = wg1[s], w2 = wg2[s];
      double uij = ui - uj.x, vij = vi - uj.y;
      double kw1 = K11 * w1 + K12 * w2, kw2 = K21 * w1 + K22 * w2;
      fb = fb - (uij * kw1 + vij * kw2);
    });
    double b = fb / dt;
    rhs[i] = b;
    R.add(0, i, b);
  }
  R.store(partials);
}

__global__ void k_fin_ppe_rhs(const double* partials, int nblk, const double* red, int n, int open, Scal* sc) {
  double s = fold(partials, 0, nblk, red, n);
  if (threadIdx.x == 0) sc->mean = open ? 0.0 : s / (double)n;     // b - 0.0 == b bit for bit
}

// Compact rows are stored as column blocks: the particles are numbered lattice-fashion (one lattice
// column of consecutive indices after another), so a row's neighbours come in runs of consecutive
// indices, one run per lattice column the kernel support crosses (seven at h = 1.3 dx).  A block
// is kBlkCols = 8 consecutive columns starting at an even column s; a row is covered greedily in
// column order (a new block starts at the first column not yet covered), and its block values hold
// A_ij at column s + e, zero where j is not a neighbour or is the row itself (the diagonal is
// applied separately, as before).  One block costs one 2-byte offset and four aligned 16-byte
// gathers of x instead of eight 2-byte offsets and eight 8-byte gathers, and the row is summed in
// the same column order as before with exact zeros between (the same bits up to the sign of a zero
// sum).  At the 1M-particle cylinder every compact row needs 7 or 8 blocks (56 slots for 44
// neighbours).  A row that would need more than kMaxBlk blocks, or whose last block would reach past
// the last row, is stored full instead (rowflag 1).
//
// The blocks are sliced by chunks of 64 rows (a wavefront): every row of chunk c has chunkb[c]
// blocks (its rows' maximum; padding blocks repeat the row's last start with zero values), stored at
// a fixed stride of kMaxBlk blocks per chunk; block b of row i (lane = i % 64):
//   start offset s - i   short4 boff[(c * kOffGroups + b / 4) * 64 + lane], component b % 4
//   FP64 values          double2 bv64[((c * kMaxBlk + b) * 4 + h) * 64 + lane], columns s+2h, s+2h+1
//   FP32 values          float4 bv32[((c * kMaxBlk + b) * 2 + h) * 64 + lane], columns s+4h .. s+4h+3
// so each load of a wavefront is one contiguous 512-byte or 1-KiB run.  bpos[k * npfb + i] records
// where slot k of the neighbour list went (b * 8 + e), for the transposed values.
static_assert(kMaxBlk % 4 == 0 && kMaxBlk * kBlkCols < 256, "block offsets in short4 groups; bpos fits a byte");
__host__ __device__ inline size_t blk64(size_t c, int b, int e, int lane) {     // element of bv64
  return (((c * kMaxBlk + b) * 4 + (e >> 1)) * kChunk + lane) * 2 + (e & 1);
}
__host__ __device__ inline size_t blk32(size_t c, int b, int e, int lane) {     // element of bv32
  return (((c * kMaxBlk + b) * 2 + (e >> 2)) * kChunk + lane) * 4 + (e & 3);
}
__host__ __device__ inline size_t blkoff(size_t c, int b, int lane) {           // element of boff
  return ((c * kOffGroups + (b >> 2)) * kChunk + lane) * 4 + (b & 3);
}
// Rows stored full (mirror neighbours, open-domain surface and scaled rows, rows the blocks do not
// fit) are also kept sliced in list order for the solver's SpMV: entry k of the t-th listed row at
// xcol/xval[((t / 64) * kNplink + k) * 64 + t % 64], so that a wavefront of listed rows reads one
// contiguous run per slot instead of 64 rows kNplink apart.
__host__ __device__ inline size_t xslot(int t, int k) {
  return ((size_t)(t / kChunk) * kNplink + k) * kChunk + (t % kChunk);
}

// Assembly, :60-122.  One thread per row.  Writes the full stored rows, the diagonal and right-hand
// side, and marks the rows the compact SpMV may take (rowflag 0: no mirror neighbour, every
// neighbour within 32767 rows); k_assemble_blocks then writes those rows' blocks.
template <class WT>
__global__ void k_assemble(Geom g, State st, const int* cnt, const int* link, const WT* wl,
                           const int* nsurf, const double* pscale, double* rhs, double* a_diag, int* a_col,
                           double* a_val, uint8_t* rowflag, const Scal* sc) {
  const int i = xcd_row(g.npfb);
  const int npfb = g.npfb;
  if (i < npfb) {
  const int n = cnt[i];
  double bi = rhs[i] - sc->mean;
  double adiag = 0.0;
  bool explicit_row = false;
  for (int k = 0; k < n; k++) {
    int s = k * npfb + i;
    int j = link[s];
    double t = -wl[s];
    adiag = adiag - t;
    if (j >= npfb) {
      explicit_row = true;
      int ij = st.irel[j];
      if (ij == i) adiag = adiag + t;
      bi = bi - st.dP[j] * t;
    } else {
      int off = j - i;
      if (off > 32767 || off < -32767) explicit_row = true;
    }
  }
  if (nsurf[i] != 0) { adiag = 1.0e30; explicit_row = true; }
  // open domain, :126-138: the off-diagonals and right-hand side of rows near the surface are scaled
  // by ppe_smooth (pscale is 1 elsewhere and on surface rows; x * 1.0 == x); scaled rows are stored
  // full (k_assemble_explicit applies the scale), so the compact rows stay unscaled
  const double ps = pscale ? pscale[i] : 1.0;
  if (ps != 1.0) explicit_row = true;
  bi = bi * ps;
  double d = 0.0 + adiag;
  if (!explicit_row) {
    for (int k = 0; k < n; k++) {
      int s = k * npfb + i;
      double val = 0.0 + (-wl[s]);
      a_col[s] = link[s];
      a_val[s] = val;
    }
  }
  // rows with mirror neighbours: k_assemble_explicit writes their slots
  a_diag[i] = d;
  rowflag[i] = explicit_row ? 1 : 0;
  rhs[i] = bi;
  }
}

// Rows stored full (:76-115), one workgroup per row (rows = the explicit rows listed by
// k_list_explicit), one thread per slot.  Slot k's column is the neighbour or the mirror's parent;
// a column's value goes to its last slot and is the sum of -w_L over its slots in slot order
// (earlier slots of the column hold 0); a self-mirror's column is the row itself and its value 0
// (it went into the diagonal).  The columns and -w_L sit in LDS, so each thread's scans read LDS.
constexpr int kExplicitThreads = 128;
static_assert(kExplicitThreads >= kNplink, "one thread per slot");
template <class WT>
__global__ void __launch_bounds__(kExplicitThreads)
k_assemble_explicit(int npfb, const int* rows, int nrows, const int* cnt, const int* link,
                    const WT* wl, const int* irel, const double* pscale, const double* a_diag, int* a_col,
                    double* a_val, int* xcol, double* xval) {
  if ((int)blockIdx.x >= nrows) return;
  __shared__ int cols[kExplicitThreads];
  __shared__ double mw[kExplicitThreads];
  const int i = rows[blockIdx.x], n = cnt[i], k = threadIdx.x;
  int ck = -1;
  if (k < n) {
    const int j = link[k * npfb + i];
    ck = j < npfb ? j : irel[j];
    cols[k] = ck;
    mw[k] = -wl[k * npfb + i];
  }
  __syncthreads();
  if (k < n) {
    double val = 0.0;
    if (ck != i) {
      bool last = true;
      for (int k2 = k + 1; k2 < n && last; k2++) if (cols[k2] == ck) last = false;
      if (last) {
        double acc = 0.0;
        for (int k2 = 0; k2 <= k; k2++) if (cols[k2] == ck) acc = acc + mw[k2];
        val = acc;
        if (pscale) val = val * pscale[i];               // ppe_smooth, :131-134
      }
    }
    a_col[k * npfb + i] = ck;
    a_val[k * npfb + i] = val;
    xcol[xslot(blockIdx.x, k)] = ck;                     // the same row in list order (xslot)
    xval[xslot(blockIdx.x, k)] = val;
  }
}

// The explicit rows in index order: rows[0..nexplicit) from the exclusive scan of rowflag.
__global__ void k_list_explicit(int npfb, const uint8_t* rowflag, const int* excl, int* rows, Scal* sc) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < npfb && rowflag[i]) rows[excl[i]] = i;
  if (i == 0) sc->nexplicit = excl[npfb];
}

// The compact rows' column blocks summed over the chunks (one workgroup; integer sums, any order).
__global__ void k_count_blocks(int nchunks, const int* chunkb, Scal* sc) {
  __shared__ int part[kBlock];
  int s = 0;
  for (int c = threadIdx.x; c < nchunks; c += blockDim.x) s += chunkb[c];
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int k = 0; k < (int)blockDim.x; k++) t += part[k];
    sc->nblocks = t;
  }
}

__global__ void k_flag_to_int(int n, const uint8_t* f, int* out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) out[i] = f[i];
}

// The compact rows' column blocks (layout above).  One thread per row, rows in index order, so a
// wavefront is one chunk; each row's column offsets and their column order sit in LDS.  Pass 1 ranks
// the offsets (a compact row has no repeated column), walks them in column order and counts blocks;
// a row with no neighbour, one needing more than kMaxBlk blocks, or one whose last block would
// reach past row npfb - 1, becomes a full row (rowflag 1).  The wavefront's maximum is the chunk's
// block count.  Pass 2 zero-fills the row's chunkb[c] blocks, writes their starts (padding blocks
// repeat the last) and puts each value at its column.  (~0.2 ms more per step at 1M than the
// four-slot groups it replaced; the SpMVs gain ~25 us each.)
__global__ void __launch_bounds__(kBlock)
k_assemble_blocks(int npfb, const int* cnt, const int* link, const double* a_val, uint8_t* rowflag,
                  int* chunkb, int16_t* boff, double* bv64, float* bv32, uint8_t* bpos) {
  __shared__ int16_t off[kNplink][kBlock];
  __shared__ uint8_t ord[kNplink][kBlock];
  const int t = threadIdx.x, i = blockIdx.x * kBlock + t, lane = t & (kChunk - 1);
  bool live = i < npfb && !rowflag[i];
  const int n = live ? cnt[i] : 0;
  int nb = 0;
  if (live) {
    for (int k = 0; k < n; k++) off[k][t] = (int16_t)(link[k * npfb + i] - i);
    for (int k = 0; k < n; k++) {
      const int16_t ok = off[k][t];
      int pos = 0;
      for (int j = 0; j < n; j++) pos += off[j][t] < ok;
      ord[pos][t] = (uint8_t)k;
    }
    int end = INT_MIN;                             // last column the blocks cover
    for (int q = 0; q < n; q++) {
      const int col = i + off[ord[q][t]][t];
      if (col > end) { nb++; end = (col & ~1) + kBlkCols - 1; }
    }
    if (n == 0 || nb > kMaxBlk || end > npfb - 1) { rowflag[i] = 1; live = false; nb = 0; }
  }
  int nbc = nb;                                    // every lane takes part in the maximum
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) { const int v = __shfl_xor(nbc, o, 64); nbc = v > nbc ? v : nbc; }
  const size_t c = (size_t)(blockIdx.x * kBlock + (t & ~(kChunk - 1))) / kChunk;
  if (lane == 0 && c * kChunk < (size_t)npfb) chunkb[c] = nbc;
  if (!live) return;
  for (int b = 0; b < nbc; b++)
    for (int e = 0; e < kBlkCols; e++) {
      bv64[blk64(c, b, e, lane)] = 0.0;
      if (bv32) bv32[blk32(c, b, e, lane)] = 0.f;
    }
  int end = INT_MIN, b = -1, s = 0;
  for (int q = 0; q < n; q++) {
    const int k = ord[q][t];
    const int col = i + off[k][t];
    if (col > end) {
      s = col & ~1; b++; end = s + kBlkCols - 1;
      boff[blkoff(c, b, lane)] = (int16_t)(s - i);
    }
    const int e = col - s;
    const double v = a_val[(size_t)k * npfb + i];
    bv64[blk64(c, b, e, lane)] = v;
    if (bv32) bv32[blk32(c, b, e, lane)] = (float)v;       // exact: the mixed path's weights are FP32
    bpos[(size_t)k * npfb + i] = (uint8_t)(b * kBlkCols + e);
  }
  for (int bb = b + 1; bb < nbc; bb++) boff[blkoff(c, bb, lane)] = (int16_t)(s - i);
}

// The compact rows of A^T (the adjoint solve): a row without mirror neighbours has the same columns in
// A and A^T (the real neighbour relation is symmetric), so A^T reuses the block starts written above
// and takes a second FP64 value plane, bv64T, filled from the transposed map's real part at the
// positions bpos recorded.
__global__ void k_transpose_blocks(int npfb, const int* cnt, const double* tval, const uint8_t* rowflag,
                                   const int* chunkb, const uint8_t* bpos, double* bv64T, float* bv32T) {
  const int t = threadIdx.x, i = blockIdx.x * kBlock + t, lane = t & (kChunk - 1);
  if (i >= npfb || rowflag[i]) return;
  const size_t c = (size_t)i / kChunk;
  const int nbc = chunkb[c], n = cnt[i];
  for (int b = 0; b < nbc; b++)
    for (int e = 0; e < kBlkCols; e++) {
      bv64T[blk64(c, b, e, lane)] = 0.0;
      if (bv32T) bv32T[blk32(c, b, e, lane)] = 0.f;
    }
  for (int k = 0; k < n; k++) {
    const int p = bpos[(size_t)k * npfb + i];
    const double v = tval[(size_t)k * npfb + i];
    bv64T[blk64(c, p / kBlkCols, p % kBlkCols, lane)] = v;
    if (bv32T) bv32T[blk32(c, p / kBlkCols, p % kBlkCols, lane)] = (float)v;   // exact: FP32 weights
  }
}

// The listed rows of A^T (their transposed map rows, c's slot order) copied in list order, as the
// forward's listed rows are (xslot), for the adjoint solve's SpMV.
__global__ void k_transpose_listed(int npfb, const int* rows, int nrows, const int* cnt, const double* tval,
                                   const int* tcol, int* xtcol, double* xtval) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= nrows) return;
  const int i = rows[t], n = cnt[i];
  for (int k = 0; k < n; k++) {
    xtcol[xslot(t, k)] = tcol[(size_t)k * npfb + i];
    xtval[xslot(t, k)] = tval[(size_t)k * npfb + i];
  }
}

vb_status launch_transpose_blocks(vb_ctx* c) {
  VB_LAUNCH(c, k_transpose_blocks, grid_1d(c->npfb), c->npfb, c->cnt, c->op.tval, c->rowflag, c->chunkb,
            c->bpos, c->op.bv64T, c->op.bv32T);
  const int nx = c->scal_host->nexplicit;       // synchronised by the assembly
  if (nx > 0)
    VB_LAUNCH(c, k_transpose_listed, grid_1d(nx), c->npfb, c->xrows, nx, c->cnt, c->op.tval, c->op.tcol,
              c->op.xtcol, c->op.xtval);
  return VB_OK;
}

// ------------------------------------------------------------------------------------------------
// SpMV rows.  x_j = (src_j - mean) / (1 + 1e-15) reproduces remove_dc + inverseM exactly.
// ------------------------------------------------------------------------------------------------
struct Mat {
  const double *diag, *val;
  const int *col, *len;
  const uint8_t* rowflag;          // null: every row uses the full stored row
  // compact rows as column blocks, sliced by chunks of 64 rows (layout at kChunk)
  const int* chunkb;               // blocks of each chunk's rows
  const short4* boff;              // block start offsets (start column = row + offset, even)
  const double2* bv64;             // FP64 values, four double2 per block
  const float4* bv32;              // FP32 values (mixed path), two float4 per block
  // full rows in list order (xslot): the solver's SpMV takes them in nxb workgroups of their own
  const int* xrows;                // the listed rows
  const int* xcol;
  const double* xval;
  int nx, nxb;                     // listed rows; their workgroups (0: full rows inline)
  // the transposed map (SPMV_TRANSPOSE): real part slot-major like the lists, mirror buckets per column,
  // and the compact rows' A^T values (the forward blocks' starts); the listed rows' A^T rows in list order
  const double2* bv64T;
  const float4* bv32T;             // the same in FP32 (mixed path)
  const int* xtcol;
  const double* xtval;
  const double *tval, *mval;
  const int *tcol, *moff, *mrow;
  int n;
};

// The full stored row, in the reference's slot order (sparse_atime, linear_solver.F90:264-272).
// acc + sum over q in [0, m) of val(q) x[idx(q)], in q order, with eight entries' loads (value and
// index, then the gathers) in flight before they are summed: each entry is otherwise a dependent
// chain of two L2 round trips, and a wall row's mirror bucket holds tens of entries
template <class V, class I>
__device__ inline double gather_sum(double acc, int m, V val, I idx, const double* __restrict__ x) {
  int q = 0;
  for (; q + 8 <= m; q += 8) {
    double v[8], g[8];
    int j[8];
#pragma unroll
    for (int u = 0; u < 8; u++) { v[u] = val(q + u); j[u] = idx(q + u); }
#pragma unroll
    for (int u = 0; u < 8; u++) g[u] = x[j[u]];
#pragma unroll
    for (int u = 0; u < 8; u++) acc = acc + v[u] * g[u];
  }
  for (; q < m; q++) acc = acc + val(q) * x[idx(q)];
  return acc;
}

__device__ inline double row_full(const Mat& A, int i, int m, double xi, const double* __restrict__ x) {
  const size_t n = A.n;
  return gather_sum(A.diag[i] * xi, m, [&](int k) { return A.val[(size_t)k * n + i]; },
                    [&](int k) { return A.col[(size_t)k * n + i]; }, x);
}

// The compact matrix arrays are read once per SpMV: non-temporal loads keep them from displacing the
// gathered vector in L2 (measured on MI355X: a copy kernel with non-temporal accesses reaches 6.2
// TB/s, a plain one 4.6).  The exception is the mixed row's FP32 value stream (row_blk32).
typedef short nt_s4 __attribute__((ext_vector_type(4)));
typedef double nt_d2 __attribute__((ext_vector_type(2)));
typedef float nt_f4 __attribute__((ext_vector_type(4)));
__device__ inline short4 ld_nt(const short4* p) {
  const nt_s4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_s4*>(p));
  return make_short4(v.x, v.y, v.z, v.w);
}
__device__ inline double2 ld_nt(const double2* p) {
  const nt_d2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_d2*>(p));
  return make_double2(v.x, v.y);
}
__device__ inline float4 ld_nt(const float4* p) {
  const nt_f4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_f4*>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}

// The row's block starts: up to kMaxBlk offsets in short4 groups, the groups the chunk uses.
__device__ inline void block_offsets(const Mat& A, size_t c, int lane, int nb, short (&o)[kMaxBlk]) {
#pragma unroll
  for (int g = 0; g < kOffGroups; g++)
    if (4 * g < nb) {
      const short4 t = ld_nt(A.boff + (c * kOffGroups + g) * kChunk + lane);
      o[4 * g] = t.x; o[4 * g + 1] = t.y; o[4 * g + 2] = t.z; o[4 * g + 3] = t.w;
    }
}
typedef double xd2 __attribute__((ext_vector_type(2)));    // two gathered x, 16-byte aligned

// The chunk of row i, as a wavefront-uniform value.  In every caller a wavefront's rows are one
// 64-row chunk (rows come in runs of a multiple of 64 from block_rows or blockIdx * 256), so the
// chunk and its block count can live in scalar registers: the block loop then unrolls with scalar
// exits instead of a per-lane bound, register rotation and a wait for every load at each step.
__device__ inline int uniform_chunk(int i) { return __builtin_amdgcn_readfirstlane(i / kChunk); }

// The block steps of a compact row as straight-line code: step B issues block B+1's loads, then
// sums block B, for B < nb (wavefront-uniform, so every exit is a scalar branch).  Written as a
// template recursion rather than a loop: with a loop the compiler kept a rolled two-block body whose
// back edge rotates the in-flight registers and waits for every outstanding load at each turn
// (59.4 against 54.8 us, mixed, 1M cylinder).  The conditional load of block B+1 joins the sum of
// block B, and at that join the compiler waits for all loads in flight, block B+1's too: each step
// waits about one memory round trip.  Measured at 1M and not kept: separate paths for the last block
// (the wait then covers block B only), the three start groups loaded unconditionally with the row's
// first loads, and the values three blocks ahead: none faster (54.5-55.6 against 54.7 us, mixed;
// the SpMV is bound by its streams' rate, not by one wavefront's latency); loading one block per
// step (64 VGPRs) gave 54.8 / 84.0 us (mixed / FP64) and ~2% slower steps.
template <int B, class Load, class Sum>
__device__ __forceinline__ void block_steps(int nb, Load&& load, Sum&& sum) {
  if constexpr (B < kMaxBlk) {
    if (B >= nb) return;
    if (B + 1 < nb) load(B + 1);
    sum(B);
    block_steps<B + 1>(nb, load, sum);
  }
}

// A compact FP64 row from its blocks (val: bv64 for A, bv64T for the compact rows of A^T):
// acc + sum over blocks and columns, in column order, of a * x (no fused multiply-add, as the FP64
// path everywhere).  Block b+1's values and gathers are issued before block b is summed.
__device__ inline double row_blk64(const Mat& A, const double2* __restrict__ bv, int i,
                                   const double* __restrict__ x, double acc) {
  const size_t c = (size_t)uniform_chunk(i);
  const int lane = i % kChunk, nb = __builtin_amdgcn_readfirstlane(A.chunkb[c]);
  short o[kMaxBlk];
  block_offsets(A, c, lane, nb, o);
  auto vs = [&](int b, int h) { return ((c * kMaxBlk + b) * 4 + h) * kChunk + lane; };
  double2 va[2][4];
  xd2 xg[2][4];
  auto load = [&](int b) {
    const int t = b & 1;
#pragma unroll
    for (int h = 0; h < 4; h++) va[t][h] = ld_nt(bv + vs(b, h));
    const xd2* xn = reinterpret_cast<const xd2*>(x + i + o[b]);
#pragma unroll
    for (int h = 0; h < 4; h++) xg[t][h] = xn[h];
  };
  auto sum = [&](int b) {
    const int t = b & 1;
#pragma unroll
    for (int h = 0; h < 4; h++) {
      acc = acc + va[t][h].x * xg[t][h].x;
      acc = acc + va[t][h].y * xg[t][h].y;
    }
  };
  load(0);
  block_steps<0>(nb, load, sum);
  return acc;
}

// ------------------------------------------------------------------------------------------------
// Mixed path compact rows: y_i = diag_i x_i + sum a32 x_j in FP64, the values read as FP32 (4 bytes
// per block column instead of 8).  The mixed path's kernel weights are FP32, so the values -w_L are
// exact in FP32 and this is the FP64 compact row's sum on the same numbers (with fused multiply-adds).
// Same block steps as the FP64 row.  CACHED (the value plane fits the Infinity Cache,
// kCachedValueBytes): the FP32 values are read with ordinary loads, so that the iteration's two SpMVs
// find them there (1M cylinder, 226 MB: 45.9 against 54.8 us with the non-temporal hint, the same
// with buffer loads of any cache scope; 139.6 against 156 ms per step with the vector kernels' own
// hints, see ld_last); otherwise non-temporally (2M: 277 against 306 ms per step, 4M: 538 against
// 582).  The FP64 row, 482 MB at 1M, is always non-temporal (83 against 93.5 us).
// (Measured alternatives: recomputing the FP32 entries from the pair distance, three FP64 gathers
// per slot, was 1.5x slower than reading the stored values; folding the p and s projections into
// the SpMV, which then folds the partial sums in every one of its workgroups, made the SpMV 1.5x
// slower than the separate projection kernels.)
// ------------------------------------------------------------------------------------------------
// (bv: bv32 for A, bv32T for the compact rows of A^T)
template <bool CACHED = false>
__device__ inline double row_blk32(const Mat& A, const float4* __restrict__ bv, int i,
                                   const double* __restrict__ x, double acc) {
  const size_t c = (size_t)uniform_chunk(i);
  const int lane = i % kChunk, nb = __builtin_amdgcn_readfirstlane(A.chunkb[c]);
  short o[kMaxBlk];
  block_offsets(A, c, lane, nb, o);
  auto vs = [&](int b, int h) { return ((c * kMaxBlk + b) * 2 + h) * kChunk + lane; };
  float4 va[2][2];
  xd2 xg[2][4];
  auto load = [&](int b) {
    const int t = b & 1;
    if constexpr (CACHED) {
      va[t][0] = bv[vs(b, 0)]; va[t][1] = bv[vs(b, 1)];
    } else {
      va[t][0] = ld_nt(bv + vs(b, 0)); va[t][1] = ld_nt(bv + vs(b, 1));
    }
    const xd2* xn = reinterpret_cast<const xd2*>(x + i + o[b]);
#pragma unroll
    for (int h = 0; h < 4; h++) xg[t][h] = xn[h];
  };
  auto sum = [&](int b) {
    const int t = b & 1;
    acc = fma((double)va[t][0].x, xg[t][0].x, acc); acc = fma((double)va[t][0].y, xg[t][0].y, acc);
    acc = fma((double)va[t][0].z, xg[t][1].x, acc); acc = fma((double)va[t][0].w, xg[t][1].y, acc);
    acc = fma((double)va[t][1].x, xg[t][2].x, acc); acc = fma((double)va[t][1].y, xg[t][2].y, acc);
    acc = fma((double)va[t][1].z, xg[t][3].x, acc); acc = fma((double)va[t][1].w, xg[t][3].y, acc);
  };
  load(0);
  block_steps<0>(nb, load, sum);
  return acc;
}

// A full row taken from its list-order copy (t-th listed row i), in the reference's slot order:
// the same sum as row_full.
__device__ inline double row_listed(const Mat& A, int t, int i, const double* __restrict__ x) {
  return gather_sum(A.diag[i] * x[i], A.len[i], [&](int k) { return A.xval[xslot(t, k)]; },
                    [&](int k) { return A.xcol[xslot(t, k)]; }, x);
}

__device__ inline double spmv_row_fp64(const Mat& A, int i, const double* __restrict__ x) {
  const int m = A.len[i];
  const double xi = x[i];
  if (!A.rowflag || A.rowflag[i]) return row_full(A, i, m, xi, x);
  return row_blk64(A, A.bv64, i, x, A.diag[i] * xi);
}

template <bool CACHED = false>
__device__ inline double spmv_row_f32(const Mat& A, int i, const double* __restrict__ x) {
  const double xi = x[i];
  if (A.rowflag[i]) return row_full(A, i, A.len[i], xi, x);
  return row_blk32<CACHED>(A, A.bv32, i, x, A.diag[i] * xi);
}

// Row c of A^T: rows without mirror neighbours as compact FP64 rows (the forward row's blocks, the
// A^T value plane, column order), the others from the transposed map in c's slot order (from its
// list-order copy when t >= 0 is the row's list position); then c's mirror bucket.
template <bool MIXED = false, bool CACHED = false>
__device__ inline double spmv_row_transpose(const Mat& A, int c, const double* __restrict__ x, int t = -1) {
  // the bucket bounds first: their loads overlap the row's chain instead of following it
  const int pb = A.moff[c], pe = A.moff[c + 1];
  double acc;
  if (t >= 0) {
    acc = gather_sum(A.diag[c] * x[c], A.len[c], [&](int k) { return A.xtval[xslot(t, k)]; },
                     [&](int k) { return A.xtcol[xslot(t, k)]; }, x);
  } else if (A.rowflag && !A.rowflag[c]) {
    const double d = A.diag[c] * x[c];
    acc = MIXED ? row_blk32<CACHED>(A, A.bv32T, c, x, d) : row_blk64(A, A.bv64T, c, x, d);
  } else {
    const size_t n = A.n;
    acc = gather_sum(A.diag[c] * x[c], A.len[c], [&](int k) { return A.tval[(size_t)k * n + c]; },
                     [&](int k) { return A.tcol[(size_t)k * n + c]; }, x);
  }
  return gather_sum(acc, pe - pb, [&](int q) { return A.mval[pb + q]; }, [&](int q) { return A.mrow[pb + q]; }, x);
}

template <int MODE, bool CACHED = false>
__device__ inline double spmv_row(const Mat& A, int i, const double* __restrict__ x) {
  if (MODE == SPMV_TRANSPOSE) return spmv_row_transpose(A, i, x);
  if (MODE == SPMV_TRANSPOSE_MIXED) return spmv_row_transpose<true, CACHED>(A, i, x);
  return MODE == SPMV_MIXED ? spmv_row_f32<CACHED>(A, i, x) : spmv_row_fp64(A, i, x);
}

// The rows of one SpMV launch, out(i, (A x)_i) for each.  With A.nxb > 0 the first nxb workgroups
// (a multiple of 8, so the others keep their XCD placement) take the listed full rows from their
// list-order copy, and the other workgroups the compact rows, skipping the listed ones: a full row
// is a chain of ~6 dependent batches of uncoalesced loads, and inside a wavefront of compact rows
// it held the whole wavefront (measured at the 1M cylinder, 2% of the rows: ~19 us of an 80 us
// SpMV).  With A.nxb = 0 every row is taken in place.
template <int MODE, bool CACHED, class Out>
__device__ inline void spmv_rows(const Mat& A, const double* __restrict__ x, Out&& out) {
  if (A.nxb > 0) {
    if ((int)blockIdx.x < A.nxb) {
      for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < A.nx; t += A.nxb * blockDim.x) {
        const int i = A.xrows[t];
        out(i, MODE >= SPMV_TRANSPOSE ? spmv_row_transpose(A, i, x, t) : row_listed(A, t, i, x));
      }
      return;
    }
    const RowRange rr = block_rows(A.n, (int)blockIdx.x - A.nxb, (int)gridDim.x - A.nxb);
    for (int i = rr.begin + threadIdx.x; i < rr.end; i += blockDim.x) {
      if (A.rowflag[i]) continue;
      const double d = A.diag[i] * x[i];
      if (MODE >= SPMV_TRANSPOSE) out(i, spmv_row_transpose<MODE == SPMV_TRANSPOSE_MIXED, CACHED>(A, i, x));
      else out(i, MODE == SPMV_MIXED ? row_blk32<CACHED>(A, A.bv32, i, x, d) : row_blk64(A, A.bv64, i, x, d));
    }
    return;
  }
  const RowRange rr = block_rows(A.n, gridDim.x);
  for (int i = rr.begin + threadIdx.x; i < rr.end; i += blockDim.x) out(i, spmv_row<MODE, CACHED>(A, i, x));
}

// ------------------------------------------------------------------------------------------------
// SPARSE_PBICGSTAB_nonull, linear_solver.F90:275-425, and SPARSE_PBICGSTAB, :82-227 (Vecs::jac set:
// the same kernels with the Jacobi divisor, zero means and the threshold 4).
//
// Per iteration: p -> [p_inv] -> SpMV -> s -> [s_inv] -> SpMV -> x,r -> finalize.  Each row kernel
// that needs a global sum from the kernel before it folds that kernel's block partials itself (every
// block folds them in the same fixed order, so all agree), and only block 0 records the scalar for
// later kernels; a kernel never reads a scalar that it writes.  Stage s writes its partials (and, in
// serial mode, its per-row terms) to its own buffer.  One one-workgroup finalize per iteration
// (after x,r) advances the iteration counters.  These eight kernels are the serial-mode iteration;
// with parallel reductions the solve runs the five-kernel form further down (k_fu_*).
// ------------------------------------------------------------------------------------------------
struct Vecs {
  double *x, *b, *r, *rb, *v, *pn, *pinv, *s, *sinv, *t;   // pn, s: before remove_dc
  // null: SPARSE_PBICGSTAB_nonull (M = 1, remove_dc of p and s); else SPARSE_PBICGSTAB, the open
  // domain's solve (:82-227): Jacobi M = jac = diag(A) (:63-75), no mean removed
  const double* jac;
};

// inverseM, :47-59: x / (M(i) + 1e-15)
__device__ inline double inv_m(const Vecs& w, int i, double x) {
  return w.jac ? x / (w.jac[i] + 1.0e-15) : x / kOnePlusEps;
}
// the same with the solve's kind fixed at compile time (OPEN: w.jac is the diagonal)
template <bool OPEN>
__device__ inline double inv_m_t(const Vecs& w, int i, double x) {
  return OPEN ? x / (w.jac[i] + 1.0e-15) : x / kOnePlusEps;
}

enum { ST_P = 0, ST_SPMV_P = 1, ST_S = 2, ST_SPMV_S = 3, ST_XR = 4, ST_INIT = 5 };
static_assert(ST_INIT < kSolverStages, "stage buffers");

struct RedBufs {
  double* partials;     // kStages * 2 * kPartialStride
  double* rows;         // serial mode: kStages * 2 * n, else null
  int n;
  __device__ double* P(int st) const { return partials + (size_t)st * 2 * kPartialStride; }
  __device__ double* R(int st) const { return rows ? rows + (size_t)st * 2 * n : nullptr; }
};

__device__ inline bool finished(const Scal* sc) { return sc->done || sc->exit_s; }

// Rows a vector kernel loads before it folds the previous kernel's partials (1M rows on 512 x 256
// threads: 8 per thread), so that the fold's latency overlaps the vector loads instead of preceding
// them; rows beyond that are loaded after the fold.
#ifndef VB_PRE
#define VB_PRE 4
#endif
constexpr int kPre = VB_PRE;

// :306-321  AX = A x (x = 0 for the PPE, so r = B); r_bar = r; the terms of the first rho (:330)
// and of sum(r)
__global__ void k_bi_init(Mat A, Vecs w, int x_given, RedBufs rb) {
  const int n = A.n;
  RowRange rr = block_rows(n, gridDim.x);
  Red<2> R(rb.R(ST_INIT), n);
  for (int i = rr.begin + threadIdx.x; i < rr.end; i += blockDim.x) {
    double r;
    if (x_given) r = w.b[i] - spmv_row_fp64(A, i, w.x);
    else { w.x[i] = 0.0; r = w.b[i]; }
    w.r[i] = r; w.rb[i] = r;
    w.v[i] = 0.0; w.pn[i] = 0.0;
    R.add(0, i, r * r);
    R.add(1, i, r);
  }
  R.store(rb.P(ST_INIT));
}

// small: the first-iteration norm threshold, 1.0d-6 (nonull, :367, :404) or the integer expression
// 10-6 = 4 (open, :171, :207)
__global__ void k_fin_init(RedBufs rb, int nblk, double tol, int itermax, double small, Scal* sc) {
  double f[2];
  fold_n<2>(rb.P(ST_INIT), nblk, rb.R(ST_INIT), rb.n, f);
  const double rho = f[0];
  if (threadIdx.x == 0) {
    sc->rho = rho; sc->rho_pp[1] = rho; sc->sum_r = f[1]; sc->rho0 = 1.0; sc->alpha = 1.0; sc->omega = 1.0; sc->beta = 0.0;
    sc->n2s0 = 1.0; sc->n2r0 = 1.0; sc->resid = 0.0; sc->tol = tol; sc->n2_small = small;
    sc->iter = 1; sc->itermax = itermax; sc->exit_s = 0; sc->status = VB_OK;
    sc->mean_p = 0.0; sc->mean_s = 0.0; sc->sum_v = 0.0; sc->sum_p = 0.0;
    sc->done = 0;
    if (itermax < 1) { sc->done = 1; sc->status = VB_WARN_MAXITER; }
    else if (rho == 0.0) { sc->done = 1; sc->status = VB_WARN_BICGSTAB_BREAKDOWN; }
  }
}

// :337-349 search direction p = r + beta (p - omega v), p = pn - mean_p of the last iteration;
// the terms of sum(p) for remove_dc
__global__ void k_bi_p(int n, Vecs w, RedBufs rb, const Scal* sc) {
  if (finished(sc)) return;
  RowRange rr = block_rows(n, gridDim.x);
  const bool first = sc->iter == 1;
  const double beta = sc->beta, omega = sc->omega, mean = sc->mean_p;
  Red<1> R(rb.R(ST_P), n);
  for (int i = rr.begin + threadIdx.x; i < rr.end; i += blockDim.x) {
    double pn = first ? w.r[i] : w.r[i] + beta * ((w.pn[i] - mean) - omega * w.v[i]);
    w.pn[i] = pn;
    R.add(0, i, pn);
  }
  R.store(rb.P(ST_P));
}

// :351-352 remove_dc(p), inverseM: p_inv = (pn - mean)/(M+1e-15) (open domain: mean 0)
__global__ void k_bi_pinv(int n, Vecs w, RedBufs rb, Scal* sc) {
  if (finished(sc)) return;
  RowRange rr = block_rows(n, gridDim.x);
  const int i0 = rr.begin + threadIdx.x, bd = blockDim.x;
  double pv[kPre];
#pragma unroll
  for (int u = 0; u < kPre; u++) if (i0 + u * bd < rr.end) pv[u] = w.pn[i0 + u * bd];
  const double mean = w.jac ? 0.0 : fold(rb.P(ST_P), 0, gridDim.x, rb.R(ST_P), n) / (double)n;   // :243
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->mean_p = mean;
#pragma unroll
  for (int u = 0; u < kPre; u++) if (i0 + u * bd < rr.end) w.pinv[i0 + u * bd] = inv_m(w, i0 + u * bd, pv[u] - mean);
  for (int i = i0 + kPre * bd; i < rr.end; i += bd) w.pinv[i] = inv_m(w, i, w.pn[i] - mean);
}

// :353-354  v = A p_inv; the terms of r_bar . v
// (FUSED: the five-kernel iteration, which also needs the terms of sum(v))
template <int MODE, bool FUSED = false, bool CACHED = false>
__global__ void k_bi_spmv_p(Mat A, Vecs w, RedBufs rb, const Scal* sc) {
  if (finished(sc)) return;
  const int n = A.n;
  Red<FUSED ? 2 : 1> R(rb.R(ST_SPMV_P), n);
  const bool scaled = FUSED && !w.jac;           // closed five-kernel form: A p_inv = c A p
  const double* x = scaled ? w.pn : w.pinv;
  auto out = [&](int i, double bb) {
    if (scaled) bb = bb * kInvOnePlusEps;
    w.v[i] = bb;
    R.add(0, i, w.rb[i] * bb);
    if (FUSED) R.add(FUSED ? 1 : 0, i, bb);
  };
  spmv_rows<MODE, CACHED>(A, x, out);
  R.store(rb.P(ST_SPMV_P));
}

// :356-364  alpha = rho/(r_bar.v + 1e-15); s = r - alpha v; terms of |s|^2 and sum(s)
// (nspmv: workgroups of the SpMV launch whose partials are folded)
__global__ void k_bi_s(int n, Vecs w, RedBufs rb, Scal* sc, int nspmv) {
  if (finished(sc)) return;
  RowRange rr = block_rows(n, gridDim.x);
  const int i0 = rr.begin + threadIdx.x, bd = blockDim.x;
  double rv[kPre], vv[kPre];
#pragma unroll
  for (int u = 0; u < kPre; u++)
    if (i0 + u * bd < rr.end) { rv[u] = w.r[i0 + u * bd]; vv[u] = w.v[i0 + u * bd]; }
  const double temp = fold(rb.P(ST_SPMV_P), 0, nspmv, rb.R(ST_SPMV_P), n);
  const double alpha = sc->rho / (temp + 1.0e-15);
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->alpha = alpha;
  Red<2> R(rb.R(ST_S), n);
  auto row = [&](int i, double r, double v) {
    double s = r - alpha * v;
    w.s[i] = s;
    R.add(0, i, s * s);
    R.add(1, i, s);
  };
#pragma unroll
  for (int u = 0; u < kPre; u++) if (i0 + u * bd < rr.end) row(i0 + u * bd, rv[u], vv[u]);
  for (int i = i0 + kPre * bd; i < rr.end; i += bd) row(i, w.r[i], w.v[i]);
  R.store(rb.P(ST_S));
}

// :366-384  early exit on |s| (x += alpha p_inv), else remove_dc(s), inverseM: s_inv
__global__ void k_bi_sinv(int n, Vecs w, RedBufs rb, Scal* sc) {
  if (finished(sc)) return;
  RowRange rr = block_rows(n, gridDim.x);
  const int i0 = rr.begin + threadIdx.x, bd = blockDim.x;
  double sv[kPre];
#pragma unroll
  for (int u = 0; u < kPre; u++) if (i0 + u * bd < rr.end) sv[u] = w.s[i0 + u * bd];
  double f[2];
  fold_n<2>(rb.P(ST_S), gridDim.x, rb.R(ST_S), n, f);
  const double n2s = f[0], sum = f[1];
  const bool first = sc->iter == 1;
  const double n2s0 = first ? ((n2s <= sc->n2_small) ? 1.0 : n2s) : sc->n2s0;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (n2s <= sc->tol * sc->tol * n2s0) {
    const double alpha = sc->alpha;
    for (int i = i0; i < rr.end; i += bd) w.x[i] = w.x[i] + alpha * w.pinv[i];
    if (lead) { sc->n2s0 = n2s0; sc->resid = sqrt(n2s / n2s0); sc->exit_s = 1; }
    return;
  }
  const double mean = w.jac ? 0.0 : sum / (double)n;
  if (lead) { sc->n2s0 = n2s0; sc->mean_s = mean; }
#pragma unroll
  for (int u = 0; u < kPre; u++) if (i0 + u * bd < rr.end) w.sinv[i0 + u * bd] = inv_m(w, i0 + u * bd, sv[u] - mean);
  for (int i = i0 + kPre * bd; i < rr.end; i += bd) w.sinv[i] = inv_m(w, i, w.s[i] - mean);
}

// :385-387  t = A s_inv; terms of t.s and t.t with the projected s
// (FUSED: the five-kernel iteration stores s already projected)
template <int MODE, bool FUSED = false, bool CACHED = false>
__global__ void k_bi_spmv_s(Mat A, Vecs w, RedBufs rb, const Scal* sc) {
  if (finished(sc)) return;
  const int n = A.n;
  const double mean = FUSED ? 0.0 : sc->mean_s;
  Red<2> R(rb.R(ST_SPMV_S), n);
  const bool scaled = FUSED && !w.jac;           // closed five-kernel form: A s_inv = c A s
  const double* x = scaled ? w.s : w.sinv;
  auto out = [&](int i, double bb) {
    if (scaled) bb = bb * kInvOnePlusEps;
    w.t[i] = bb;
    double si = w.s[i] - mean;
    R.add(0, i, bb * si);
    R.add(1, i, bb * bb);
  };
  spmv_rows<MODE, CACHED>(A, x, out);
  R.store(rb.P(ST_SPMV_S));
}

// :389-401  omega; x = x + alpha p_inv + omega s_inv; r = s - omega t; terms of |r|^2 and the next
// r_bar . r
__global__ void k_bi_xr(int n, Vecs w, RedBufs rb, Scal* sc, int nspmv) {
  if (finished(sc)) return;
  RowRange rr = block_rows(n, gridDim.x);
  const int i0 = rr.begin + threadIdx.x, bd = blockDim.x;
  double xv[kPre], pv[kPre], qv[kPre], sv[kPre], tv[kPre], bv[kPre];
#pragma unroll
  for (int u = 0; u < kPre; u++) {
    const int i = i0 + u * bd;
    if (i < rr.end) { xv[u] = w.x[i]; pv[u] = w.pinv[i]; qv[u] = w.sinv[i]; sv[u] = w.s[i]; tv[u] = w.t[i]; bv[u] = w.rb[i]; }
  }
  double f[2];
  fold_n<2>(rb.P(ST_SPMV_S), nspmv, rb.R(ST_SPMV_S), n, f);
  const double ts = f[0], tt = f[1];
  const double omega = ts / (tt + 1.0e-15);
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->omega = omega;
  const double alpha = sc->alpha, ms = sc->mean_s;
  Red<2> R(rb.R(ST_XR), n);
  auto row = [&](int i, double x, double pinv, double sinv, double s, double t, double rbi) {
    w.x[i] = x + alpha * pinv + omega * sinv;
    double r = (s - ms) - omega * t;
    w.r[i] = r;
    R.add(0, i, r * r);
    R.add(1, i, rbi * r);
  };
#pragma unroll
  for (int u = 0; u < kPre; u++)
    if (i0 + u * bd < rr.end) row(i0 + u * bd, xv[u], pv[u], qv[u], sv[u], tv[u], bv[u]);
  for (int i = i0 + kPre * bd; i < rr.end; i += bd) row(i, w.x[i], w.pinv[i], w.sinv[i], w.s[i], w.t[i], w.rb[i]);
  R.store(rb.P(ST_XR));
}

// :403-419 convergence; next rho (:330) and beta (:342-343)
__global__ void k_fin_xr(RedBufs rb, int nblk, Scal* sc) {
  if (finished(sc)) return;
  double f[2];
  fold_n<2>(rb.P(ST_XR), nblk, rb.R(ST_XR), rb.n, f);
  const double n2r = f[0], rho = f[1];
  if (threadIdx.x == 0) {
    if (sc->iter == 1) sc->n2r0 = (n2r <= sc->n2_small) ? 1.0 : n2r;
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

// ------------------------------------------------------------------------------------------------
// The same iteration in five kernels (parallel reductions only; serial mode keeps the kernels above,
// which reproduce the reference's operation order):
//
//   k_fu_p      [finish iteration k-1: convergence, rho, beta]  p = r + beta (p - omega v),
//               p -= mean(p)                                                   -> sum(p) terms
//   k_bi_spmv_p v = A p_inv = c A p                                            -> r_bar.v, sum(v)
//   k_fu_s      alpha; s = r - alpha v, s -= mean(s)                           -> |s|^2 (before
//               remove_dc, as the reference's early-exit test)
//   k_bi_spmv_s t = A s_inv = c A s                                            -> t.s, t.t
//   k_fu_xr     early exit on |s|, omega; x += alpha c p + omega c s; r = s - omega t
//                                                                              -> |r|^2, r_bar.r, sum(r)
//
// (c = 1 / (1 + 1e-15), the closed domain's inverseM; the open domain's Jacobi directions
// p_inv = p / (diag + 1e-15) and s_inv are stored by k_fu_p and k_fu_s and read in their place.)
//
// The two remove_dc means are not folded from the new vectors (that takes a kernel boundary each)
// but follow from sums already folded: sum(p_k) = sum(r) + beta (sum(p_{k-1}) - omega sum(v)) and
// sum(s) = sum(r) - alpha sum(v), with sum(p_{k-1}) the terms of the projected p written one
// iteration earlier.  They equal the folded means up to rounding of the sums, so the projections
// remove the constant component as the reference's do, and the iterates agree with the
// eight-kernel ones to rounding (not bit for bit).
//
// Scalars: a kernel still never reads a scalar that it writes.  k_fu_p both finishes iteration k-1
// (writing rho_k) and uses rho_{k-1}, so rho is kept by iteration parity in rho_pp; the iteration
// number comes from the host (while the solve runs, the device's counter equals it).  The last
// iteration of each polling batch is finished by the one-workgroup k_fu_fin, and the first k_fu_p
// of the next batch then reads beta and sum(r) from the scalars.  The sum(p) terms go to the second
// quantity of the s stage, so k_fu_xr folds them with |s|^2 in one pass.
// ------------------------------------------------------------------------------------------------
static_assert(3 * kMaxBlocks <= 2 * kPartialStride, "three partial quantities per vector-kernel stage");

struct FinState { bool done; double beta, sum_r; };

// :403-419, :329-343 for iteration k from its folded |r|^2, r_bar.r and sum(r); lead records them
__device__ inline FinState finish_iteration(const double (&f)[3], int k, Scal* sc, bool lead) {
  const double n2r = f[0], rho = f[1];
  const double n2r0 = k == 1 ? ((n2r <= sc->n2_small) ? 1.0 : n2r) : sc->n2r0;
  FinState o{false, 0.0, f[2]};
  int status = VB_OK;
  const bool converged = n2r <= sc->tol * sc->tol * n2r0;
  const double rho0 = sc->rho_pp[k & 1];
  if (converged) o.done = true;
  else if (k + 1 > sc->itermax) { o.done = true; status = VB_WARN_MAXITER; }
  else if (rho == 0.0) { o.done = true; status = VB_WARN_BICGSTAB_BREAKDOWN; }
  else o.beta = rho / (rho0 + 1.0e-15) * sc->alpha / (sc->omega + 1.0e-15);
  if (lead) {
    if (k == 1) sc->n2r0 = n2r0;
    sc->resid = sqrt(n2r / n2r0);
    if (!converged) { sc->iter = k + 1; sc->rho0 = rho0; sc->rho = rho; sc->rho_pp[(k + 1) & 1] = rho; }
    sc->beta = o.beta; sc->sum_r = o.sum_r;
    if (o.done) { sc->done = 1; sc->status = status; }
  }
  return o;
}

// A vector element read for the last time in the iteration (NT: non-temporal) and the solution's
// update (NT: non-temporal store: x is read again only in the next iteration's last kernel).  NT
// when the mixed SpMVs read their FP32 values cacheably (solve's CACHED): streaming the vectors that
// are not read again keeps more of the matrix and of the next kernels' vectors in the caches
// (measured at the 1M cylinder: 141.3 against 135.3 ms per mixed step; where the matrix is read
// non-temporally, FP64 at 1M: 214.2 against 215.0 ms, and mixed at 2M: 277 against 287 ms, the
// vectors keep ordinary accesses).  With NT the vector kernels also load 2 * kPre rows per thread
// before their fold, all of a thread's rows at 1M (132.1 against 135.2 ms per mixed step; FP64 215.6
// against 213.8).  256-thread workgroups: the launch bound lets k_fu_xr hold its 16 rows' operands
// without spilling.
template <bool NT> __device__ __forceinline__ double ld_last(const double* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT> __device__ __forceinline__ void st_last(double v, double* p) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// Iteration k; fin = 1: finish iteration k-1 here (its x,r partials), 0: k_fu_fin or k_fin_init did
// (OPEN: the open domain's SPARSE_PBICGSTAB, Jacobi and no mean removed; NT: see ld_last)
template <bool OPEN, bool NT>
__global__ void __launch_bounds__(kBlock) k_fu_p(int n, Vecs w, RedBufs rb, Scal* sc, int k, int fin) {
  constexpr int P = NT ? 2 * kPre : kPre;   // rows loaded before the fold, see ld_last
  if (finished(sc)) return;
  RowRange rr = block_rows(n, gridDim.x);
  const int i0 = rr.begin + threadIdx.x, bd = blockDim.x;
  double rv[P], pv[P], vv[P];
#pragma unroll
  for (int u = 0; u < P; u++) {
    const int i = i0 + u * bd;
    if (i < rr.end) { rv[u] = w.r[i]; pv[u] = ld_last<NT>(w.pn + i); vv[u] = ld_last<NT>(w.v + i); }
  }
  double beta, sum_r;
  if (fin) {
    double f[3];
    fold_n<3>(rb.P(ST_XR), gridDim.x, nullptr, n, f);
    const FinState o = finish_iteration(f, k - 1, sc, blockIdx.x == 0 && threadIdx.x == 0);
    if (o.done) return;
    beta = o.beta; sum_r = o.sum_r;
  } else {
    beta = sc->beta; sum_r = sc->sum_r;
  }
  const bool first = k == 1;
  const double omega = sc->omega;
  const double sum_p = first ? sum_r : sum_r + beta * (sc->sum_p - omega * sc->sum_v);
  const double mean = OPEN ? 0.0 : sum_p / (double)n;                    // remove_dc, :243
  Red<1> R(nullptr, n);
  auto row = [&](int i, double r, double p, double v) {
    const double pn = first ? r : r + beta * (p - omega * v);
    const double pp = pn - mean;
    w.pn[i] = pp;
    if (OPEN) w.pinv[i] = inv_m_t<OPEN>(w, i, pp);   // closed: p_inv = c p is never stored
    R.add(0, i, pp);
  };
#pragma unroll
  for (int u = 0; u < P; u++) if (i0 + u * bd < rr.end) row(i0 + u * bd, rv[u], pv[u], vv[u]);
  for (int i = i0 + P * bd; i < rr.end; i += bd) row(i, w.r[i], ld_last<NT>(w.pn + i), ld_last<NT>(w.v + i));
  R.store(rb.P(ST_S) + gridDim.x);
}

template <bool OPEN, bool NT>
__global__ void __launch_bounds__(kBlock) k_fu_s(int n, Vecs w, RedBufs rb, Scal* sc, int nspmv, int k) {
  constexpr int P = NT ? 2 * kPre : kPre;   // rows loaded before the fold, see ld_last
  if (finished(sc)) return;
  RowRange rr = block_rows(n, gridDim.x);
  const int i0 = rr.begin + threadIdx.x, bd = blockDim.x;
  double rv[P], vv[P];
#pragma unroll
  for (int u = 0; u < P; u++)
    if (i0 + u * bd < rr.end) { rv[u] = ld_last<NT>(w.r + i0 + u * bd); vv[u] = ld_last<NT>(w.v + i0 + u * bd); }
  double f[2];
  fold_n<2>(rb.P(ST_SPMV_P), nspmv, nullptr, n, f);
  const double alpha = sc->rho_pp[k & 1] / (f[0] + 1.0e-15);
  const double mean = OPEN ? 0.0 : (sc->sum_r - alpha * f[1]) / (double)n;
  if (blockIdx.x == 0 && threadIdx.x == 0) { sc->alpha = alpha; sc->sum_v = f[1]; sc->mean_s = mean; }
  Red<1> R(nullptr, n);
  auto row = [&](int i, double r, double v) {
    const double s = r - alpha * v;
    R.add(0, i, s * s);
    const double sp = s - mean;
    w.s[i] = sp;
    if (OPEN) w.sinv[i] = inv_m_t<OPEN>(w, i, sp);   // closed: s_inv = c s is never stored
  };
#pragma unroll
  for (int u = 0; u < P; u++) if (i0 + u * bd < rr.end) row(i0 + u * bd, rv[u], vv[u]);
  for (int i = i0 + P * bd; i < rr.end; i += bd) row(i, ld_last<NT>(w.r + i), ld_last<NT>(w.v + i));
  R.store(rb.P(ST_S));
}

// (OPEN: the directions p_inv, s_inv are stored; closed: they are c p and c s, c = 1 / (1 + 1e-15))
template <bool OPEN, bool NT>
__global__ void __launch_bounds__(kBlock) k_fu_xr(int n, Vecs w, RedBufs rb, Scal* sc, int nspmv, int k) {
  constexpr int P = NT ? 2 * kPre : kPre;   // rows loaded before the fold, see ld_last
  if (finished(sc)) return;
  RowRange rr = block_rows(n, gridDim.x);
  const int i0 = rr.begin + threadIdx.x, bd = blockDim.x;
  const double* pd = OPEN ? w.pinv : w.pn;
  double xv[P], pv[P], qv[P], sv[P], tv[P], bv[P];
#pragma unroll
  for (int u = 0; u < P; u++) {
    const int i = i0 + u * bd;
    if (i < rr.end) {
      xv[u] = ld_last<NT>(w.x + i); pv[u] = ld_last<NT>(pd + i); sv[u] = ld_last<NT>(w.s + i);
      tv[u] = ld_last<NT>(w.t + i); bv[u] = w.rb[i];
      qv[u] = OPEN ? w.sinv[i] : sv[u];
    }
  }
  double fs[2], ft[2];
  fold_n<2>(rb.P(ST_S), gridDim.x, nullptr, n, fs);          // |s|^2, sum(p)
  fold_n<2>(rb.P(ST_SPMV_S), nspmv, nullptr, n, ft);         // t.s, t.t
  const double n2s = fs[0];
  const double n2s0 = k == 1 ? ((n2s <= sc->n2_small) ? 1.0 : n2s) : sc->n2s0;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  const double alpha = sc->alpha;
  const double ca = OPEN ? alpha : alpha * kInvOnePlusEps;   // the factor of the stored direction
  if (lead) {
    if (k == 1) sc->n2s0 = n2s0;
    sc->sum_p = fs[1];
  }
  if (n2s <= sc->tol * sc->tol * n2s0) {                     // :374-381
    for (int i = i0; i < rr.end; i += bd) st_last<NT>(ld_last<NT>(w.x + i) + ca * ld_last<NT>(pd + i), w.x + i);
    if (lead) { sc->resid = sqrt(n2s / n2s0); sc->exit_s = 1; }
    return;
  }
  const double omega = ft[0] / (ft[1] + 1.0e-15);
  const double co = OPEN ? omega : omega * kInvOnePlusEps;
  if (lead) sc->omega = omega;
  Red<3> R(nullptr, n);
  auto row = [&](int i, double x, double pdir, double sdir, double s, double t, double rbi) {
    st_last<NT>(x + ca * pdir + co * sdir, w.x + i);
    const double r = s - omega * t;
    w.r[i] = r;
    R.add(0, i, r * r);
    R.add(1, i, rbi * r);
    R.add(2, i, r);
  };
#pragma unroll
  for (int u = 0; u < P; u++)
    if (i0 + u * bd < rr.end) row(i0 + u * bd, xv[u], pv[u], qv[u], sv[u], tv[u], bv[u]);
  for (int i = i0 + P * bd; i < rr.end; i += bd)
    row(i, ld_last<NT>(w.x + i), ld_last<NT>(pd + i), OPEN ? w.sinv[i] : ld_last<NT>(w.s + i), ld_last<NT>(w.s + i),
        ld_last<NT>(w.t + i), w.rb[i]);
  R.store(rb.P(ST_XR));
}

// finishes iteration k at the end of a polling batch (one workgroup)
__global__ void k_fu_fin(RedBufs rb, int nblk, Scal* sc, int k) {
  if (finished(sc)) return;
  double f[3];
  fold_n<3>(rb.P(ST_XR), nblk, nullptr, rb.n, f);
  finish_iteration(f, k, sc, threadIdx.x == 0);
}

#define VB_FIN(c, kern, ...)                                                               \
  do {                                                                                     \
    hipLaunchKernelGGL(kern, dim3(1), dim3(kBlock), 0, (c)->stream, __VA_ARGS__);          \
    VB_HIP(c, hipGetLastError());                                                          \
  } while (0)

static Mat step_matrix(vb_ctx* c) {
  Mat A;
  A.diag = c->a_diag; A.val = c->a_val; A.col = c->a_col; A.len = c->cnt;
  A.rowflag = c->rowflag; A.chunkb = c->chunkb; A.boff = (const short4*)c->boff;
  A.bv64 = (const double2*)c->bv64; A.bv32 = (const float4*)c->bv32; A.n = c->npfb;
  A.xrows = c->xrows; A.xcol = c->xcol; A.xval = c->xval; A.nx = 0; A.nxb = 0;
  A.tval = c->op.tval; A.mval = c->op.mval; A.tcol = c->op.tcol; A.moff = c->op.moff; A.mrow = c->op.mrow;
  A.bv64T = (const double2*)c->op.bv64T; A.bv32T = (const float4*)c->op.bv32T;
  A.xtcol = c->op.xtcol; A.xtval = c->op.xtval;
  return A;
}

// The solvers' SpMVs (forward and adjoint) take the assembled step matrix's listed full rows in
// workgroups of their own (spmv_rows); nexplicit is the count the assembly synchronised.
static void split_listed_rows(vb_ctx* c, Mat& A) {
  if (!A.rowflag) return;
  A.nx = c->scal_host->nexplicit;
  int b = (A.nx + kSpmvBlock - 1) / kSpmvBlock;
  if (b > kSpmvListBlocks) b = kSpmvListBlocks;
  A.nxb = (b + 7) / 8 * 8;
}

static RedBufs red_bufs(vb_ctx* c, int n) { return RedBufs{c->partials, c->red, n}; }

template <int MODE, bool CACHED = false>
static vb_status solve(vb_ctx* c, const Mat& A, Vecs w, int x_given, double tol, int itermax) {
  const int n = A.n;
  const int nb = grid_for(n);          // streaming vector kernels
  const int ns = spmv_grid(n) + A.nxb; // SpMV kernels, one row per thread (+ the listed rows' workgroups)
  RedBufs rb = red_bufs(c, n);
  Scal* sc = c->scal;
  const bool fused = rb.rows == nullptr;   // serial mode: the reference's eight-kernel order
  constexpr bool NT = CACHED;            // see ld_last
  const double small = w.jac ? (double)(10 - 6) : 1.0e-6;
  VB_LAUNCH(c, k_bi_init, nb, A, w, x_given, rb);
  VB_FIN(c, k_fin_init, rb, nb, tol, itermax, small, sc);
  int launched = 0, batch = 8;
  for (;;) {
    for (int b = 0; b < batch; b++) {
      const int k = launched + b + 1;
      if (fused) {
        if (w.jac) VB_LAUNCH(c, (k_fu_p<true, NT>), nb, n, w, rb, sc, k, b > 0 ? 1 : 0);
        else VB_LAUNCH(c, (k_fu_p<false, NT>), nb, n, w, rb, sc, k, b > 0 ? 1 : 0);
        VB_LAUNCH_BS(c, (k_bi_spmv_p<MODE, true, CACHED>), ns, kSpmvBlock, A, w, rb, sc);
        if (w.jac) VB_LAUNCH(c, (k_fu_s<true, NT>), nb, n, w, rb, sc, ns, k);
        else VB_LAUNCH(c, (k_fu_s<false, NT>), nb, n, w, rb, sc, ns, k);
        VB_LAUNCH_BS(c, (k_bi_spmv_s<MODE, true, CACHED>), ns, kSpmvBlock, A, w, rb, sc);
        if (w.jac) VB_LAUNCH(c, (k_fu_xr<true, NT>), nb, n, w, rb, sc, ns, k);
        else VB_LAUNCH(c, (k_fu_xr<false, NT>), nb, n, w, rb, sc, ns, k);
        continue;
      }
      VB_LAUNCH(c, k_bi_p, nb, n, w, rb, sc);
      VB_LAUNCH(c, k_bi_pinv, nb, n, w, rb, sc);
      VB_LAUNCH_BS(c, k_bi_spmv_p<MODE>, ns, kSpmvBlock, A, w, rb, sc);
      VB_LAUNCH(c, k_bi_s, nb, n, w, rb, sc, ns);
      VB_LAUNCH(c, k_bi_sinv, nb, n, w, rb, sc);
      VB_LAUNCH_BS(c, k_bi_spmv_s<MODE>, ns, kSpmvBlock, A, w, rb, sc);
      VB_LAUNCH(c, k_bi_xr, nb, n, w, rb, sc, ns);
      VB_FIN(c, k_fin_xr, rb, nb, sc);
    }
    if (fused) VB_FIN(c, k_fu_fin, rb, nb, sc, launched + batch);
    launched += batch;
    VB_HIP(c, hipMemcpyAsync(c->scal_host, c->scal, sizeof(Scal), hipMemcpyDeviceToHost, c->stream));
    VB_HIP(c, hipStreamSynchronize(c->stream));
    if (c->scal_host->done || c->scal_host->exit_s) break;
    if (launched > itermax + 2) break;
    batch = launched < 32 ? 8 : 32;
  }
  return VB_OK;
}

// open: the open domain's Jacobi solve on the assembled diagonal
static Vecs step_vecs(vb_ctx* c, double* x, bool open) {
  return Vecs{x, c->rhs, c->r, c->rb, c->vv, c->pn, c->pinv, c->ss, c->sinv, c->tt, open ? c->a_diag : nullptr};
}

// Right-hand side and assembly (:27-122) without the solve: the full and compact rows, the explicit
// rows, and the transposed map marked stale.
vb_status assemble_ppe(vb_ctx* c) {
  const int nb = spmv_grid(c->npfb);    // a neighbour sum: one row per thread, like the SpMV
  vb_status sw = with_weights(c, [&](auto* w1, auto* w2, auto*) -> vb_status {
    VB_LAUNCH_BS(c, k_ppe_rhs<weight_t<decltype(w1)>>, nb, kSpmvBlock, c->g, c->st, c->cnt, c->link, w1, w2,
                 c->k11, c->k12, c->k21, c->k22, c->rhs, c->partials, c->red, c->scal);
    return VB_OK;
  });
  if (sw != VB_OK) return sw;
  VB_FIN(c, k_fin_ppe_rhs, c->partials, nb, c->red, c->npfb, c->g.open, c->scal);
  sw = with_weights(c, [&](auto*, auto*, auto* wl) -> vb_status {
    VB_LAUNCH(c, k_assemble<weight_t<decltype(wl)>>, xcd_grid(c->npfb), c->g, c->st, c->cnt, c->link, wl,
              c->nsurf, c->pscale, c->rhs, c->a_diag, c->a_col, c->a_val, c->rowflag, c->scal);
    return VB_OK;
  });
  if (sw != VB_OK) return sw;
  VB_LAUNCH(c, k_assemble_blocks, grid_1d(c->npfb), c->npfb, c->cnt, c->link, c->a_val, c->rowflag,
            c->chunkb, c->boff, c->bv64, c->prm.precision == VB_PREC_MIXED ? c->bv32 : nullptr, c->bpos);
  // the rows stored full (rowflag 1): listed, counted (one synchronisation), one workgroup each
  VB_LAUNCH(c, k_flag_to_int, grid_1d(c->npfb), c->npfb, c->rowflag, c->xflag);
  vb_status ss;
  if ((ss = launch_exclusive_scan(c, c->xflag, c->xscan, c->npfb)) != VB_OK) return ss;
  VB_LAUNCH(c, k_list_explicit, grid_1d(c->npfb), c->npfb, c->rowflag, c->xscan, c->xrows, c->scal);
  hipLaunchKernelGGL(k_count_blocks, dim3(1), dim3(kBlock), 0, c->stream, (int)ppe_chunks(c->npfb), c->chunkb,
                     c->scal);
  VB_HIP(c, hipGetLastError());
  if ((ss = sync_scalars(c)) != VB_OK) return ss;
  // the mixed SpMVs' FP32 value plane as read: (blocks) x (64 lanes) x (8 columns) x 4 bytes
  c->vcache = c->prm.precision == VB_PREC_MIXED && !c->spmv_stream &&
              (double)c->scal_host->nblocks * kChunk * kBlkCols * sizeof(float) <= kCachedValueBytes;
  const int nx = c->scal_host->nexplicit;
  if (nx > 0) {
    sw = with_weights(c, [&](auto*, auto*, auto* wl) -> vb_status {
      hipLaunchKernelGGL(k_assemble_explicit<weight_t<decltype(wl)>>, dim3(nx), dim3(kExplicitThreads), 0,
                         c->stream, c->npfb, c->xrows, nx, c->cnt, c->link, wl, c->st.irel, c->pscale, c->a_diag,
                         c->a_col, c->a_val, c->xcol, c->xval);
      VB_HIP(c, hipGetLastError());
      return VB_OK;
    });
    if (sw != VB_OK) return sw;
  }
  c->op.valid = false; c->op.step_matrix = true;        // a new matrix: the transposed map is rebuilt on use
  return VB_OK;
}

vb_status launch_ppe(vb_ctx* c, double tol, int itermax) {
  vb_status sa = assemble_ppe(c);
  if (sa != VB_OK) return sa;
  Mat A = step_matrix(c);
  Vecs w = step_vecs(c, c->st.p, c->g.open != 0);
  // FP64 from the stored rows in the reference's slot order: on request, and in serial mode
  if ((c->spmv_stored || c->red) && c->prm.precision != VB_PREC_MIXED) A.rowflag = nullptr;
  if (!c->red) split_listed_rows(c, A);
  if (c->prm.precision == VB_PREC_MIXED)
    return c->vcache ? solve<SPMV_MIXED, true>(c, A, w, 0, tol, itermax) : solve<SPMV_MIXED>(c, A, w, 0, tol, itermax);
  return solve<SPMV_FP64>(c, A, w, 0, tol, itermax);
}

// SPARSE_PBICGSTAB_nonull (open = false) or SPARSE_PBICGSTAB (open = true) on a caller's system in
// Numerical-Recipes storage (host arrays).
vb_status solve_nr(vb_ctx* c, const double* A_nr, const int* ija, double* x, const double* B, int n,
                   int itermax, double tol, bool open) {
  if (n != c->npfb) { c->err = "system size must equal the context's npfb"; return VB_ERR_ARG; }
  c->op.valid = false; c->op.step_matrix = false;       // the stored matrix becomes the caller's
  std::vector<int> len(n), col((size_t)n * kNplink, 0);
  std::vector<double> val((size_t)n * kNplink, 0.0), diag(n);
  for (int i = 0; i < n; i++) {
    diag[i] = A_nr[i];
    int b = ija[i], e = ija[i + 1];
    int m = e - b;
    if (m < 0 || m > kNplink) { c->err = "row longer than VB_NPLINK slots"; return VB_ERR_ARG; }
    len[i] = m;
    for (int k = 0; k < m; k++) {
      int col1 = ija[b - 1 + k];
      if (col1 < 1 || col1 > n) { c->err = "column index out of range"; return VB_ERR_ARG; }
      col[(size_t)k * n + i] = col1 - 1;
      val[(size_t)k * n + i] = A_nr[b - 1 + k];
    }
  }
  const size_t nw = (size_t)n * kNplink;
  VB_HIP(c, hipMemcpyAsync(c->a_len, len.data(), (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
  VB_HIP(c, hipMemcpyAsync(c->a_col, col.data(), nw * 4, hipMemcpyHostToDevice, c->stream));
  VB_HIP(c, hipMemcpyAsync(c->a_val, val.data(), nw * 8, hipMemcpyHostToDevice, c->stream));
  VB_HIP(c, hipMemcpyAsync(c->a_diag, diag.data(), (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
  VB_HIP(c, hipMemcpyAsync(c->rhs, B, (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
  VB_HIP(c, hipMemcpyAsync(c->xsys, x, (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
  Mat A = step_matrix(c);
  A.len = c->a_len;
  A.rowflag = nullptr;
  vb_status s = solve<SPMV_FP64>(c, A, step_vecs(c, c->xsys, open), 1, tol, itermax);
  if (s != VB_OK) return s;
  VB_HIP(c, hipMemcpyAsync(x, c->xsys, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
  VB_HIP(c, hipStreamSynchronize(c->stream));
  return VB_OK;
}

// The adjoint PPE solve A^T lambda = g on the matrix the last step assembled, by FP64 iterative
// refinement around the reference's own BiCGSTAB (SPARSE_PBICGSTAB_nonull for closed domains,
// SPARSE_PBICGSTAB with Jacobi M = diag(A^T) = diag(A) for open ones) with the transposed map as its
// matrix.  g is read from and lambda written to the operator's device vectors op.g and op.lam.
//
// Why refinement.  In a wall-bounded domain A is not symmetric (wall particles see the mirrors of
// fluid particles but have none of their own), so the left null vector of A is far from the
// constant (|A^T 1| / |A| = 0.14 on the 2.6k-particle channel) while A 1 = 0.  A^T lambda = g is
// consistent for mean-zero g, and the DC-projected iteration keeps lambda mean-zero -- the solution
// the adjoint needs, since the forward solve finds A p = b - c 1 (the constant c making it
// consistent; its multiplier's condition is sum(lambda) = 0).  But on gradient-shaped right-hand
// sides the BiCGSTAB recurrence stalls near 3e-9 of |g| (the reference algorithm on the CPU does the
// same), so a single solve cannot meet 1e-8.  Each round here solves A^T d = r / |r| to 1e-6 (the
// unit scale also keeps the reference's first-iteration threshold from turning the criterion
// absolute), updates lambda += |r| d, and recomputes r = g - A^T lambda in FP64; two or three rounds
// reach 1e-12.  A round's inner tolerance is only what the requested one still needs (0.3 tol |g| /
// |r|, within [1e-6, 0.1]).  Closed domains remove the means of g and of each r first (the adjoint of
// the forward's DC removal of its right-hand side).  Stops when |r| <= tol |g| or after itermax inner
// iterations in total.
__global__ void k_fill_zero(double* a, int n);
constexpr double kRefineInnerTol = 1.0e-6;
constexpr int kRefineRounds = 12;

// r = g - A^T lambda (lambda = 0: r = g), partial sums of r
__global__ void k_ref_resid(Mat A, const double* g, const double* lam, int first, double* r, double* partials) {
  const int n = A.n;
  RowRange rr = block_rows(n, gridDim.x);
  Red<1> R(nullptr, n);
  for (int i = rr.begin + threadIdx.x; i < rr.end; i += blockDim.x) {
    const double v = first ? g[i] : g[i] - spmv_row_transpose(A, i, lam);
    r[i] = v;
    R.add(0, i, v);
  }
  R.store(partials);
}

// out[0] = mean(r) (0 when open)
__global__ void k_ref_mean(const double* partials, int nblk, int n, int open, double* out) {
  const double f = fold(partials, 0, nblk, nullptr, n);
  if (threadIdx.x == 0) out[0] = open ? 0.0 : f / (double)n;
}

// partial sums of (r - mean)^2: a second pass, since sum(r^2) - n mean^2 cancels when r's constant
// part dominates (a right-hand side with a constant added)
__global__ void k_ref_dev(int n, const double* r, const double* sc, double* partials) {
  RowRange rr = block_rows(n, gridDim.x);
  const double m = sc[0];
  Red<1> R(nullptr, n);
  for (int i = rr.begin + threadIdx.x; i < rr.end; i += blockDim.x) {
    const double d = r[i] - m;
    R.add(0, i, d * d);
  }
  R.store(partials);
}

// out[1] = |r - mean|
__global__ void k_ref_fin(const double* partials, int nblk, int n, double* out) {
  const double q = fold(partials, 0, nblk, nullptr, n);
  if (threadIdx.x == 0) out[1] = sqrt(q);
}

__global__ void k_ref_scale(int n, const double* r, const double* sc, double* b) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) b[i] = (r[i] - sc[0]) / sc[1];
}

__global__ void k_ref_update(int n, double* lam, const double* d, const double* sc, int first) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) lam[i] = first ? sc[1] * d[i] : lam[i] + sc[1] * d[i];
}

vb_status adjoint_solve(vb_ctx* c, double tol, int itermax, int* iters, double* resid, int* status) {
  if (!c->op.step_matrix) {
    c->err = "the adjoint solve needs the matrix of a step's vb_ppe_solve";
    return VB_ERR_ARG;
  }
  vb_status s;
  if (!c->op.valid && (s = build_transpose(c)) != VB_OK) return s;
  if (!c->op.rref && (s = alloc_operator(c, 0)) != VB_OK) return s;
  OperatorMap& op = c->op;
  const int n = c->npfb, nb = grid_for(n), open = c->g.open ? 1 : 0;
  Mat A = step_matrix(c);
  Mat As = A;                                   // the solve's SpMVs: listed rows in workgroups of their own
  if (!c->red) split_listed_rows(c, As);
  Vecs w{op.dref, op.bref, c->r, c->rb, c->vv, c->pn, c->pinv, c->ss, c->sinv, c->tt, open ? c->a_diag : nullptr};
  double h[2], gnorm = 0.0, rnorm = 0.0;
  int total = 0, st = VB_OK;
  for (int round = 0; round < kRefineRounds; round++) {
    VB_LAUNCH(c, k_ref_resid, nb, A, op.g, op.lam, round == 0 ? 1 : 0, op.rref, c->partials);
    VB_FIN(c, k_ref_mean, c->partials, nb, n, open, op.refs);
    VB_LAUNCH(c, k_ref_dev, nb, n, op.rref, op.refs, c->partials);
    VB_FIN(c, k_ref_fin, c->partials, nb, n, op.refs);
    VB_HIP(c, hipMemcpyAsync(h, op.refs, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    VB_HIP(c, hipStreamSynchronize(c->stream));
    rnorm = h[1];
    if (round == 0) gnorm = rnorm;
    if (rnorm == 0.0 || rnorm <= tol * gnorm) {
      if (round == 0) VB_LAUNCH(c, k_fill_zero, grid_1d(n), op.lam, n);
      break;
    }
    if (total >= itermax) { st = VB_WARN_MAXITER; break; }
    // a round solves to 1e-6, or only as far as the requested tolerance still needs (with a factor
    // 3 of margin, and at least a tenfold reduction) once the residual is close to it
    const double inner = std::min(0.1, std::max(kRefineInnerTol, 0.3 * tol * gnorm / rnorm));
    VB_LAUNCH(c, k_ref_scale, grid_1d(n), n, op.rref, op.refs, op.bref);
    // the inner solve reads the compact rows' A^T values as FP32 on the mixed path (exact there); the
    // refinement residual above is FP64
    s = c->prm.precision != VB_PREC_MIXED ? solve<SPMV_TRANSPOSE>(c, As, w, 0, inner, itermax - total)
        : c->vcache ? solve<SPMV_TRANSPOSE_MIXED, true>(c, As, w, 0, inner, itermax - total)
                    : solve<SPMV_TRANSPOSE_MIXED>(c, As, w, 0, inner, itermax - total);
    if (s != VB_OK) return s;
    total += std::min(c->scal_host->iter, itermax - total);
    if (c->scal_host->status == VB_WARN_BICGSTAB_BREAKDOWN) st = VB_WARN_BICGSTAB_BREAKDOWN;
    VB_LAUNCH(c, k_ref_update, grid_1d(n), n, op.lam, op.dref, op.refs, round == 0 ? 1 : 0);
    if (st == VB_WARN_BICGSTAB_BREAKDOWN) break;
    if (round == kRefineRounds - 1) st = VB_WARN_MAXITER;
  }
  if (st == VB_OK && !(rnorm <= tol * gnorm || rnorm == 0.0)) st = VB_WARN_MAXITER;
  if (iters) *iters = total;
  if (resid) *resid = gnorm > 0.0 ? rnorm / gnorm : 0.0;
  if (status) *status = st;
  return VB_OK;
}

__global__ void k_fill_zero(double* a, int n) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) a[i] = 0.0;
}

// One application of a solver's own SpMV row to x (no reductions): the forward FP64 solve's row
// (compact rows in column order, the others full) or the adjoint solve's A^T row, for direct checks.
template <int MODE>
__global__ void k_apply_solver_row(Mat A, const double* x, double* y) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < A.n) y[i] = spmv_row<MODE>(A, i, x);
}

vb_status apply_solver_spmv(vb_ctx* c, bool transpose, const double* x, double* y) {
  if (!c->op.step_matrix) {
    c->err = "the solver's SpMV needs the matrix of a step's vb_ppe_solve";
    return VB_ERR_ARG;
  }
  if (transpose && !c->op.valid) {
    vb_status s = build_transpose(c);
    if (s != VB_OK) return s;
  }
  // the compact rows gather x two elements at a time from even columns
  if (reinterpret_cast<uintptr_t>(x) % 16 != 0) {
    c->err = "the solver's SpMV needs x aligned to 16 bytes";
    return VB_ERR_ARG;
  }
  Mat A = step_matrix(c);
  if (c->spmv_stored || c->red) A.rowflag = nullptr;
  if (transpose) VB_LAUNCH(c, k_apply_solver_row<SPMV_TRANSPOSE>, grid_1d(c->npfb), A, x, y);
  else VB_LAUNCH(c, k_apply_solver_row<SPMV_FP64>, grid_1d(c->npfb), A, x, y);
  return VB_OK;
}

// ------------------------------------------------------------------------------------------------
// measurement hook: time the SpMV kernel of the current solve configuration
// ------------------------------------------------------------------------------------------------
__global__ void k_bi_set_live(Scal* sc, int* counts_out, const int* cnt, const uint8_t* rowflag, int n) {
  __shared__ int sh[2][kBlock];
  int s = 0, e = 0;
  for (int i = threadIdx.x; i < n; i += kBlock) {
    s += cnt[i];
    if (rowflag && rowflag[i]) e += cnt[i];
  }
  sh[0][threadIdx.x] = s; sh[1][threadIdx.x] = e;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) { sh[0][threadIdx.x] += sh[0][threadIdx.x + w]; sh[1][threadIdx.x] += sh[1][threadIdx.x + w]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { sc->done = 0; sc->exit_s = 0; counts_out[0] = sh[0][0]; counts_out[1] = sh[1][0]; }
}

vb_status profile_spmv(vb_ctx* c, int reps, double* avg_us, double* alg_bytes) {
  const int n = c->npfb;
  const bool mixed = c->prm.precision == VB_PREC_MIXED;
  Mat A = step_matrix(c);
  if ((c->spmv_stored || c->red) && !mixed) A.rowflag = nullptr;
  if (!c->red) split_listed_rows(c, A);
  const int nb = spmv_grid(n) + A.nxb;           // the launch the solve makes
  Vecs w = step_vecs(c, c->st.p, c->g.open != 0);
  int* d_counts = c->seg_counts;
  hipLaunchKernelGGL(k_bi_set_live, dim3(1), dim3(kBlock), 0, c->stream, c->scal, d_counts, c->cnt,
                     A.rowflag, n);
  VB_HIP(c, hipGetLastError());
  hipEvent_t e0, e1;
  VB_HIP(c, hipEventCreate(&e0));
  VB_HIP(c, hipEventCreate(&e1));
  const bool fused = c->red == nullptr;
  for (int r = 0; r <= reps; r++) {
    if (r == 1) VB_HIP(c, hipEventRecord(e0, c->stream));
    RedBufs rb{c->partials, nullptr, n};
    if (fused) {                     // the kernel the solve launches (five-kernel iteration)
      if (mixed && c->vcache) VB_LAUNCH_BS(c, (k_bi_spmv_p<SPMV_MIXED, true, true>), nb, kSpmvBlock, A, w, rb, c->scal);
      else if (mixed) VB_LAUNCH_BS(c, (k_bi_spmv_p<SPMV_MIXED, true>), nb, kSpmvBlock, A, w, rb, c->scal);
      else VB_LAUNCH_BS(c, (k_bi_spmv_p<SPMV_FP64, true>), nb, kSpmvBlock, A, w, rb, c->scal);
    } else {
      if (mixed) VB_LAUNCH_BS(c, k_bi_spmv_p<SPMV_MIXED>, nb, kSpmvBlock, A, w, rb, c->scal);
      else VB_LAUNCH_BS(c, k_bi_spmv_p<SPMV_FP64>, nb, kSpmvBlock, A, w, rb, c->scal);
    }
  }
  VB_HIP(c, hipEventRecord(e1, c->stream));
  VB_HIP(c, hipEventSynchronize(e1));
  float ms = 0.f;
  VB_HIP(c, hipEventElapsedTime(&ms, e0, e1));
  int counts[2] = {0, 0};
  VB_HIP(c, hipMemcpy(counts, d_counts, 8, hipMemcpyDeviceToHost));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *avg_us = 1000.0 * ms / reps;
  // Algorithmic bytes of one launch, per nonzero whatever the storage: per row length 4, row flag 1,
  // diagonal 8, the gathered vector 8 (each element once), r_bar 8, v written 8; per off-diagonal
  // nonzero a 2-byte relative column and the value (8 bytes FP64, 4 mixed) on compact rows, a 4-byte
  // column and an 8-byte value on full rows.  (The column blocks store one 2-byte start per block
  // instead of a column per nonzero, and zeros where a block column is not a neighbour.)
  const double nnz = counts[0], nnz_full = A.rowflag ? counts[1] : counts[0];
  *alg_bytes = 37.0 * n + (mixed ? 6.0 : 10.0) * (nnz - nnz_full) + 12.0 * nnz_full;
  return VB_OK;
}

}  // namespace vb
