// Shared device helpers, geometry and context layout for the gfx950 VEBISPH step.
//
// Everything here is FP64 unless named otherwise.  The library is compiled with -ffp-contract=off
// so that every a*b+c is two rounded operations, as in the reference's FP64 CPU code: the pair
// test, the kernel weights and the operators then agree with the reference bit for bit, and the
// only differences left are the order of the global reductions (dot products, means).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/vebisph_gpu.h"

namespace vb {

#ifndef VB_SPMV_BLOCK
#define VB_SPMV_BLOCK 256
#endif
#ifndef VB_MAX_BLOCKS
#define VB_MAX_BLOCKS 512
#endif
constexpr double kPi = 3.14159265358979323846;   // common_parameter.F90:12
constexpr int kBlock = 256;                // threads per workgroup (4 wavefronts of 64)
// SpMV launch: one row per thread (up to kSpmvMaxBlocks workgroups), so that every row's chain of
// dependent gathers has its own wavefront slot.  Measured on the 1M-row PPE: one row per thread
// beats two by 8-15%, and 256-thread workgroups beat 512.
constexpr int kSpmvBlock = VB_SPMV_BLOCK;
constexpr int kSpmvMaxBlocks = 16384;
// Grid cap of the solver's streaming vector kernels (several rows per thread).  The kernel after
// each reduction folds the previous kernel's per-workgroup partials in every workgroup, so this
// cap also bounds the fold traffic (workgroups x partials).  Measured on the 1M-row PPE (ms per
// step, mixed): 256 -> 224.7, 512 -> 219.9, 1024 -> 223.9, 2048 -> 238.3.
constexpr int kMaxBlocks = VB_MAX_BLOCKS;
// Workgroups of an SpMV launch that take the PPE's listed full rows (vb_ppe.hip: spmv_rows), at most
constexpr int kSpmvListBlocks = 1024;
constexpr int kPartialStride = kSpmvMaxBlocks + kSpmvListBlocks > kMaxBlocks ? kSpmvMaxBlocks + kSpmvListBlocks
                                                                            : kMaxBlocks;
static_assert(kSpmvBlock % 64 == 0 && kSpmvBlock <= 1024, "SpMV block");
static_assert(kMaxBlocks % 8 == 0 && kSpmvMaxBlocks % 8 == 0, "grid caps");
constexpr int kSolverStages = 6;           // BiCGSTAB reduction stages with their own buffers
constexpr int kNplink = VB_NPLINK;
// The PPE's compact rows as column blocks (layout in vb_ppe.hip): rows in chunks of 64 (a
// wavefront), blocks of kBlkCols columns, at most kMaxBlk blocks per row.
constexpr int kChunk = 64;
constexpr int kBlkCols = 8;
constexpr int kMaxBlk = 12;
constexpr int kOffGroups = kMaxBlk / 4;
inline size_t ppe_chunks(int npfb) { return ((size_t)npfb + kChunk - 1) / kChunk; }
inline size_t ppe_block_values(int npfb) { return ppe_chunks(npfb) * kChunk * kMaxBlk * kBlkCols; }
inline size_t ppe_block_offsets(int npfb) { return ppe_chunks(npfb) * kChunk * kMaxBlk; }
inline size_t ppe_list_slots(int npfb) { return ppe_chunks(npfb) * kChunk * kNplink; }
// MI355X's Infinity Cache (the memory-side cache in front of HBM) holds 256 MiB.  The mixed SpMVs
// read their FP32 value plane with ordinary (cacheable) loads when it takes at most 15/16 of that,
// so that the two SpMVs of an iteration and the next iteration's find it there; a larger plane is
// streamed with non-temporal loads (vb_ppe.hip: row_blk32, solve).
constexpr double kCachedValueBytes = 256.0 * 1024 * 1024 * 15 / 16;
constexpr double kEps = 2.220446049250313e-16;   // epsilon(1d0)

// Boundary description and derived constants, passed by value to kernels (< 4 KiB).
struct Geom {
  int nbp, ncirc, npfb, nbS;
  double bnode[VB_MAX_PATCHES][2], bedge[VB_MAX_PATCHES][2];
  double bnorm[VB_MAX_PATCHES][2];         // bound_norm, mirror_boundaries_mod.F90:464-474
  double blen[VB_MAX_PATCHES];             // sqrt(b_edge . b_edge)
  double bvel[VB_MAX_PATCHES], bth[VB_MAX_PATCHES], bang[VB_MAX_PATCHES];
  int btype[VB_MAX_PATCHES], bpp[VB_MAX_PATCHES];   // bpp 0-based
  double ccen[VB_MAX_CIRCLES][2], crad[VB_MAX_CIRCLES], comg[VB_MAX_CIRCLES];
  double cvel[VB_MAX_CIRCLES][2], cth[VB_MAX_CIRCLES];
  double h, dx, dv, sup, sup2, uno, ad7, ad7h, eta2, eta_small, oneoverh;
  double xmint, xmaxt, ymint, ymaxt;
  int ncx, ncy, nct;
  double bf[2];
  int open;                                // i_open_domain: free surface (freesurf_mod.F90)
  int shift_coeff;                         // i_shift_coeff (new_fick_shift.F90:43-51)
};

// Particle arrays (device pointers).  Rows 0..npfb-1 are wall+fluid, npfb..np-1 mirrors.
struct State {
  double *x, *y, *u, *v, *th, *t11, *t12, *t21, *t22, *trc, *p;
  uint8_t* inbin;
  double *x0, *y0, *u0, *v0;
  int *irel, *vrel;                        // by row (mirror rows only)
  double* dP;
};

// Linear parts of the affine maps the mirror phase applies, recorded (only during the adjoint's
// re-run of a step; all null otherwise) for the reverse sweep:
//   mjac[6m..6m+5]  mirror m (row npfb + m): J, the 2x2 Jacobian (row-major) of its position in its
//                   parent's position, then d dP_m / d x_parent (2);
//   msgn[m]         bit 0: its phase-3 velocity is minus the parent's (plus a constant), bit 1: so is
//                   its temperature, bit 2: the map is not affine (a circle mirror);
//   ejac[4i..]      particle i: the 2x2 Jacobian of replace_escapees' move (identity if not moved);
//   eflag[i]        1 moved across a patch, 2 moved by a circle (not affine), 0 not moved.
struct LinRec {
  double* mjac;
  int* msgn;
  double* ejac;
  int* eflag;
};

// Device scalars of one step (one struct in device memory).
struct Scal {
  double dt, umax, mean, wab1;
  int np, nmirror, n_inbin, maxneigh, err, seg_base;
  // BiCGSTAB
  double rho, rho0, alpha, omega, beta, n2s0, n2r0, resid, mean_p, mean_s, tol;
  double n2_small;                         // first-iteration norm threshold (1e-6, or 4 when open)
  int iter, itermax, done, exit_s, status;
  // BiCGSTAB in five kernels per iteration: rho of iteration k in rho_pp[k & 1]; sums of r, v and
  // of the projected p, each recorded by the kernel that folds it
  double rho_pp[2], sum_r, sum_v, sum_p;
  int nexplicit;                           // PPE rows with mirror neighbours (full stored rows)
  int nblocks;                             // column blocks of the compact rows, summed over chunks
  // output quantities: RMS velocity (secondary_properties), max |u|, |v|, |p| over rows 1..npfb
  double u_rms, fmax[3];
  // diagnostics: per circle Fx, Fy, boundary-particle count; the Nusselt number
  double circ[3 * VB_MAX_CIRCLES], nusselt;
};

// Error bits accumulated on the device (vb_status codes as bit positions).
enum { kErrOutOfGrid = 1 << VB_ERR_OUT_OF_GRID, kErrCellOverflow = 1 << VB_ERR_CELL_OVERFLOW,
       kErrAllBinned = 1 << VB_ERR_ALL_BINNED, kErrCapacity = 1 << VB_ERR_CAPACITY };

// ------------------------------------------------------------------------------------------------
// Kernel functions, sphtools.F90:11-76.  Integer powers are left-to-right products, as the
// reference's compiler evaluates q**4 and q**5.
// ------------------------------------------------------------------------------------------------
__host__ __device__ inline double p4(double q) { return ((q * q) * q) * q; }
__host__ __device__ inline double p5(double q) { return (((q * q) * q) * q) * q; }

__host__ __device__ inline double fac(double qq, double ad7h) {
  int isel = (int)floor(qq);
  if (isel == 0) {
    double q3 = 3.0 - qq, q2 = 2.0 - qq, q1 = 1.0 - qq;
    return ad7h * (((-5.0 * p4(q3)) + 30.0 * p4(q2)) - 75.0 * p4(q1));
  } else if (isel == 1) {
    double q3 = 3.0 - qq, q2 = 2.0 - qq;
    return ad7h * ((-5.0 * p4(q3)) + 30.0 * p4(q2));
  } else if (isel == 2) {
    double q3 = 3.0 - qq;
    return ad7h * (-5.0 * p4(q3));
  }
  return 0.0;
}

__host__ __device__ inline double wab(double qq, double ad7) {
  int isel = (int)floor(qq);
  if (isel == 0) {
    double q3 = 3.0 - qq, q2 = 2.0 - qq, q1 = 1.0 - qq;
    return ad7 * ((p5(q3) - 6.0 * p5(q2)) + 15.0 * p5(q1));
  } else if (isel == 1) {
    double q3 = 3.0 - qq, q2 = 2.0 - qq;
    return ad7 * (p5(q3) - 6.0 * p5(q2));
  } else if (isel == 2) {
    double q3 = 3.0 - qq;
    return ad7 * p5(q3);
  }
  return 0.0;
}

// Visits the neighbour slots k = 0..n-1 of row i in order: body(s, j, g) with s = k*npfb + i (the
// slot-major index), j = link[s] and g = gather(j), the neighbour's operands.  Four slots at a time
// load their indices, then their operands, before any body runs, so a row's chain of dependent loads
// is a quarter as long; the bodies still run in slot order (same arithmetic, same bits).
template <class Gather, class Body>
__device__ inline void for_slots(const int* __restrict__ link, int i, int n, int npfb, Gather gather, Body body) {
  int k = 0;
  for (; k + 4 <= n; k += 4) {
    const int s0 = k * npfb + i, s1 = s0 + npfb, s2 = s1 + npfb, s3 = s2 + npfb;
    const int j0 = link[s0], j1 = link[s1], j2 = link[s2], j3 = link[s3];
    const auto g0 = gather(j0);
    const auto g1 = gather(j1);
    const auto g2 = gather(j2);
    const auto g3 = gather(j3);
    body(s0, j0, g0); body(s1, j1, g1); body(s2, j2, g2); body(s3, j3, g3);
  }
  for (; k < n; k++) {
    const int s = k * npfb + i, j = link[s];
    body(s, j, gather(j));
  }
}

// ------------------------------------------------------------------------------------------------
// Deterministic reductions.  Wavefront tree by shuffles, the block's 4 waves through LDS, one
// partial per block (plain store); a one-workgroup finalize kernel then folds the partials in a
// fixed order and computes the next scalar.  The kernel boundary makes the partials visible, so no
// device-scope fences are needed (an agent-scope release on gfx950 writes back the XCD's L2, which
// per workgroup costs more than the reduction).  Same inputs and grid => same bits, run after run.
//
// Serial mode (parity runs): producers store every row's term instead, and the finalize sums them
// in index order on one thread, i.e. exactly the reference's single-thread loops.
// ------------------------------------------------------------------------------------------------
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

__device__ inline double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) { double w = __shfl_down(v, o, 64); v = w > v ? w : v; }
  return v;
}

// Sum of NV values over the block (any multiple of 64 threads up to 1024); result in thread 0.
template <int NV>
__device__ inline void block_sum(double (&v)[NV]) {
  __shared__ double sh[NV][16];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; q++) {
    double w = wave_sum(v[q]);
    if (lane == 0) sh[q][wid] = w;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; q++) {
      double s = sh[q][0];
      for (int w = 1; w < nw; w++) s += sh[q][w];
      v[q] = s;
    }
  }
  __syncthreads();
}

// Per-row reduction terms: accumulated per thread, or stored per row in serial mode.
template <int NV>
struct Red {
  double acc[NV];
  double* rows;      // non-null in serial mode: rows[q*n + i]
  int n;
  __device__ Red(double* r, int n_) : rows(r), n(n_) {
#pragma unroll
    for (int q = 0; q < NV; q++) acc[q] = 0.0;
  }
  __device__ inline void add(int q, int i, double v) {
    if (rows) rows[q * n + i] = v; else acc[q] += v;
  }
  // block partials (plain stores); partials[q*gridDim.x + blockIdx.x]
  __device__ inline void store(double* partials) {
    if (rows) return;
    block_sum<NV>(acc);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int q = 0; q < NV; q++) partials[q * gridDim.x + blockIdx.x] = acc[q];
    }
  }
};

// Fold of NV quantities inside one workgroup (every thread of the workgroup calls it; any multiple
// of 64 threads): the partials of quantity q, partials[q*nblk + b], are read coalesced (thread t
// takes b = t, t + blockDim, ...), summed per thread, per wavefront by shuffles, then the wavefront
// sums in wavefront order by every thread.  The order depends only on nblk and blockDim, so all
// workgroups of a kernel get the same bits.  In serial mode rows[q*n + 0..n-1] are summed in index
// order by one thread instead.  Result in all threads.
template <int NV>
__device__ inline void fold_n(const double* partials, int nblk, const double* rows, int n, double (&out)[NV]) {
  __shared__ double sh[NV][16];
  if (rows) {
    if (threadIdx.x == 0) {
#pragma unroll
      for (int q = 0; q < NV; q++) {
        double s = 0.0;
        for (int i = 0; i < n; i++) s = s + rows[(size_t)q * n + i];
        sh[q][0] = s;
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; q++) out[q] = sh[q][0];
    __syncthreads();
    return;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  // Loads go out kFoldBatch per quantity at a time and are summed afterwards in b order (a
  // missing b adds an exact 0.0), so a thread waits on one round trip per batch, not per partial.
  constexpr int kFoldBatch = 8;
  double s[NV];
#pragma unroll
  for (int q = 0; q < NV; q++) s[q] = 0.0;
  for (int b0 = threadIdx.x; b0 < nblk; b0 += kFoldBatch * blockDim.x) {
    double v[NV][kFoldBatch];
#pragma unroll
    for (int u = 0; u < kFoldBatch; u++) {
      const int b = b0 + u * blockDim.x;
#pragma unroll
      for (int q = 0; q < NV; q++) v[q][u] = b < nblk ? partials[q * nblk + b] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kFoldBatch; u++)
#pragma unroll
      for (int q = 0; q < NV; q++) s[q] += v[q][u];
  }
#pragma unroll
  for (int q = 0; q < NV; q++) {
    const double t = wave_sum(s[q]);
    if (lane == 0) sh[q][wid] = t;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; q++) {
    double s = sh[q][0];
    for (int w = 1; w < nw; w++) s += sh[q][w];
    out[q] = s;
  }
  __syncthreads();
}

__device__ inline double fold(const double* partials, int q, int nblk, const double* rows, int n) {
  double r[1];
  fold_n<1>(partials + (size_t)q * nblk, nblk, rows ? rows + (size_t)q * n : nullptr, n, r);
  return r[0];
}

// Rows owned by a block: contiguous chunks, with consecutive chunks placed on the same XCD
// (workgroups are dealt round-robin to the 8 XCDs, each with its own L2), so a block's neighbour
// gathers mostly hit rows its XCD already holds.
struct RowRange { int begin, end; };
__device__ inline RowRange block_rows(int n, int b, int nblk) {
  int lb = b;
  if ((nblk & 7) == 0) lb = (b & 7) * (nblk >> 3) + (b >> 3);
  int chunk = (n + nblk - 1) / nblk;
  chunk = (chunk + blockDim.x - 1) / blockDim.x * blockDim.x;
  int r0 = lb * chunk;
  int r1 = min(n, r0 + chunk);
  return {r0, r1};
}
__device__ inline RowRange block_rows(int n, int nblk) { return block_rows(n, (int)blockIdx.x, nblk); }

inline int grid_for(int n) {
  int b = (n + kBlock - 1) / kBlock;
  if (b > kMaxBlocks) b = kMaxBlocks;
  if (b > 8) b = (b + 7) / 8 * 8;          // multiple of 8 for the XCD mapping
  if (b < 1) b = 1;
  return b;
}

inline int spmv_grid(int n) {
  int b = (n + kSpmvBlock - 1) / kSpmvBlock;
  if (b > kSpmvMaxBlocks) b = kSpmvMaxBlocks;
  if (b > 8) b = (b + 7) / 8 * 8;
  if (b < 1) b = 1;
  return b;
}

inline int grid_1d(long n) { long b = (n + kBlock - 1) / kBlock; return (int)(b < 1 ? 1 : b); }

// One row per thread, XCD-aware: the row of this thread in a launch of xcd_grid(n) workgroups of
// kBlock threads, or n for none.  The workgroups of one XCD take consecutive row ranges (block_rows),
// so the neighbour data a row gathers is fetched once into that XCD's L2; with blockIdx-ordered rows,
// consecutive workgroups land on different XCDs and every XCD whose rows touch a particle fetches it
// from HBM again (measured up to 2.2x the algorithmic bytes on the neighbour sums).
__device__ inline int xcd_row(int n) {
  const RowRange rr = block_rows(n, gridDim.x);
  const int i = rr.begin + (int)threadIdx.x;
  return i < rr.end ? i : n;
}
inline int xcd_grid(long n) {
  long b = (n + kBlock - 1) / kBlock;
  if (b > 8) b = (b + 7) / 8 * 8;          // multiple of 8 for the XCD mapping
  return (int)(b < 1 ? 1 : b);
}

}  // namespace vb
