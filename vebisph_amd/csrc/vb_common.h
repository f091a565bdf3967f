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

constexpr int kBlock = 256;                // threads per workgroup (4 wavefronts of 64)
constexpr int kMaxBlocks = 2048;           // grid cap for reduction kernels (8 per CU)
constexpr int kNplink = VB_NPLINK;
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
};

// Particle arrays (device pointers).  Rows 0..npfb-1 are wall+fluid, npfb..np-1 mirrors.
struct State {
  double *x, *y, *u, *v, *th, *t11, *t12, *t21, *t22, *trc, *p;
  uint8_t* inbin;
  double *x0, *y0, *u0, *v0;
  int *irel, *vrel;                        // by row (mirror rows only)
  double* dP;
};

// Device scalars of one step (one struct in device memory).
struct Scal {
  double dt, umax, mean, wab1;
  int np, nmirror, n_inbin, maxneigh, err, seg_base;
  // BiCGSTAB
  double rho, rho0, alpha, omega, beta, n2s0, n2r0, resid, mean_p, mean_s, tol;
  int iter, itermax, done, exit_s, status;
  unsigned counter;                        // last-block-done ticket
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

// ------------------------------------------------------------------------------------------------
// Deterministic reductions.  Wavefront tree by DPP/shuffle, then the block's 4 waves through LDS,
// then one partial per block; the last block to finish (acq_rel ticket) folds the partials in
// index order.  Same inputs and grid => same bits, run after run.
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

// Sum of NV values over the block; result valid in thread 0.
template <int NV>
__device__ inline void block_sum(double (&v)[NV]) {
  __shared__ double sh[NV][kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
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
      for (int w = 1; w < kBlock / 64; w++) s += sh[q][w];
      v[q] = s;
    }
  }
  __syncthreads();
}

// Writes this block's NV partials, takes a ticket, and returns true in ALL threads of the block that
// finished last.  That block then sees every other block's partials (release/acquire at agent
// scope, which also orders the per-XCD L2s) and must reset the ticket.
template <int NV>
__device__ inline bool publish_partials(double (&v)[NV], double* partials, unsigned* ticket) {
  __shared__ bool last;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; q++)
      __hip_atomic_store(&partials[q * gridDim.x + blockIdx.x], v[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    unsigned t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    last = (t == gridDim.x - 1);
    if (last) __threadfence();
  }
  __syncthreads();
  return last;
}

// In the last block: fold the partials of quantity q in block order; result in all threads.
__device__ inline double fold_partials(const double* partials, int q, int nblk) {
  __shared__ double sh[kBlock];
  double s = 0.0;
  // each thread sums a contiguous run, then a fixed tree: deterministic for a given nblk
  int per = (nblk + kBlock - 1) / kBlock;
  int b0 = threadIdx.x * per, b1 = min(nblk, b0 + per);
  for (int b = b0; b < b1; b++)
    s += __hip_atomic_load(&partials[q * nblk + b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

// Rows owned by a block: contiguous chunks, with consecutive chunks placed on the same XCD
// (workgroups are dealt round-robin to the 8 XCDs, each with its own L2), so a block's neighbour
// gathers mostly hit rows its XCD already holds.
struct RowRange { int begin, end; };
__device__ inline RowRange block_rows(int n, int nblk) {
  int b = blockIdx.x;
  int lb = b;
  if ((nblk & 7) == 0) lb = (b & 7) * (nblk >> 3) + (b >> 3);
  int chunk = (n + nblk - 1) / nblk;
  chunk = (chunk + kBlock - 1) / kBlock * kBlock;
  int r0 = lb * chunk;
  int r1 = min(n, r0 + chunk);
  return {r0, r1};
}

inline int grid_for(int n) {
  int b = (n + kBlock - 1) / kBlock;
  if (b > kMaxBlocks) b = kMaxBlocks;
  if (b > 8) b = (b + 7) / 8 * 8;          // multiple of 8 for the XCD mapping
  if (b < 1) b = 1;
  return b;
}

inline int grid_1d(long n) { long b = (n + kBlock - 1) / kBlock; return (int)(b < 1 ? 1 : b); }

}  // namespace vb
