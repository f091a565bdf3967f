// Host-side context of the gfx950 step: device buffers (capacity-sized, allocated once), the
// stream, and the launchers each phase file provides.
#pragma once

#include <string>

#include <type_traits>

#include "vb_common.h"

// The assembled PPE operator's transposed map (vb_operator.hip), built on first use after each
// assembly.
struct OperatorMap {
  double* tval = nullptr;             // real part of A^T, slot-major like the neighbour list
  int* tcol = nullptr;
  int *mcount = nullptr, *moff = nullptr, *mfill = nullptr, *nmiss = nullptr;
  int *mkey = nullptr, *mrow = nullptr;   // mirror entries per column, in (row, slot) order
  double* mval = nullptr;
  size_t cap_m = 0;
  double *g = nullptr, *lam = nullptr;   // adjoint solve: right-hand side and solution (npfb)
  double* bv64T = nullptr;            // A^T values of the compact rows (layout of vb_ctx::bv64)
  float* bv32T = nullptr;             //   the same in FP32 (mixed path; exact: its weights are FP32)
  int* xtcol = nullptr;               // A^T rows of the listed (full) rows in list order (xslot)
  double* xtval = nullptr;
  int* rslot = nullptr;               // reverse slot: the slot of c in i's list, beside tcol
  double *rref = nullptr, *bref = nullptr, *dref = nullptr;   // adjoint refinement: residual,
  double* refs = nullptr;                                       // scaled residual, correction, scalars
  double *ga = nullptr, *gb = nullptr;     // transposed gradient: per-row coefficients (npfb)
  double *gin = nullptr, *gout = nullptr;  // gradient operator staging (2 npfb each)
  bool valid = false;                 // the map matches the stored matrix
  bool step_matrix = false;           // the stored matrix is a step's PPE (not a caller's system)
};

struct AdjointCtx;                   // the adjoint sweep's record and work arrays (vb_adjoint.hip)

struct vb_ctx {
  vb_params prm;
  vb::Geom g;
  int npfb, nbS, cap, mcap;          // rows: cap = npfb + mcap, mcap = mirror capacity
  int itime;
  double time, dt_visc;
  hipStream_t stream;
  std::string err;

  vb::State st;                      // current particle arrays
  vb::LinRec rec{};                  // mirror/escape Jacobians: set only during the adjoint's re-run
  double *u_alt, *v_alt, *th_alt;    // prediction writes u*, v*, theta* here; then swapped in
  int* nsurf;

  // mirrors: segment compaction scratch
  int* seg_counts;                   // per-block counts / offsets
  int seg_blocks;

  // cells and neighbours
  int *cellkey, *cellslot, *nc, *ist, *ip;
  int* scan_bsum;                    // block sums of launch_exclusive_scan (one per 1024 elements)
  int *cnt, *link;                   // link[k*npfb + i], 0-based
  double *wg1, *wg2, *wl;            // slot-major like link (FP64 path)
  float *wg1f, *wg2f, *wlf;          // the same planes in FP32 (mixed path; then wg1.. are null)
  bool wf32;                         // the weight planes are FP32 (precision = VB_PREC_MIXED)
  double *k11, *k12, *k21, *k22;
  double *usx, *usy;
  double* kick;                      // per-particle B perturbation (reference_kick)
  double *gpx, *gpy;
  double* vort;                      // secondary_properties: vorticity of rows 0..npfb-1
  double* conc;                      // open domain: concentration (calc_part_dist_vars), else 0
  double *snx, *sny;                 // open domain: surface normal -h sum w_G (before smoothing)
  double* pscale;                    // open domain: PPE row scale, ppe_smooth or 1 (surface rows)

  // PPE (ELL, slot-major) and BiCGSTAB vectors
  double *a_diag, *a_val, *rhs;
  int *a_col, *a_len;                // a_len: row lengths of a caller-supplied system
  uint8_t* rowflag;                  // 1: SpMV uses the full stored row (mirror neighbours)
  int *xflag, *xscan, *xrows;        // the rows with rowflag 1, listed for their assembly
  int* xcol;                         // the listed rows' columns and values in list order (xslot)
  double* xval;
  int* chunkb;                       // compact rows: column blocks per 64-row chunk
  int16_t* boff;                     // compact rows, column blocks: start offsets
  double* bv64;                      //   FP64 values
  float* bv32;                       //   FP32 values (mixed path)
  uint8_t* bpos;                     //   block position of each neighbour slot (for A^T)

  bool spmv_stored;                  // FP64 path: full stored rows everywhere (VB_SPMV=stored)
  bool spmv_stream = false;          // mixed path: FP32 values always streamed (VB_SPMV=stream)
  bool vcache = false;               // mixed path: FP32 value plane read cacheably (kCachedValueBytes)
  double *r, *rb, *vv, *pn, *pinv, *ss, *sinv, *tt, *xsys;

  double* partials;                  // reduction partials, 2 * kPartialStride per solver stage
  double* red;                       // per-row reduction terms, 2 * npfb per stage (serial mode only)
  OperatorMap op;                    // A and A^T of the assembled PPE (the adjoint's operator)
  AdjointCtx* adj = nullptr;         // recorded forward run for vb_adjoint_backward
  vb::Scal* scal;                    // device scalars
  vb::Scal* scal_host;               // pinned mirror
  int nmirror_host, np_host;
};

namespace vb {
// phase launchers (vb_step.hip)
vb_status launch_set_time_step(vb_ctx* c);
vb_status launch_put_in_bin(vb_ctx* c);
vb_status launch_mirrors(vb_ctx* c);
vb_status launch_neighbours(vb_ctx* c);
vb_status launch_weights(vb_ctx* c);
vb_status launch_freesurf(vb_ctx* c);
vb_status launch_shift(vb_ctx* c);
vb_status launch_predict(vb_ctx* c);
vb_status launch_mirror_vel(vb_ctx* c);
vb_status launch_pressure_update(vb_ctx* c);
vb_status launch_final(vb_ctx* c);
vb_status launch_mirror_p(vb_ctx* c);               // mirror pressures p_parent + dP only
void free_adjoint(vb_ctx* c);                       // (vb_adjoint.hip)
vb_status launch_secondary_properties(vb_ctx* c);   // vort, Scal::u_rms
vb_status launch_field_maxima(vb_ctx* c);           // Scal::fmax
vb_status launch_circle_forces(vb_ctx* c);          // Scal::circ
vb_status launch_nusselt(vb_ctx* c, double scale);  // Scal::nusselt
// PPE (vb_ppe.hip)
vb_status launch_ppe(vb_ctx* c, double tol, int itermax);
vb_status assemble_ppe(vb_ctx* c);   // right-hand side and matrix only (launch_ppe = this + the solve)
vb_status profile_spmv(vb_ctx* c, int reps, double* avg_us, double* alg_bytes);
vb_status solve_nr(vb_ctx* c, const double* A_nr, const int* ija, double* x, const double* B, int n,
                   int itermax, double tol, bool open);
// the assembled PPE operator (vb_operator.hip): y = A x or A^T x on device vectors of npfb rows
vb_status apply_ppe_operator(vb_ctx* c, bool transpose, const double* x, double* y);
// the same through the solvers' own SpMV rows (forward FP64 compact rows, adjoint A^T rows; vb_ppe.hip)
vb_status apply_solver_spmv(vb_ctx* c, bool transpose, const double* x, double* y);
vb_status build_transpose(vb_ctx* c);
vb_status alloc_operator(vb_ctx* c, size_t bucket_cap);
vb_status launch_transpose_blocks(vb_ctx* c);
vb_status apply_grad_operator(vb_ctx* c, bool transpose, const double* x, double* y);
// A^T lambda = g, FP64 iterative refinement around the reference's BiCGSTAB on the transposed map
// (vb_ppe.hip); g, lambda in op; total inner iterations, |g - A^T lambda| / |g| and a vb_status
vb_status adjoint_solve(vb_ctx* c, double tol, int itermax, int* iters, double* resid, int* status);
// v -= mean(v) on the device (vb_operator.hip; uses the reduction partials)
vb_status remove_mean(vb_ctx* c, double* v, int n);
void free_operator(vb_ctx* c);
// out[0..n-1] = exclusive prefix sums of in[0..n-1], out[n] = total (vb_step.hip)
vb_status launch_exclusive_scan(vb_ctx* c, const int* in, int* out, int n);
// helpers (vb_capi.hip)
vb_status hip_check(vb_ctx* c, hipError_t e, const char* what);
vb_status sync_scalars(vb_ctx* c);
}  // namespace vb

#define VB_HIP(c, call)                                                     \
  do {                                                                      \
    vb_status _s = vb::hip_check((c), (call), #call);                       \
    if (_s != VB_OK) return _s;                                             \
  } while (0)

// Calls f(wg1, wg2, wl) with the context's kernel-weight planes -- FP32 on the mixed path, else
// FP64 -- and returns its status; kernels that read the planes take their element type as a template
// argument (weight_t<decltype(w1)>), widening each weight to FP64 as they load it.
template <class F>
inline vb_status with_weights(vb_ctx* c, F&& f) {
  if (c->wf32) return f(c->wg1f, c->wg2f, c->wlf);
  return f(c->wg1, c->wg2, c->wl);
}
template <class P>
using weight_t = std::remove_cv_t<std::remove_pointer_t<P>>;

#define VB_LAUNCH(c, kern, grid, ...)                                       \
  do {                                                                      \
    hipLaunchKernelGGL(kern, dim3(grid), dim3(vb::kBlock), 0, (c)->stream, __VA_ARGS__); \
    VB_HIP(c, hipGetLastError());                                           \
  } while (0)

#define VB_LAUNCH_BS(c, kern, grid, bs, ...)                                \
  do {                                                                      \
    hipLaunchKernelGGL(kern, dim3(grid), dim3(bs), 0, (c)->stream, __VA_ARGS__); \
    VB_HIP(c, hipGetLastError());                                           \
  } while (0)
