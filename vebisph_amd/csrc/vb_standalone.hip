// Stand-alone device kernels of the C ABI: elementwise pieces of the step that callers (and the
// parity tests) can run on their own device arrays.
#include <mutex>

#include "vb_ctx.h"
#include "vb_logconf.h"

namespace vb {

// nnf_models_mod.F90:8-110.  The reference keeps these with no call sites; they are elementwise.
__global__ void k_nnf(int model, int n, const double* sr, double mu, double* out) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double s = sr[i], r;
  switch (model) {
    case 0: r = s; break;                                               // :8-21 copies mu(i)
    case 1: r = mu * pow(s + 1.0, 1.5 - 1.0); break;                    // :23-39, n = 1.5
    case 2: {                                                           // :41-61
      double ty = 1.0e-2, al = 1.0e2;
      r = (s <= ty / (mu * al)) ? mu * al : mu + ty / s;
    } break;
    case 3: {                                                           // :63-82
      double m0 = mu, mi = 0.05 * m0;
      r = mi + (m0 - mi) / (1.0 + pow(1.0e-3 * s, 3.0));
    } break;
    case 4: {                                                           // :84-108
      double m0 = mu, mi = 1.0e-3;
      r = mi + (m0 - mi) * pow(1.0 + pow(0.14 * s, 1.5), (0.1 - 1.0) / 1.5);
    } break;
    default: r = 0.0;
  }
  out[i] = r;
}

__global__ void k_kernel_functions(int n, const double* q, double ad7, double ad7h, double* f, double* w) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  f[i] = fac(q[i], ad7h);
  w[i] = wab(q[i], ad7);
}

__global__ void k_logconf_n(int n, const double* tau, const double* trc, const double* gv, double dt,
                            double wi, double eps, const double* kick, int use_kick, double* tout,
                            double* trout) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double t[4] = {tau[4 * i], tau[4 * i + 1], tau[4 * i + 2], tau[4 * i + 3]};
  double G[4] = {gv[4 * i], gv[4 * i + 1], gv[4 * i + 2], gv[4 * i + 3]};
  double o[4], tr;
  logconf(t, trc[i], G, dt, wi, eps, use_kick ? kick[i] : 0.0, use_kick != 0, o, &tr);
  tout[4 * i] = o[0]; tout[4 * i + 1] = o[1]; tout[4 * i + 2] = o[2]; tout[4 * i + 3] = o[3];
  trout[i] = tr;
}

// Adjoint of the map for n particles: with the map's outputs (tau', trace') and their cotangents
// (tau'_bar, trace'_bar), the cotangents of its inputs J^T (tau'_bar, trace'_bar) for
// (tau, trace, grad_vel, Wi, dt), per particle (Wi and dt are shared by all particles: a caller sums
// their per-particle contributions).  Column j of the Jacobian J is one pass of the map in Dual
// arithmetic seeded along input j: eleven passes, exact derivatives of the branch each value takes.
__global__ void k_logconf_adjoint(int n, const double* tau, const double* trc, const double* gv, double dt,
                                  double wi, double eps, const double* kick, int use_kick, const double* tbar,
                                  const double* trbar, double* tau_bar, double* trc_bar, double* gv_bar,
                                  double* wi_bar, double* dt_bar) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  // inputs: tau (0-3), trace (4), grad_vel (5-8), Wi (9), dt (10)
  const double x[11] = {tau[4 * i], tau[4 * i + 1], tau[4 * i + 2], tau[4 * i + 3], trc[i],
                        gv[4 * i],  gv[4 * i + 1],  gv[4 * i + 2],  gv[4 * i + 3],  wi, dt};
  const double ob[5] = {tbar[4 * i], tbar[4 * i + 1], tbar[4 * i + 2], tbar[4 * i + 3], trbar[i]};
  const double k = use_kick ? kick[i] : 0.0;
  double xb[11];
  for (int j = 0; j < 11; j++) {
    Dual in[11];
    for (int q = 0; q < 11; q++) in[q] = Dual(x[q], q == j ? 1.0 : 0.0);
    Dual o[4], tr;
    logconf(in, in[4], in + 5, in[10], in[9], eps, k, use_kick != 0, o, &tr);
    xb[j] = (((ob[0] * o[0].d + ob[1] * o[1].d) + ob[2] * o[2].d) + ob[3] * o[3].d) + ob[4] * tr.d;
  }
  for (int q = 0; q < 4; q++) { tau_bar[4 * i + q] = xb[q]; gv_bar[4 * i + q] = xb[5 + q]; }
  trc_bar[i] = xb[4];
  wi_bar[i] = xb[9];
  dt_bar[i] = xb[10];
}

// minstd stream of predictor_mod.F90:357,369-371 (see vb_capi.hip), computed on the device by
// jumping the generator ahead: x_k = 48271^k mod (2^31-1).
__device__ inline uint64_t mulmod(uint64_t a, uint64_t b) { return (a * b) % 2147483647ull; }
__device__ inline uint64_t powmod(uint64_t e) {
  uint64_t r = 1, b = 48271;
  while (e) { if (e & 1) r = mulmod(r, b); b = mulmod(b, b); e >>= 1; }
  return r;
}
__global__ void k_kick_table(int n, double* kick) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  uint64_t a = powmod(2ull * i + 1), b = mulmod(a, 48271);
  uint64_t f = ((a << 30) | ((b - 1) & ((1ull << 30) - 1))) >> 7;
  kick[i] = (1.0e13 * kEps) * (ldexp((double)f, -54) - 0.5);
}

// Scratch of vb_logconf_noalloc: the per-particle kick table lives in this file-scope device
// array instead of a hipMalloc'd buffer.  Each GPU holds its own copy until the process ends; host
// code reaches it only through the address hipGetSymbolAddress returns for the current device.
constexpr int kScratchMax = 1 << 20;            // particles per call (8 MiB of doubles)
static __device__ __align__(256) double scratch[kScratchMax];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices];      // device address of `scratch`, per device

// The current device's copy of `scratch`, looked up once per device.
static vb_status scratch_address(double** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return VB_ERR_HIP;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return VB_ERR_HIP;
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return VB_OK;
}

}  // namespace vb

using namespace vb;

extern "C" {

vb_status vb_nnf(int32_t model, int32_t n, const double* d_sr, double mu, double* d_out) {
  if (n < 0 || model < 0 || model > 4) return VB_ERR_ARG;
  if (n == 0) return VB_OK;
  hipLaunchKernelGGL(k_nnf, dim3(grid_1d(n)), dim3(kBlock), 0, 0, model, n, d_sr, mu, d_out);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(0) != hipSuccess) return VB_ERR_HIP;
  return VB_OK;
}

vb_status vb_kernel_functions(int32_t n, const double* d_q, double ad7, double ad7h, double* d_f, double* d_w) {
  if (n < 0) return VB_ERR_ARG;
  if (n == 0) return VB_OK;
  hipLaunchKernelGGL(k_kernel_functions, dim3(grid_1d(n)), dim3(kBlock), 0, 0, n, d_q, ad7, ad7h, d_f, d_w);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(0) != hipSuccess) return VB_ERR_HIP;
  return VB_OK;
}

vb_status vb_logconf(int32_t n, const double* d_tau, const double* d_trace, const double* d_gradvel,
                     double dt, double wi, double eps_ptt, int32_t kick, double* d_tau_out, double* d_trace_out) {
  if (n < 0) return VB_ERR_ARG;
  if (n == 0) return VB_OK;
  double* dk = nullptr;
  if (kick) {
    if (hipMalloc((void**)&dk, (size_t)n * 8) != hipSuccess) return VB_ERR_HIP;
    hipLaunchKernelGGL(k_kick_table, dim3(grid_1d(n)), dim3(kBlock), 0, 0, n, dk);
  }
  hipLaunchKernelGGL(k_logconf_n, dim3(grid_1d(n)), dim3(kBlock), 0, 0, n, d_tau, d_trace, d_gradvel, dt,
                     wi, eps_ptt, dk, kick, d_tau_out, d_trace_out);
  hipError_t e = hipGetLastError();
  hipError_t e2 = hipStreamSynchronize(0);
  if (dk) (void)hipFree(dk);
  return (e != hipSuccess || e2 != hipSuccess) ? VB_ERR_HIP : VB_OK;
}

vb_status vb_logconf_noalloc(int32_t n, const double* d_tau, const double* d_trace, const double* d_gradvel,
                             double dt, double wi, double eps_ptt, int32_t kick, double* d_tau_out,
                             double* d_trace_out) {
  if (n < 0) return VB_ERR_ARG;
  if (n > kScratchMax) return VB_ERR_CAPACITY;
  if (n == 0) return VB_OK;
  double* dk = nullptr;
  if (kick) {
    vb_status s = scratch_address(&dk);
    if (s != VB_OK) return s;
    hipLaunchKernelGGL(k_kick_table, dim3(grid_1d(n)), dim3(kBlock), 0, 0, n, dk);
  }
  hipLaunchKernelGGL(k_logconf_n, dim3(grid_1d(n)), dim3(kBlock), 0, 0, n, d_tau, d_trace, d_gradvel, dt,
                     wi, eps_ptt, dk, kick, d_tau_out, d_trace_out);
  hipError_t e = hipGetLastError();
  hipError_t e2 = hipStreamSynchronize(0);
  return (e != hipSuccess || e2 != hipSuccess) ? VB_ERR_HIP : VB_OK;
}

vb_status vb_logconf_adjoint(int32_t n, const double* d_tau, const double* d_trace, const double* d_gradvel,
                             double dt, double wi, double eps_ptt, int32_t kick, const double* d_tau_out_bar,
                             const double* d_trace_out_bar, double* d_tau_bar, double* d_trace_bar,
                             double* d_gradvel_bar, double* d_wi_bar, double* d_dt_bar) {
  if (n < 0) return VB_ERR_ARG;
  if (n == 0) return VB_OK;
  const void* need[] = {d_tau, d_trace, d_gradvel, d_tau_out_bar, d_trace_out_bar, d_tau_bar, d_trace_bar,
                        d_gradvel_bar, d_wi_bar, d_dt_bar};
  for (const void* p : need)
    if (!p) return VB_ERR_ARG;
  double* dk = nullptr;
  if (kick) {
    if (hipMalloc((void**)&dk, (size_t)n * 8) != hipSuccess) return VB_ERR_HIP;
    hipLaunchKernelGGL(k_kick_table, dim3(grid_1d(n)), dim3(kBlock), 0, 0, n, dk);
  }
  hipLaunchKernelGGL(k_logconf_adjoint, dim3(grid_1d(n)), dim3(kBlock), 0, 0, n, d_tau, d_trace, d_gradvel,
                     dt, wi, eps_ptt, dk, kick, d_tau_out_bar, d_trace_out_bar, d_tau_bar, d_trace_bar,
                     d_gradvel_bar, d_wi_bar, d_dt_bar);
  hipError_t e = hipGetLastError();
  hipError_t e2 = hipStreamSynchronize(0);
  if (dk) (void)hipFree(dk);
  return (e != hipSuccess || e2 != hipSuccess) ? VB_ERR_HIP : VB_OK;
}

}  // extern "C"
