// C ABI of include/vebisph_gpu.h: context lifetime, state transfer, the phase entry points, exports,
// and the stand-alone kernels.  Host code only orchestrates; all particle work runs on the device.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "vb_ctx.h"

namespace vb {

vb_status hip_check(vb_ctx* c, hipError_t e, const char* what) {
  if (e == hipSuccess) return VB_OK;
  if (c) c->err = std::string(what) + ": " + hipGetErrorString(e);
  return VB_ERR_HIP;
}

vb_status sync_scalars(vb_ctx* c) {
  VB_HIP(c, hipMemcpyAsync(c->scal_host, c->scal, sizeof(Scal), hipMemcpyDeviceToHost, c->stream));
  VB_HIP(c, hipStreamSynchronize(c->stream));
  return VB_OK;
}

static vb_status check_err_bits(vb_ctx* c) {
  vb_status s = sync_scalars(c);
  if (s != VB_OK) return s;
  int e = c->scal_host->err;
  if (e & kErrAllBinned) { c->err = "all fluid particles are in the bin"; return VB_ERR_ALL_BINNED; }
  if (e & kErrOutOfGrid) { c->err = "particle outside the cell grid"; return VB_ERR_OUT_OF_GRID; }
  if (e & kErrCellOverflow) { c->err = "more than nplink particles in a cell"; return VB_ERR_CELL_OVERFLOW; }
  if (e & kErrCapacity) { c->err = "neighbour or mirror capacity exceeded"; return VB_ERR_CAPACITY; }
  return VB_OK;
}

// predictor_mod.F90:357,369-371 perturbation: minstd stream from its default state, two draws per
// real(8) value, one value per particle in index order, rescaled to 1e13*epsilon*(r - 1/2).
static std::vector<double> make_kick(int n) {
  std::vector<double> k(n);
  uint64_t x = 1;
  const uint64_t m = 2147483647ull;
  for (int i = 0; i < n; i++) {
    x = (x * 48271ull) % m; uint64_t a = x;
    x = (x * 48271ull) % m; uint64_t b = (x - 1) & ((1ull << 30) - 1);
    uint64_t f = ((a << 30) | b) >> 7;
    double r = ldexp((double)f, -54);
    k[i] = (1.0e13 * kEps) * (r - 0.5);
  }
  return k;
}

template <typename T>
static vb_status dalloc(vb_ctx* c, T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  VB_HIP(c, hipMalloc((void**)p, n * sizeof(T)));
  VB_HIP(c, hipMemsetAsync(*p, 0, n * sizeof(T), c->stream));
  return VB_OK;
}

static void free_all(vb_ctx* c) {
  void* ptrs[] = {c->st.x, c->st.y, c->st.u, c->st.v, c->st.th, c->st.t11, c->st.t12, c->st.t21,
                  c->st.t22, c->st.trc, c->st.p, c->st.inbin, c->st.x0, c->st.y0, c->st.u0, c->st.v0,
                  c->st.irel, c->st.vrel, c->st.dP, c->u_alt, c->v_alt, c->th_alt, c->nsurf,
                  c->seg_counts, c->cellkey, c->cellslot, c->nc, c->ist, c->ip, c->scan_bsum, c->cnt, c->link,
                  c->wg1, c->wg2, c->wl, c->wg1f, c->wg2f, c->wlf, c->k11, c->k12, c->k21, c->k22,
                  c->usx, c->usy, c->kick, c->gpx, c->gpy, c->vort, c->conc, c->snx, c->sny, c->pscale, c->a_diag, c->a_val, c->rhs,
                  c->a_col, c->a_len, c->rowflag, c->chunkb, c->xflag, c->xscan, c->xrows, c->xcol, c->xval,
                  c->boff, c->bv64, c->bv32, c->bpos,
                  c->r, c->rb, c->vv,
                  c->pn, c->pinv, c->ss, c->sinv, c->tt,
                  c->xsys, c->partials, c->red, c->scal};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  free_operator(c);
  free_adjoint(c);
  if (c->scal_host) (void)hipHostFree(c->scal_host);
  if (c->stream) (void)hipStreamDestroy(c->stream);
}

static vb_status build_geom(vb_ctx* c, const vb_params* prm, const vb_boundary* b) {
  Geom& g = c->g;
  memset(&g, 0, sizeof g);
  if (b->nb_patches < 1 || b->nb_patches > VB_MAX_PATCHES || b->nb_circles < 0 ||
      b->nb_circles > VB_MAX_CIRCLES) {
    c->err = "unsupported number of boundary patches or circles";
    return VB_ERR_ARG;
  }
  g.nbp = b->nb_patches; g.ncirc = b->nb_circles; g.npfb = c->npfb; g.nbS = c->nbS;
  g.open = prm->i_open_domain; g.shift_coeff = prm->i_shift_coeff;
  for (int i = 0; i < g.nbp; i++) {
    g.bnode[i][0] = b->b_node[2 * i]; g.bnode[i][1] = b->b_node[2 * i + 1];
    g.btype[i] = b->b_type[i]; g.bvel[i] = b->b_vel[i]; g.bth[i] = b->b_thermal[i];
    g.bpp[i] = b->b_type[i] == 2 ? b->b_periodic_parent[i] - 1 : 0;
    if (g.btype[i] < 0 || g.btype[i] > 2) {
      c->err = "only wall (1), periodic (2) and non-boundary (0) patches are supported";
      return VB_ERR_ARG;
    }
    if (g.btype[i] == 2 && (g.bpp[i] < 0 || g.bpp[i] >= g.nbp)) { c->err = "bad b_periodic_parent"; return VB_ERR_ARG; }
  }
  for (int i = 0; i < g.nbp; i++) {                 // datclass.F90:730-744
    int ip1 = (i + 1) % g.nbp;
    g.bedge[i][0] = g.bnode[ip1][0] - g.bnode[i][0];
    g.bedge[i][1] = g.bnode[ip1][1] - g.bnode[i][1];
  }
  for (int i = 0; i < g.nbp; i++) {                 // bound_norm, mirror_boundaries_mod.F90:464-474
    double t = sqrt(g.bedge[i][0] * g.bedge[i][0] + g.bedge[i][1] * g.bedge[i][1]);
    g.blen[i] = t;
    g.bnorm[i][0] = -1.0 * g.bedge[i][1] / t;
    g.bnorm[i][1] = g.bedge[i][0] / t;
  }
  for (int i = 0; i < g.nbp; i++) {                 // corner angles, input.F90:255-273
    int im1 = (i + g.nbp - 1) % g.nbp, ip1 = (i + 1) % g.nbp;
    double pd = g.bedge[im1][1] * g.bnode[ip1][0] - g.bedge[im1][0] * g.bnode[ip1][1] +
                g.bnode[i][0] * g.bnode[im1][1] - g.bnode[i][1] * g.bnode[im1][0];
    pd = -1.0 * pd / sqrt(g.bedge[im1][0] * g.bedge[im1][0] + g.bedge[im1][1] * g.bedge[im1][1]);
    double t1 = g.blen[i];
    if (pd <= 0.0 && fabs(pd) > fabs(t1)) pd = -t1;
    else if (pd > 0.0 && fabs(pd) > fabs(t1)) pd = t1;
    g.bang[i] = asin(pd / t1);
  }
  for (int i = 0; i < g.ncirc; i++) {
    g.ccen[i][0] = b->c_centre[2 * i]; g.ccen[i][1] = b->c_centre[2 * i + 1];
    g.crad[i] = b->c_radius[i]; g.comg[i] = b->c_omega[i];
    g.cvel[i][0] = b->c_vel[2 * i]; g.cvel[i][1] = b->c_vel[2 * i + 1]; g.cth[i] = b->c_thermal[i];
  }
  // constants, input.F90:164-172 and kernel_calculation.F90:37-38
  const double h = prm->h, h2 = h * h;
  g.h = h; g.dx = prm->dx; g.dv = prm->dv;
  g.sup = 3.0 * h; g.sup2 = 9.0 * h2; g.uno = 1.0 / (3.0 * h);
  g.ad7 = 7.0 / 478.0 / kPi / h / h; g.ad7h = g.ad7 / h;
  g.eta2 = h * h * 1.0e-10;
  g.eta_small = (double)0.01f * h * h;
  g.oneoverh = 1.0 / h;
  g.bf[0] = prm->body_force[0]; g.bf[1] = prm->body_force[1];
  // padded box and cell grid, sph2D_incom.F90:64-70, cell_griddin.F90:13-24
  double xb0 = g.bnode[0][0], xb1 = xb0, yb0 = g.bnode[0][1], yb1 = yb0;
  for (int i = 1; i < g.nbp; i++) {
    xb0 = std::min(xb0, g.bnode[i][0]); xb1 = std::max(xb1, g.bnode[i][0]);
    yb0 = std::min(yb0, g.bnode[i][1]); yb1 = std::max(yb1, g.bnode[i][1]);
  }
  g.xmint = xb0 - g.sup * 1.5; g.xmaxt = xb1 + g.sup * 1.5;
  g.ymint = yb0 - g.sup * 1.5; g.ymaxt = yb1 + g.sup * 1.5;
  g.ncx = (int)((g.xmaxt - g.xmint) * g.uno) + 1;
  g.ncy = (int)((g.ymaxt - g.ymint) * g.uno) + 1;
  g.nct = g.ncx * g.ncy;
  return VB_OK;
}

}  // namespace vb

using namespace vb;

extern "C" {

vb_status vb_create(const vb_params* prm, const vb_boundary* bnd, int32_t npfb, int32_t nbS, vb_ctx** out) {
  if (!prm || !bnd || !out || npfb < 1 || nbS < 0 || nbS > npfb) return VB_ERR_ARG;
  *out = nullptr;
  // Slot indices k*npfb + i (k < VB_NPLINK) and the transposed map's keys i*VB_NPLINK + k are 32-bit:
  // more than (2^31 - 1) / VB_NPLINK = 28,633,115 particles per context would wrap them.  Checked
  // before any device call.
  if ((int64_t)npfb * kNplink > (int64_t)INT32_MAX) return VB_ERR_CAPACITY;
  vb_ctx* c = new vb_ctx();
  memset((void*)&c->st, 0, sizeof c->st);
  c->prm = *prm;
  c->npfb = npfb; c->nbS = nbS;
  c->mcap = (npfb <= 500 ? 10 : 2) * npfb;          // nmirror_esti, mirror_boundaries_mod.F90:53-54
  c->cap = npfb + c->mcap;
  c->itime = prm->itime;
  c->time = prm->time;
  vb_status s = VB_OK;
  if ((prm->i_open_domain != 0 && prm->i_open_domain != 1) || (prm->i_shift_coeff != 0 && prm->i_shift_coeff != 1)) {
    c->err = "i_open_domain and i_shift_coeff must be 0 or 1";
    s = VB_ERR_ARG;
  }
  if (prm->fluid_model < 0 || prm->fluid_model > 2) {
    c->err = "fluid_model must be 0 or 2 (viscoelastic) or 1 (Newtonian)";
    s = VB_ERR_ARG;
  }
  if (s == VB_OK) s = build_geom(c, prm, bnd);
  if (s == VB_OK) s = hip_check(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "stream");
  const size_t cap = c->cap, n = npfb, nw = (size_t)npfb * kNplink;
  State& st = c->st;
#define A(p, cnt) if (s == VB_OK) s = dalloc(c, &(p), (cnt))
  A(st.x, cap); A(st.y, cap); A(st.u, cap); A(st.v, cap); A(st.th, cap); A(st.t11, cap); A(st.t12, cap);
  A(st.t21, cap); A(st.t22, cap); A(st.trc, cap); A(st.p, cap); A(st.inbin, cap);
  A(st.x0, n); A(st.y0, n); A(st.u0, n); A(st.v0, n); A(st.irel, cap); A(st.vrel, cap); A(st.dP, cap);
  A(c->u_alt, cap); A(c->v_alt, cap); A(c->th_alt, cap); A(c->nsurf, cap);
  c->seg_blocks = grid_1d(std::max<long>((long)cap, (long)npfb * std::max(1, c->g.ncirc)));
  A(c->seg_counts, c->seg_blocks);
  A(c->cellkey, cap); A(c->cellslot, cap); A(c->nc, c->g.nct + 1); A(c->ist, c->g.nct + 1); A(c->ip, cap);
  A(c->scan_bsum, (std::max<size_t>(cap, (size_t)c->g.nct) + 1023) / 1024 + 1);
  A(c->cnt, n); A(c->link, nw);
  // kernel weights: FP32 planes on the mixed path (the SPH sums read them widened), else FP64
  c->wf32 = prm->precision == VB_PREC_MIXED;
  if (c->wf32) { A(c->wg1f, nw); A(c->wg2f, nw); A(c->wlf, nw); }
  else { A(c->wg1, nw); A(c->wg2, nw); A(c->wl, nw); }
  A(c->k11, n); A(c->k12, n); A(c->k21, n); A(c->k22, n);
  A(c->usx, n); A(c->usy, n); A(c->kick, n); A(c->gpx, n); A(c->gpy, n); A(c->vort, n); A(c->conc, n);
  if (prm->i_open_domain == 1) { A(c->snx, n); A(c->sny, n); A(c->pscale, n); }
  A(c->a_diag, n); A(c->a_val, nw); A(c->rhs, n); A(c->a_col, nw); A(c->a_len, n);
  // compact rows: kMaxBlk column blocks for every row of every chunk; the full rows' list-order copy
  // for every row
  A(c->chunkb, ppe_chunks(npfb) + 1);
  A(c->xflag, n + 1); A(c->xscan, n + 1); A(c->xrows, n);
  A(c->xcol, ppe_list_slots(npfb)); A(c->xval, ppe_list_slots(npfb));
  A(c->rowflag, n); A(c->boff, ppe_block_offsets(npfb)); A(c->bv64, ppe_block_values(npfb)); A(c->bpos, nw);
  // VB_SPMV=stored: every FP64 SpMV row from the stored matrix (measurement switch; same results).
  const char* sp = getenv("VB_SPMV");
  c->spmv_stored = sp && strcmp(sp, "stored") == 0;
  // VB_SPMV=stream: the mixed SpMVs stream their FP32 values whatever the plane's size (measurement
  // switch; the same results, bit for bit)
  c->spmv_stream = sp && strcmp(sp, "stream") == 0;
  if (prm->precision == VB_PREC_MIXED) A(c->bv32, ppe_block_values(npfb));
  A(c->r, n); A(c->rb, n); A(c->vv, n); A(c->pn, n); A(c->pinv, n); A(c->ss, n); A(c->sinv, n);
  A(c->tt, n); A(c->xsys, n);
  A(c->partials, (size_t)kSolverStages * 2 * kPartialStride); A(c->scal, 1);
  if (prm->serial_reductions) A(c->red, (size_t)kSolverStages * 2 * n);
#undef A
  if (s == VB_OK) s = hip_check(c, hipHostMalloc((void**)&c->scal_host, sizeof(Scal)), "hipHostMalloc");
  if (s == VB_OK) {
    // trace_conf starts at 2 (input.F90:111)
    std::vector<double> two(cap, 2.0);
    s = hip_check(c, hipMemcpyAsync(st.trc, two.data(), cap * 8, hipMemcpyHostToDevice, c->stream), "trc");
    std::vector<double> kick = make_kick(npfb);
    if (s == VB_OK)
      s = hip_check(c, hipMemcpyAsync(c->kick, kick.data(), n * 8, hipMemcpyHostToDevice, c->stream), "kick");
    Scal h0;
    memset(&h0, 0, sizeof h0);
    h0.wab1 = wab(prm->dx / prm->h, c->g.ad7);            // new_fick_shift.F90:132-133
    h0.np = npfb;
    if (s == VB_OK) s = hip_check(c, hipMemcpyAsync(c->scal, &h0, sizeof h0, hipMemcpyHostToDevice, c->stream), "scal");
    if (s == VB_OK) s = hip_check(c, hipStreamSynchronize(c->stream), "sync");
  }
  // set_time_step's viscous limit, div_free.F90:142
  double a = sqrt(prm->Ra / prm->Pr), b = sqrt(prm->Ra * prm->Pr);
  c->dt_visc = prm->dt_coef_visc * prm->h * prm->h * (a < b ? a : b);
  c->np_host = npfb; c->nmirror_host = 0;
  if (s != VB_OK) {
    free_all(c);
    delete c;
    return s;
  }
  *out = c;
  return VB_OK;
}

vb_status vb_destroy(vb_ctx* c) {
  if (!c) return VB_ERR_ARG;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_all(c);
  delete c;
  return VB_OK;
}

const char* vb_last_error(const vb_ctx* c) { return c ? c->err.c_str() : "null context"; }

vb_status vb_upload_state(vb_ctx* c, int32_t n, const double* x, const double* y, const double* u,
                          const double* v, const double* theta, const double* t11, const double* t12,
                          const double* t21, const double* t22, const double* trc, const double* p,
                          const int32_t* inbin) {
  if (!c || n < 0 || n > c->cap) return VB_ERR_ARG;
  const size_t b = (size_t)n * 8;
  const double* src[] = {x, y, u, v, theta, t11, t12, t21, t22, trc, p};
  double* dst[] = {c->st.x, c->st.y, c->st.u, c->st.v, c->st.th, c->st.t11, c->st.t12, c->st.t21,
                   c->st.t22, c->st.trc, c->st.p};
  for (int q = 0; q < 11; q++)
    if (src[q]) VB_HIP(c, hipMemcpyAsync(dst[q], src[q], b, hipMemcpyHostToDevice, c->stream));
  if (inbin) {
    std::vector<uint8_t> ib(n);
    for (int i = 0; i < n; i++) ib[i] = inbin[i] ? 1 : 0;
    VB_HIP(c, hipMemcpyAsync(c->st.inbin, ib.data(), n, hipMemcpyHostToDevice, c->stream));
    VB_HIP(c, hipStreamSynchronize(c->stream));
  }
  VB_HIP(c, hipStreamSynchronize(c->stream));
  return VB_OK;
}

vb_status vb_download_state(vb_ctx* c, int32_t n, double* x, double* y, double* u, double* v,
                            double* theta, double* t11, double* t12, double* t21, double* t22,
                            double* trc, double* p, int32_t* inbin) {
  if (!c || n < 0 || n > c->cap) return VB_ERR_ARG;
  const size_t b = (size_t)n * 8;
  double* dst[] = {x, y, u, v, theta, t11, t12, t21, t22, trc, p};
  const double* src[] = {c->st.x, c->st.y, c->st.u, c->st.v, c->st.th, c->st.t11, c->st.t12,
                         c->st.t21, c->st.t22, c->st.trc, c->st.p};
  for (int q = 0; q < 11; q++)
    if (dst[q]) VB_HIP(c, hipMemcpyAsync(dst[q], src[q], b, hipMemcpyDeviceToHost, c->stream));
  std::vector<uint8_t> ib(inbin ? n : 0);
  if (inbin) VB_HIP(c, hipMemcpyAsync(ib.data(), c->st.inbin, n, hipMemcpyDeviceToHost, c->stream));
  VB_HIP(c, hipStreamSynchronize(c->stream));
  for (int i = 0; inbin && i < n; i++) inbin[i] = ib[i];
  return VB_OK;
}

vb_status vb_set_time_step(vb_ctx* c) { return c ? launch_set_time_step(c) : VB_ERR_ARG; }
vb_status vb_put_in_bin(vb_ctx* c) { return c ? launch_put_in_bin(c) : VB_ERR_ARG; }
vb_status vb_create_mirror_particles(vb_ctx* c) { return c ? launch_mirrors(c) : VB_ERR_ARG; }

vb_status vb_particle_allocation(vb_ctx* c) {
  if (!c) return VB_ERR_ARG;
  vb_status s = launch_neighbours(c);
  return s != VB_OK ? s : check_err_bits(c);
}

vb_status vb_kernel_calculation(vb_ctx* c) { return c ? launch_weights(c) : VB_ERR_ARG; }
vb_status vb_locate_free_surf_part(vb_ctx* c) { return c ? launch_freesurf(c) : VB_ERR_ARG; }
vb_status vb_new_fick_shift(vb_ctx* c) { return c ? launch_shift(c) : VB_ERR_ARG; }
vb_status vb_prediction_step(vb_ctx* c) { return c ? launch_predict(c) : VB_ERR_ARG; }
vb_status vb_mirror_velocities(vb_ctx* c) { return c ? launch_mirror_vel(c) : VB_ERR_ARG; }

vb_status vb_ppe_solve(vb_ctx* c, double tol, int32_t itermax, int32_t* iters, double* residual) {
  if (!c) return VB_ERR_ARG;
  if (tol <= 0.0) tol = 1.0e-8;                   // ppe_...:143-144
  if (itermax <= 0) itermax = 1000;
  vb_status s = launch_ppe(c, tol, itermax);
  if (s != VB_OK) return s;
  if (iters || residual) {
    if ((s = sync_scalars(c)) != VB_OK) return s;
    if (iters) *iters = c->scal_host->iter;
    if (residual) *residual = c->scal_host->resid;
  }
  return (vb_status)c->scal_host->status;
}

static vb_status caller_system(vb_ctx* c, const double* A, const int32_t* ija, double* x, const double* B,
                               int32_t n, int32_t nmax, int32_t itermax, double tol, int32_t* iter,
                               double* residual, bool open) {
  if (!c || !A || !ija || !x || !B || n < 1 || nmax < n + 1) return VB_ERR_ARG;
  if (ija[0] != n + 2) { c->err = "Sparse storage mode is not configured properly"; return VB_ERR_ARG; }
  if (ija[n] > nmax + 1) { c->err = "ija(n+1) beyond nmax"; return VB_ERR_ARG; }
  vb_status s = solve_nr(c, A, ija, x, B, n, itermax, tol, open);
  if (s != VB_OK) return s;
  if (iter) *iter = c->scal_host->iter;
  if (residual) *residual = c->scal_host->resid;
  return (vb_status)c->scal_host->status;
}

vb_status vb_sparse_pbicgstab_nonull(vb_ctx* c, const double* A, const int32_t* ija, double* x,
                                     const double* B, int32_t n, int32_t nmax, int32_t itermax,
                                     double tol, int32_t* iter, double* residual) {
  return caller_system(c, A, ija, x, B, n, nmax, itermax, tol, iter, residual, false);
}

vb_status vb_sparse_pbicgstab(vb_ctx* c, const double* A, const int32_t* ija, double* x, const double* B,
                              int32_t n, int32_t nmax, int32_t itermax, double tol, int32_t* iter,
                              double* residual) {
  return caller_system(c, A, ija, x, B, n, nmax, itermax, tol, iter, residual, true);
}

vb_status vb_pressure_update(vb_ctx* c) { return c ? launch_pressure_update(c) : VB_ERR_ARG; }

vb_status vb_final_advection(vb_ctx* c) {
  if (!c) return VB_ERR_ARG;
  vb_status s = launch_final(c);
  if (s != VB_OK) return s;
  if ((s = sync_scalars(c)) != VB_OK) return s;
  c->time = c->time + c->scal_host->dt;          // sph2D_incom.F90:98
  return VB_OK;
}

vb_status vb_div_free(vb_ctx* c) {
  if (!c) return VB_ERR_ARG;
  vb_status s;
  if ((s = vb_set_time_step(c)) != VB_OK) return s;
  if ((s = vb_put_in_bin(c)) != VB_OK) return s;
  if ((s = vb_create_mirror_particles(c)) != VB_OK) return s;
  if ((s = vb_particle_allocation(c)) != VB_OK) return s;
  if ((s = vb_kernel_calculation(c)) != VB_OK) return s;
  if ((s = vb_locate_free_surf_part(c)) != VB_OK) return s;
  if ((s = vb_new_fick_shift(c)) != VB_OK) return s;
  if ((s = vb_prediction_step(c)) != VB_OK) return s;
  if ((s = vb_mirror_velocities(c)) != VB_OK) return s;
  vb_status ls = vb_ppe_solve(c, 1.0e-8, 1000, nullptr, nullptr);
  if (ls != VB_OK && ls != VB_WARN_BICGSTAB_BREAKDOWN && ls != VB_WARN_MAXITER) return ls;
  if ((s = vb_pressure_update(c)) != VB_OK) return s;
  if ((s = vb_final_advection(c)) != VB_OK) return s;
  return ls;
}

vb_status vb_get_scalars(vb_ctx* c, int32_t* ints, double* dbls) {
  if (!c) return VB_ERR_ARG;
  vb_status s = sync_scalars(c);
  if (s != VB_OK) return s;
  const Scal& h = *c->scal_host;
  int32_t iv[12] = {c->npfb + h.nmirror, c->npfb, c->nbS, h.nmirror, h.maxneigh, h.iter, c->itime,
                    h.n_inbin, c->g.ncx, c->g.ncy, c->g.nct, h.status};
  double dv[8] = {h.dt, h.resid, c->time, h.umax, c->g.xmint, c->g.xmaxt, c->g.ymint, c->g.ymaxt};
  if (ints) memcpy(ints, iv, sizeof iv);
  if (dbls) memcpy(dbls, dv, sizeof dv);
  return VB_OK;
}

vb_status vb_profile_spmv(vb_ctx* c, int32_t reps, double* avg_us, double* alg_bytes) {
  if (!c || reps < 1 || !avg_us || !alg_bytes) return VB_ERR_ARG;
  return profile_spmv(c, reps, avg_us, alg_bytes);
}

// Four 16-byte non-temporal loads in flight per thread, then four stores: the fastest plain copy
// measured on MI355X (6.2 TB/s; a grid-stride copy of single loads reaches 4.6).
typedef double cal_d2 __attribute__((ext_vector_type(2)));
__global__ void vb_copy_calibration(const double2* __restrict__ src, double2* __restrict__ dst, size_t n2) {
  const cal_d2* s = reinterpret_cast<const cal_d2*>(src);
  cal_d2* d = reinterpret_cast<cal_d2*>(dst);
  const size_t base = (size_t)blockIdx.x * blockDim.x * 4 + threadIdx.x;
  cal_d2 v[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const size_t i = base + (size_t)u * blockDim.x;
    if (i < n2) v[u] = __builtin_nontemporal_load(s + i);
  }
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const size_t i = base + (size_t)u * blockDim.x;
    if (i < n2) __builtin_nontemporal_store(v[u], d + i);
  }
}

vb_status vb_profile_copy(vb_ctx* c, int64_t bytes, int32_t reps, double* avg_us) {
  if (!c || reps < 1 || !avg_us || bytes < 16 || bytes % 16 != 0) return VB_ERR_ARG;
  void *src = nullptr, *dst = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  vb_status s = VB_OK;
  const size_t n2 = (size_t)bytes / 16;
  float ms = 0.f;
  do {
    if ((s = hip_check(c, hipMalloc(&src, bytes), "hipMalloc")) != VB_OK) break;
    if ((s = hip_check(c, hipMalloc(&dst, bytes), "hipMalloc")) != VB_OK) break;
    if ((s = hip_check(c, hipMemsetAsync(src, 0, bytes, c->stream), "hipMemsetAsync")) != VB_OK) break;
    if ((s = hip_check(c, hipEventCreate(&e0), "hipEventCreate")) != VB_OK) break;
    if ((s = hip_check(c, hipEventCreate(&e1), "hipEventCreate")) != VB_OK) break;
    for (int r = 0; r <= reps && s == VB_OK; r++) {
      if (r == 1) s = hip_check(c, hipEventRecord(e0, c->stream), "hipEventRecord");
      hipLaunchKernelGGL(vb_copy_calibration, dim3((unsigned)((n2 + 4 * kBlock - 1) / (4 * kBlock))), dim3(kBlock),
                         0, c->stream, (const double2*)src, (double2*)dst, n2);
      if (s == VB_OK) s = hip_check(c, hipGetLastError(), "vb_copy_calibration");
    }
    if (s != VB_OK) break;
    if ((s = hip_check(c, hipEventRecord(e1, c->stream), "hipEventRecord")) != VB_OK) break;
    if ((s = hip_check(c, hipEventSynchronize(e1), "hipEventSynchronize")) != VB_OK) break;
    s = hip_check(c, hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
  } while (0);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (src) (void)hipFree(src);
  if (dst) (void)hipFree(dst);
  if (s == VB_OK) *avg_us = 1000.0 * ms / reps;
  return s;
}

vb_status vb_export_mirrors(vb_ctx* c, int32_t m, int32_t* irel, int32_t* vrel, double* dP) {
  if (!c || m < 0 || m > c->mcap) return VB_ERR_ARG;
  VB_HIP(c, hipMemcpyAsync(irel, c->st.irel + c->npfb, (size_t)m * 4, hipMemcpyDeviceToHost, c->stream));
  VB_HIP(c, hipMemcpyAsync(vrel, c->st.vrel + c->npfb, (size_t)m * 4, hipMemcpyDeviceToHost, c->stream));
  VB_HIP(c, hipMemcpyAsync(dP, c->st.dP + c->npfb, (size_t)m * 8, hipMemcpyDeviceToHost, c->stream));
  VB_HIP(c, hipStreamSynchronize(c->stream));
  for (int j = 0; j < m; j++) irel[j] += 1;
  return VB_OK;
}

vb_status vb_export_cells(vb_ctx* c, int32_t np, int32_t* cellpart, int32_t nct1, int32_t* ist, int32_t* ip) {
  if (!c || np != c->np_host || nct1 != c->g.nct + 1 || !cellpart || !ist || !ip) return VB_ERR_ARG;
  VB_HIP(c, hipMemcpyAsync(cellpart, c->cellkey, (size_t)np * 4, hipMemcpyDeviceToHost, c->stream));
  VB_HIP(c, hipMemcpyAsync(ist, c->ist, (size_t)nct1 * 4, hipMemcpyDeviceToHost, c->stream));
  VB_HIP(c, hipMemcpyAsync(ip, c->ip, (size_t)np * 4, hipMemcpyDeviceToHost, c->stream));
  VB_HIP(c, hipStreamSynchronize(c->stream));
  const int32_t binned = ist[nct1 - 1];
  for (int32_t k = 0; k < np; k++) {
    cellpart[k] = cellpart[k] < 1 ? 0 : cellpart[k];
    ip[k] = k < binned ? ip[k] + 1 : 0;
  }
  for (int32_t q = 0; q < nct1; q++) ist[q] += 1;
  return VB_OK;
}

vb_status vb_export_neighbours(vb_ctx* c, int32_t n, int32_t w, int32_t* cnt, int32_t* link) {
  if (!c || n != c->npfb || w < 1 || w > kNplink) return VB_ERR_ARG;
  std::vector<int32_t> l((size_t)n * w);
  VB_HIP(c, hipMemcpyAsync(cnt, c->cnt, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
  VB_HIP(c, hipMemcpyAsync(l.data(), c->link, (size_t)n * w * 4, hipMemcpyDeviceToHost, c->stream));
  VB_HIP(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < w; k++)
    for (int i = 0; i < n; i++) {
      size_t s = (size_t)k * n + i;
      link[s] = k < cnt[i] ? l[s] + 1 : 0;
    }
  return VB_OK;
}

vb_status vb_export_weights(vb_ctx* c, int32_t n, int32_t w, double* wG, double* wL, double* kg) {
  if (!c || n != c->npfb || w < 1 || w > kNplink) return VB_ERR_ARG;
  std::vector<int32_t> cnt(n);
  const size_t nw = (size_t)n * w;
  VB_HIP(c, hipMemcpyAsync(cnt.data(), c->cnt, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
  std::vector<float> wf(c->wf32 ? 3 * nw : 0);            // FP32 planes, widened below
  if (c->wf32) {
    VB_HIP(c, hipMemcpyAsync(wf.data(), c->wg1f, nw * 4, hipMemcpyDeviceToHost, c->stream));
    VB_HIP(c, hipMemcpyAsync(wf.data() + nw, c->wg2f, nw * 4, hipMemcpyDeviceToHost, c->stream));
    VB_HIP(c, hipMemcpyAsync(wf.data() + 2 * nw, c->wlf, nw * 4, hipMemcpyDeviceToHost, c->stream));
  } else {
    VB_HIP(c, hipMemcpyAsync(wG, c->wg1, nw * 8, hipMemcpyDeviceToHost, c->stream));
    VB_HIP(c, hipMemcpyAsync(wG + nw, c->wg2, nw * 8, hipMemcpyDeviceToHost, c->stream));
    VB_HIP(c, hipMemcpyAsync(wL, c->wl, nw * 8, hipMemcpyDeviceToHost, c->stream));
  }
  double* kk[4] = {c->k11, c->k12, c->k21, c->k22};
  for (int q = 0; q < 4; q++)
    VB_HIP(c, hipMemcpyAsync(kg + (size_t)q * n, kk[q], (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
  VB_HIP(c, hipStreamSynchronize(c->stream));
  if (c->wf32)
    for (size_t s = 0; s < nw; s++) { wG[s] = wf[s]; wG[nw + s] = wf[nw + s]; wL[s] = wf[2 * nw + s]; }
  for (int k = 0; k < w; k++)
    for (int i = 0; i < n; i++)
      if (k >= cnt[i]) {
        size_t s = (size_t)k * n + i;
        wG[s] = 0.0; wG[nw + s] = 0.0; wL[s] = 0.0;
      }
  return VB_OK;
}

vb_status vb_export_ushift(vb_ctx* c, int32_t n, double* us) {
  if (!c || n != c->npfb) return VB_ERR_ARG;
  VB_HIP(c, hipMemcpyAsync(us, c->usx, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
  VB_HIP(c, hipMemcpyAsync(us + n, c->usy, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
  VB_HIP(c, hipStreamSynchronize(c->stream));
  return VB_OK;
}

vb_status vb_secondary_properties(vb_ctx* c, int32_t n, double* vort, double* u_rms) {
  if (!c || n < 0 || n > c->npfb) return VB_ERR_ARG;
  vb_status s = launch_secondary_properties(c);
  if (s != VB_OK) return s;
  if (vort && n > 0)
    VB_HIP(c, hipMemcpyAsync(vort, c->vort, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
  if ((s = sync_scalars(c)) != VB_OK) return s;
  if (u_rms) *u_rms = c->scal_host->u_rms;
  return VB_OK;
}

vb_status vb_ppe_operator(vb_ctx* c, int32_t transpose, int32_t n, const double* x, double* y) {
  if (!c || n != c->npfb || !x || !y) {
    if (c) c->err = "vb_ppe_operator: n must equal the context's npfb";
    return VB_ERR_ARG;
  }
  if (transpose < 0 || transpose > 3) {
    c->err = "vb_ppe_operator: transpose must be 0..3";
    return VB_ERR_ARG;
  }
  const size_t bytes = (size_t)n * sizeof(double);
  VB_HIP(c, hipMemcpyAsync(c->vv, x, bytes, hipMemcpyHostToDevice, c->stream));
  vb_status s = transpose >= 2 ? apply_solver_spmv(c, transpose == 2, c->vv, c->tt)
                               : apply_ppe_operator(c, transpose != 0, c->vv, c->tt);
  if (s != VB_OK) return s;
  VB_HIP(c, hipMemcpyAsync(y, c->tt, bytes, hipMemcpyDeviceToHost, c->stream));
  VB_HIP(c, hipStreamSynchronize(c->stream));
  return VB_OK;
}

vb_status vb_grad_operator(vb_ctx* c, int32_t transpose, int32_t n, const double* x, double* y) {
  if (!c || n != c->npfb || !x || !y) {
    if (c) c->err = "vb_grad_operator: n must equal the context's npfb";
    return VB_ERR_ARG;
  }
  const size_t n1 = (size_t)n * sizeof(double), n2 = 2 * n1;
  if (!c->op.gin) {
    VB_HIP(c, hipMalloc((void**)&c->op.gin, n2));
    VB_HIP(c, hipMalloc((void**)&c->op.gout, n2));
  }
  VB_HIP(c, hipMemcpyAsync(c->op.gin, x, transpose ? n2 : n1, hipMemcpyHostToDevice, c->stream));
  vb_status s = apply_grad_operator(c, transpose != 0, c->op.gin, c->op.gout);
  if (s != VB_OK) return s;
  VB_HIP(c, hipMemcpyAsync(y, c->op.gout, transpose ? n1 : n2, hipMemcpyDeviceToHost, c->stream));
  VB_HIP(c, hipStreamSynchronize(c->stream));
  return VB_OK;
}

vb_status vb_ppe_adjoint_solve(vb_ctx* c, int32_t n, const double* g, double* lambda, double tol,
                               int32_t itermax, int32_t* iters, double* residual) {
  if (!c || n != c->npfb || !g || !lambda) {
    if (c) c->err = "vb_ppe_adjoint_solve: n must equal the context's npfb";
    return VB_ERR_ARG;
  }
  if (tol <= 0.0) tol = 1.0e-8;
  if (itermax <= 0) itermax = 1000;
  const size_t bytes = (size_t)n * sizeof(double);
  if (!c->op.g) {
    vb_status sa = alloc_operator(c, 0);
    if (sa != VB_OK) return sa;
  }
  VB_HIP(c, hipMemcpyAsync(c->op.g, g, bytes, hipMemcpyHostToDevice, c->stream));
  int it = 0, st = VB_OK;
  double res = 0.0;
  vb_status s = adjoint_solve(c, tol, itermax, &it, &res, &st);
  if (s != VB_OK) return s;
  VB_HIP(c, hipMemcpyAsync(lambda, c->op.lam, bytes, hipMemcpyDeviceToHost, c->stream));
  VB_HIP(c, hipStreamSynchronize(c->stream));
  if (iters) *iters = it;
  if (residual) *residual = res;
  return (vb_status)st;
}

vb_status vb_free_surface(vb_ctx* c, int32_t n, int32_t* n_surf, double* conc) {
  if (!c || n < 0 || n > std::max(c->np_host, c->npfb)) return VB_ERR_ARG;
  const int nr = std::min<int>(n, c->npfb);
  if (n_surf && n > 0)
    VB_HIP(c, hipMemcpyAsync(n_surf, c->nsurf, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
  if (conc && nr > 0)
    VB_HIP(c, hipMemcpyAsync(conc, c->conc, (size_t)nr * 8, hipMemcpyDeviceToHost, c->stream));
  VB_HIP(c, hipStreamSynchronize(c->stream));
  if (conc) for (int i = nr; i < n; i++) conc[i] = 0.0;
  return VB_OK;
}

vb_status vb_circle_forces(vb_ctx* c, int32_t nc, double* force, int32_t* counts) {
  if (!c || nc != c->g.ncirc || (nc > 0 && (!force || !counts))) {
    if (c) c->err = "vb_circle_forces: nc must equal the context's circle count";
    return VB_ERR_ARG;
  }
  if (nc == 0) return VB_OK;
  vb_status s = launch_circle_forces(c);
  if (s != VB_OK) return s;
  if ((s = sync_scalars(c)) != VB_OK) return s;
  for (int ib = 0; ib < nc; ib++) {
    force[2 * ib] = c->scal_host->circ[3 * ib];
    force[2 * ib + 1] = c->scal_host->circ[3 * ib + 1];
    counts[ib] = (int32_t)c->scal_host->circ[3 * ib + 2];
  }
  return VB_OK;
}

vb_status vb_nusselt(vb_ctx* c, double* nu) {
  if (!c || !nu) return VB_ERR_ARG;
  const Geom& g = c->g;
  double ylo = g.bnode[0][1], yhi = ylo, tlo = 0.0, thi = 0.0;
  bool wall = false;
  for (int i = 0; i < g.nbp; i++) {
    ylo = std::min(ylo, g.bnode[i][1]); yhi = std::max(yhi, g.bnode[i][1]);
    if (g.btype[i] != 1) continue;
    if (!wall) { tlo = thi = g.bth[i]; wall = true; }
    tlo = std::min(tlo, g.bth[i]); thi = std::max(thi, g.bth[i]);
  }
  if (!wall || thi - tlo <= 0.0) {
    c->err = "vb_nusselt: the walls carry no temperature difference";
    return VB_ERR_ARG;
  }
  const double scale = sqrt(c->prm.Ra * c->prm.Pr) * (yhi - ylo) / (thi - tlo);
  vb_status s = launch_nusselt(c, scale);
  if (s != VB_OK) return s;
  if ((s = sync_scalars(c)) != VB_OK) return s;
  *nu = c->scal_host->nusselt;
  return VB_OK;
}

vb_status vb_field_maxima(vb_ctx* c, double* maxima) {
  if (!c || !maxima) return VB_ERR_ARG;
  vb_status s = launch_field_maxima(c);
  if (s != VB_OK) return s;
  if ((s = sync_scalars(c)) != VB_OK) return s;
  for (int q = 0; q < 3; q++) maxima[q] = c->scal_host->fmax[q];
  return VB_OK;
}

vb_status vb_export_grad_p(vb_ctx* c, int32_t n, double* gp) {
  if (!c || n != c->npfb) return VB_ERR_ARG;
  VB_HIP(c, hipMemcpyAsync(gp, c->gpx, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
  VB_HIP(c, hipMemcpyAsync(gp + n, c->gpy, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
  VB_HIP(c, hipStreamSynchronize(c->stream));
  return VB_OK;
}

vb_status vb_export_ppe(vb_ctx* c, int32_t n, int32_t w, double* diag, int32_t* col, double* val, double* rhs) {
  if (!c || n != c->npfb || w < 1 || w > kNplink) return VB_ERR_ARG;
  std::vector<int32_t> cnt(n);
  const size_t nw = (size_t)n * w;
  VB_HIP(c, hipMemcpyAsync(cnt.data(), c->cnt, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
  VB_HIP(c, hipMemcpyAsync(diag, c->a_diag, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
  VB_HIP(c, hipMemcpyAsync(col, c->a_col, nw * 4, hipMemcpyDeviceToHost, c->stream));
  VB_HIP(c, hipMemcpyAsync(val, c->a_val, nw * 8, hipMemcpyDeviceToHost, c->stream));
  VB_HIP(c, hipMemcpyAsync(rhs, c->rhs, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
  VB_HIP(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < w; k++)
    for (int i = 0; i < n; i++) {
      size_t s = (size_t)k * n + i;
      if (k < cnt[i]) col[s] += 1; else { col[s] = 0; val[s] = 0.0; }
    }
  return VB_OK;
}

}  // extern "C"
