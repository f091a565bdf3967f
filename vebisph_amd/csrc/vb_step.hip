// gfx950 kernels for the per-step phases of div_free (everything except the PPE).
//
// One thread per particle (or per mirror / candidate / cell); SoA FP64 arrays; neighbour lists and
// pair weights slot-major (link[k*npfb + i]) so that a wavefront reading slot k of 64 consecutive
// particles issues one coalesced load.  Every kernel mirrors one reference loop; each carries the
// file:line it restates.
#include "vb_ctx.h"
#include "vb_logconf.h"

namespace vb {

// ------------------------------------------------------------------------------------------------
// Phase 1: set_time_step (div_free.F90:126-146) and r = r0 + u0 dt (:39-53)
// ------------------------------------------------------------------------------------------------
__global__ void k_umax(State st, int nbS, int npfb, double cfl_h, double dt_visc, Scal* sc,
                       double* partials) {
  double m = 1.0e-10;
  for (int i = nbS + blockIdx.x * kBlock + threadIdx.x; i < npfb; i += gridDim.x * kBlock) {
    double ua = sqrt(st.u[i] * st.u[i] + st.v[i] * st.v[i]);
    if (ua > m) m = ua;
  }
  m = wave_max(m);
  __shared__ double sh[kBlock / 64];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = sh[0];
    for (int w = 1; w < kBlock / 64; w++) b = sh[w] > b ? sh[w] : b;
    partials[blockIdx.x] = b;
  }
}

// max is order-independent: one wavefront folds the partials (strided loads, all in flight at
// once, then a shuffle max); also resets the step's counters
__global__ void k_fin_umax(const double* partials, int nblk, double cfl_h, double dt_visc, Scal* sc) {
  double um = 1.0e-10;
  for (int b = threadIdx.x; b < nblk; b += 64) um = partials[b] > um ? partials[b] : um;
  um = wave_max(um);
  if (threadIdx.x != 0) return;
  sc->umax = um;
  double dt_cfl = cfl_h / um;
  sc->dt = dt_cfl < dt_visc ? dt_cfl : dt_visc;
  sc->n_inbin = 0;
  sc->err = 0;
  sc->nmirror = 0;
  sc->maxneigh = 0;
}

__global__ void k_advect(State st, int nbS, int npfb, const Scal* sc) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= npfb) return;
  double x = st.x[i], y = st.y[i], u = st.u[i], v = st.v[i];
  st.x0[i] = x; st.y0[i] = y; st.u0[i] = u; st.v0[i] = v;
  if (i >= nbS) {
    double dt = sc->dt;
    st.x[i] = x + u * dt;
    st.y[i] = y + v * dt;
  }
}

// Phase 2: put_in_bin.F90:9-54
__global__ void k_put_in_bin(Geom g, State st, int nbS, int npfb, Scal* sc) {
  int i = nbS + blockIdx.x * kBlock + threadIdx.x;
  if (i >= npfb) return;
  bool b = st.inbin[i] != 0;
  double x = st.x[i], y = st.y[i];
  if (!b) {
    if (x < g.xmint || x > g.xmaxt || y < g.ymint || y > g.ymaxt) b = true;
    if (isnan(x) || isnan(y)) b = true;
  }
  if (b) {
    st.inbin[i] = 1;
    st.x[i] = 2.0 * g.xmint - g.xmaxt;
    st.y[i] = 2.0 * g.ymint - g.ymaxt;
    st.u[i] = 0.0; st.v[i] = 0.0;
    atomicAdd(&sc->n_inbin, 1);
  }
}

__global__ void k_check_all_binned(int nfluid, Scal* sc) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && sc->n_inbin == nfluid) sc->err |= kErrAllBinned;
}

// ------------------------------------------------------------------------------------------------
// Phase 3: create_mirror_particles (mirror_boundaries_mod.F90:18-334)
// ------------------------------------------------------------------------------------------------
__device__ inline double bound_dist(const Geom& g, double px, double py, int ib) {   // :447-462
  int ip1 = (ib + 1) % g.nbp;
  double pd = g.bedge[ib][1] * px - g.bedge[ib][0] * py + g.bnode[ip1][0] * g.bnode[ib][1] -
              g.bnode[ip1][1] * g.bnode[ib][0];
  return -1.0 * pd / g.blen[ib];
}

__device__ inline double pos_along(const Geom& g, double px, double py, int ib, double pd) {  // :476-491
  if (fabs(g.bedge[ib][0]) > fabs(g.bedge[ib][1]))
    return (px - g.bnode[ib][0] - pd * g.bnorm[ib][0]) / g.bedge[ib][0];
  return (py - g.bnode[ib][1] - pd * g.bnorm[ib][1]) / g.bedge[ib][1];
}

// Gradient of pos_along(x, ib, bound_dist(x, ib)) in x (bound_dist's gradient is bnorm[ib]).
__device__ inline void grad_pos_along(const Geom& g, int ib, double* d) {
  const double nx = g.bnorm[ib][0], ny = g.bnorm[ib][1];
  if (fabs(g.bedge[ib][0]) > fabs(g.bedge[ib][1])) {
    d[0] = (1.0 - nx * nx) / g.bedge[ib][0]; d[1] = (-nx * ny) / g.bedge[ib][0];
  } else {
    d[0] = (-ny * nx) / g.bedge[ib][1]; d[1] = (1.0 - ny * ny) / g.bedge[ib][1];
  }
}

// Jacobian of the reflection x - 2 bound_dist(x, ib) n_ib across patch ib's line: I - 2 n n^T.
__device__ inline void jac_reflect(const Geom& g, int ib, double* J) {
  const double nx = g.bnorm[ib][0], ny = g.bnorm[ib][1];
  J[0] = 1.0 - 2.0 * nx * nx; J[1] = -2.0 * nx * ny; J[2] = -2.0 * ny * nx; J[3] = 1.0 - 2.0 * ny * ny;
}

// Jacobian of the periodic transfer from patch `from` to patch `to`:
// x' = node_to + e_to - t e_to - pd n_to with t = pos_along(x, from), pd = bound_dist(x, from).
__device__ inline void jac_periodic(const Geom& g, int from, int to, double* J) {
  double dt[2];
  grad_pos_along(g, from, dt);
  const double nfx = g.bnorm[from][0], nfy = g.bnorm[from][1];
  J[0] = -g.bedge[to][0] * dt[0] - g.bnorm[to][0] * nfx; J[1] = -g.bedge[to][0] * dt[1] - g.bnorm[to][0] * nfy;
  J[2] = -g.bedge[to][1] * dt[0] - g.bnorm[to][1] * nfx; J[3] = -g.bedge[to][1] * dt[1] - g.bnorm[to][1] * nfy;
}

__device__ inline void mat2_mul(const double* A, const double* B, double* C) {
  const double c0 = A[0] * B[0] + A[1] * B[2], c1 = A[0] * B[1] + A[1] * B[3];
  const double c2 = A[2] * B[0] + A[3] * B[2], c3 = A[2] * B[1] + A[3] * B[3];
  C[0] = c0; C[1] = c1; C[2] = c2; C[3] = c3;
}

// replace_escapees, :336-407.  Per particle the circle and patch tests run in the reference's order.
// rec.ejac (the adjoint's re-run only): the Jacobian of the particle's move is recorded.
__global__ void k_escapees(Geom g, State st, const Scal* sc, LinRec rec) {
  int i = g.nbS + blockIdx.x * kBlock + threadIdx.x;
  if (i >= g.npfb) return;
  double x = st.x[i], y = st.y[i];
  double J[4] = {1.0, 0.0, 0.0, 1.0};
  int flag = 0;
  for (int ib = 0; ib < g.ncirc; ib++) {
    double R = g.crad[ib];
    double rx = x - g.ccen[ib][0], ry = y - g.ccen[ib][1];
    double rad = sqrt(rx * rx + ry * ry);
    if (rad < fabs(R) && R > 0.0) { x = x + (rx / rad) * 2.0 * (R - rad); y = y + (ry / rad) * 2.0 * (R - rad); flag = 2; }
    if (rad > fabs(R) && R < 0.0) { x = x - (rx / rad) * 2.0 * (rad + R); y = y - (ry / rad) * 2.0 * (rad + R); flag = 2; }
  }
  bool moved = false;
  if (!st.inbin[i]) {
    for (int ib = 0; ib < g.nbp; ib++) {
      int t = g.btype[ib];
      if (t != 1 && t != 2) continue;
      double pd = bound_dist(g, x, y, ib);
      if (pd > 0.0) continue;
      if (fabs(pd) > g.sup + 1.0e-10) continue;
      double tmp = pos_along(g, x, y, ib, pd);
      double M[4];
      if (t == 1) {
        if (tmp < 0.0 || tmp > 1.0) continue;
        x = x - 2.0 * pd * g.bnorm[ib][0];
        y = y - 2.0 * pd * g.bnorm[ib][1];
        jac_reflect(g, ib, M);
      } else {
        int ibp = g.bpp[ib];
        double tx = -1.0 * tmp * g.bedge[ibp][0] - pd * g.bnorm[ibp][0];
        double ty = -1.0 * tmp * g.bedge[ibp][1] - pd * g.bnorm[ibp][1];
        x = g.bnode[ibp][0] + g.bedge[ibp][0] + tx;
        y = g.bnode[ibp][1] + g.bedge[ibp][1] + ty;
        jac_periodic(g, ib, ibp, M);
      }
      mat2_mul(M, J, J);
      if (!flag) flag = 1;
      double dt = sc->dt;
      st.x0[i] = x - st.u[i] * dt; st.y0[i] = y - st.v[i] * dt;
      st.u0[i] = st.u[i]; st.v0[i] = st.v[i];
      moved = true;
    }
  }
  if (g.ncirc || moved) { st.x[i] = x; st.y[i] = y; }
  if (rec.ejac) {
    rec.ejac[4 * i] = J[0]; rec.ejac[4 * i + 1] = J[1]; rec.ejac[4 * i + 2] = J[2]; rec.ejac[4 * i + 3] = J[3];
    rec.eflag[i] = flag;
  }
}

// A mirror segment is one reference loop that appends mirrors in candidate order.  The append order
// is reproduced by an order-preserving compaction: count per block, exclusive scan, emit.
enum SegKind { SEG_CIRCLE, SEG_WALL, SEG_PERIODIC, SEG_CORNER_WW, SEG_CORNER_PP, SEG_CORNER_MIX };
struct Seg { int kind, ib, ibp, ncand; double cross; };
struct MRec { int irel, vrel; double x, y, u, v, th, dP; };

__device__ inline bool near_patch(const Geom& g, double x, double y, int ib, double& bp1, double& bp2) {
  double pd = bound_dist(g, x, y, ib);                     // create_boundary_boxes :420-433
  if (fabs(pd) > g.sup + 1.0e-10) return false;
  double tmp = pos_along(g, x, y, ib, pd);
  if (tmp < 0.0 || tmp > 1.0) return false;
  bp1 = tmp; bp2 = pd;
  return true;
}

__device__ inline bool near_corner(const Geom& g, double x, double y, int ib) {    // :435-441
  double ax = x - g.bnode[ib][0], ay = y - g.bnode[ib][1];
  return !(sqrt(ax * ax + ay * ay) > g.sup + 1.0e-10);
}

__device__ inline double dPdot(const Geom& g, double dx, double dy) { return g.bf[0] * dx + g.bf[1] * dy; }

__device__ bool seg_eval(const Geom& g, const Seg& s, const State& st, int c, int mlimit, MRec* r) {
  switch (s.kind) {
    case SEG_CIRCLE: {                                     // :61-96
      int i = g.nbS + c / g.ncirc, ib = c % g.ncirc;
      double R = g.crad[ib];
      double cdir = R / fabs(R);
      double xi = st.x[i], yi = st.y[i];
      double rx = xi - g.ccen[ib][0], ry = yi - g.ccen[ib][1];
      double rad = sqrt(rx * rx + ry * ry);
      if (cdir > 0.0) {
        if (!(rad <= R + g.sup)) return false;
        if (r) {
          rx = rx / rad; ry = ry / rad;
          r->irel = i; r->vrel = 999 + ib + 1;
          r->x = xi - rx * 2.0 * (rad - R); r->y = yi - ry * 2.0 * (rad - R);
          double tx = ry, ty = -rx;
          r->u = 2.0 * g.cvel[ib][0] + 2.0 * tx * R * g.comg[ib] - 1.0 * st.u[i];
          r->v = 2.0 * g.cvel[ib][1] + 2.0 * ty * R * g.comg[ib] - 1.0 * st.v[i];
          r->th = 2.0 * g.cth[ib] - 1.0 * st.th[i];
          r->dP = dPdot(g, r->x - xi, r->y - yi);
        }
        return true;
      } else if (cdir < 0.0) {
        if (!(rad >= fabs(R) - g.sup)) return false;
        if (r) {
          rx = rx / rad; ry = ry / rad;
          r->irel = i; r->vrel = 999 + ib + 1;
          r->x = xi + rx * 2.0 * (fabs(R) - rad); r->y = yi + ry * 2.0 * (fabs(R) - rad);
          double tx = ry, ty = -rx;
          r->u = 2.0 * g.cvel[ib][0] + 2.0 * tx * fabs(R) * g.comg[ib] - 1.0 * st.u[i];
          r->v = 2.0 * g.cvel[ib][1] + 2.0 * ty * fabs(R) * g.comg[ib] - 1.0 * st.v[i];
          r->th = 2.0 * g.cth[ib] - 1.0 * st.th[i];
          r->dP = dPdot(g, r->x - xi, r->y - yi);
        }
        return true;
      }
      return false;
    }
    case SEG_WALL: {                                       // :106-153
      int i = c, ib = s.ib, ibp1 = (ib + 1) % g.nbp;
      double xi = st.x[i], yi = st.y[i], bp1, bp2;
      if (!near_patch(g, xi, yi, ib, bp1, bp2)) return false;
      if (bp2 <= 1.0e-10) return false;
      if (g.bang[ib] < 0.0) {
        double ta = 0.5 * (3.14159265358979323846 + g.bang[ib]);
        if (bp2 + 0.1 * g.dx >= bp1 * tan(ta) * g.blen[ib]) return false;
      }
      if (g.bang[ibp1] < 0.0) {
        double ta = 0.5 * (3.14159265358979323846 + g.bang[ibp1]);
        if (bp2 >= (1.0 - bp1) * tan(ta) * g.blen[ib]) return false;
      }
      if (r) {
        r->irel = i; r->vrel = ib + 1;
        r->x = xi - 2.0 * bp2 * g.bnorm[ib][0]; r->y = yi - 2.0 * bp2 * g.bnorm[ib][1];
        double tx = g.bnorm[ib][1], ty = -g.bnorm[ib][0];
        r->u = 2.0 * g.bvel[ib] * tx - st.u[i];
        r->v = 2.0 * g.bvel[ib] * ty - st.v[i];
        r->th = 2.0 * g.bth[ib] - st.th[i];
        r->dP = dPdot(g, r->x - xi, r->y - yi);
      }
      return true;
    }
    case SEG_PERIODIC: {                                   // :156-187
      int i = c, ib = s.ib, ibp = s.ibp;
      double bp1, bp2;
      if (!near_patch(g, st.x[i], st.y[i], ibp, bp1, bp2)) return false;
      if (bp2 <= 0.0) return false;
      if (r) {
        r->irel = i; r->vrel = ib + 1;
        double tx = -1.0 * bp1 * g.bedge[ib][0] - bp2 * g.bnorm[ib][0];
        double ty = -1.0 * bp1 * g.bedge[ib][1] - bp2 * g.bnorm[ib][1];
        r->x = g.bnode[ib][0] + g.bedge[ib][0] + tx;
        r->y = g.bnode[ib][1] + g.bedge[ib][1] + ty;
        r->u = st.u[i]; r->v = st.v[i]; r->th = st.th[i]; r->dP = 0.0;
      }
      return true;
    }
    case SEG_CORNER_WW: {                                  // :220-250
      int i = g.nbS + c, ib = s.ib, ibm1 = (ib + g.nbp - 1) % g.nbp;
      double xi = st.x[i], yi = st.y[i];
      if (!near_corner(g, xi, yi, ib)) return false;
      if (r) {
        double tbx = xi - g.bnode[ib][0], tby = yi - g.bnode[ib][1];
        r->irel = i; r->vrel = -(ib + 1);
        r->x = xi - 2.0 * tbx; r->y = yi - 2.0 * tby;
        double tmp = s.cross;   // the reference reuses the corner's cross product here
        double tx = g.bnorm[ib][1], ty = -g.bnorm[ib][0];
        double uu = 2.0 * tmp * tx - st.u[i], vv = 2.0 * tmp * ty - st.v[i];
        tx = g.bnorm[ibm1][1]; ty = -g.bnorm[ibm1][0];
        r->u = uu + 2.0 * tmp * tx; r->v = vv + 2.0 * tmp * ty;
        r->th = st.th[i] + 2.0 * g.bth[ib];
        r->dP = dPdot(g, r->x - xi, r->y - yi);
      }
      return true;
    }
    case SEG_CORNER_PP: {                                  // :253-280
      int i = g.nbS + c, ib = s.ib, ibp = s.ibp;
      if (!near_corner(g, st.x[i], st.y[i], ibp)) return false;
      if (r) {
        r->irel = i; r->vrel = -(ib + 1);
        double tx = g.bnode[ib][0] - g.bnode[ibp][0], ty = g.bnode[ib][1] - g.bnode[ibp][1];
        r->x = st.x[i] + tx; r->y = st.y[i] + ty;
        r->u = st.u[i]; r->v = st.v[i]; r->th = st.th[i]; r->dP = 0.0;
      }
      return true;
    }
    case SEG_CORNER_MIX: {                                 // :283-319
      if (c >= mlimit) return false;
      int j = g.npfb + c, ib = s.ib, ibb = s.ibp, ibm1 = (ib + g.nbp - 1) % g.nbp;
      double xj = st.x[j], yj = st.y[j];
      double pd = bound_dist(g, xj, yj, ibb);
      if (pd > g.sup + 1.0e-10) return false;
      if (pd < 1.0e-10) return false;
      double t = pos_along(g, xj, yj, ibb, pd);
      if (!((t < 0.0 && ibb == ib) || (t > 1.0 && ibb == ibm1))) return false;
      if (r) {
        r->irel = st.irel[j]; r->vrel = -(ib + 1);
        r->x = xj - 2.0 * pd * g.bnorm[ibb][0]; r->y = yj - 2.0 * pd * g.bnorm[ibb][1];
        double tx = g.bnorm[ibb][1], ty = -g.bnorm[ibb][0];
        r->u = 2.0 * g.bvel[ibb] * tx - st.u[j];
        r->v = 2.0 * g.bvel[ibb] * ty - st.v[j];
        r->th = 2.0 * g.bth[ibb] - st.th[j];
        double d = st.dP[j];
        r->dP = d + dPdot(g, r->x - xj, r->y - yj);
      }
      return true;
    }
  }
  return false;
}

__device__ inline int block_ballot_prefix(bool f, int* total) {
  __shared__ int wc[kBlock / 64];
  unsigned long long b = __ballot(f);
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) wc[wid] = __popcll(b);
  __syncthreads();
  int pre = 0, tot = 0;
  for (int w = 0; w < kBlock / 64; w++) { if (w < wid) pre += wc[w]; tot += wc[w]; }
  pre += __popcll(b & ((1ull << lane) - 1ull));
  __syncthreads();
  *total = tot;
  return pre;
}

__global__ void k_seg_count(Geom g, Seg s, State st, const Scal* sc, int* counts) {
  int c = blockIdx.x * kBlock + threadIdx.x;
  int mlimit = s.kind == SEG_CORNER_MIX ? sc->nmirror : 0;
  bool f = c < s.ncand && seg_eval(g, s, st, c, mlimit, nullptr);
  int tot;
  block_ballot_prefix(f, &tot);
  if (threadIdx.x == 0) counts[blockIdx.x] = tot;
}

// exclusive scan of n ints in place by one block of 1024 threads; returns the total to thread 0..
__device__ int block_scan_1024(int* a, int n) {
  __shared__ int sh[1024];
  int per = (n + 1023) / 1024;
  int b0 = threadIdx.x * per, b1 = min(n, b0 + per);
  int s = 0;
  for (int k = b0; k < b1; k++) s += a[k];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {          // Hillis-Steele inclusive scan of the run sums
    int v = threadIdx.x >= (unsigned)o ? sh[threadIdx.x - o] : 0;
    __syncthreads();
    sh[threadIdx.x] += v;
    __syncthreads();
  }
  int run = threadIdx.x ? sh[threadIdx.x - 1] : 0;
  int total = sh[1023];
  for (int k = b0; k < b1; k++) { int t = a[k]; a[k] = run; run += t; }
  __syncthreads();
  return total;
}

__global__ void k_seg_scan(int* counts, int nb, Scal* sc, int mcap) {
  int total = block_scan_1024(counts, nb);
  if (threadIdx.x == 0) {
    int base = sc->nmirror;
    sc->seg_base = base;
    if (base + total > mcap) { sc->err |= kErrCapacity; total = mcap - base; }
    sc->nmirror = base + total;
  }
}

// The linear part of mirror m's maps in its parent's values (see LinRec), for the segment that
// emitted it from candidate c: J of the position, d dP / d x_parent = (J - I)^T f where
// dP = f . (x_m - x_parent) (a mirror of a mirror adds to the intermediate mirror's), and the signs
// of the velocity and temperature maps.
__device__ void seg_record(const Geom& g, const Seg& s, int c, LinRec rec, int m) {
  double J[4] = {1.0, 0.0, 0.0, 1.0}, dp[2] = {0.0, 0.0};
  int sg = 0;
  const double fx = g.bf[0], fy = g.bf[1];
  switch (s.kind) {
    case SEG_CIRCLE: sg = 4; break;
    case SEG_WALL:
      jac_reflect(g, s.ib, J);
      dp[0] = (J[0] - 1.0) * fx + J[2] * fy; dp[1] = J[1] * fx + (J[3] - 1.0) * fy;
      sg = 3;
      break;
    case SEG_PERIODIC: jac_periodic(g, s.ibp, s.ib, J); break;
    case SEG_CORNER_WW:
      J[0] = -1.0; J[3] = -1.0;
      dp[0] = -2.0 * fx; dp[1] = -2.0 * fy;
      sg = 1;
      break;
    case SEG_CORNER_PP: break;
    case SEG_CORNER_MIX: {
      const double* Jj = rec.mjac + 6 * (size_t)c;       // the intermediate mirror npfb + c
      double R[4];
      jac_reflect(g, s.ibp, R);
      mat2_mul(R, Jj, J);
      dp[0] = Jj[4] + (J[0] - Jj[0]) * fx + (J[2] - Jj[2]) * fy;
      dp[1] = Jj[5] + (J[1] - Jj[1]) * fx + (J[3] - Jj[3]) * fy;
      sg = rec.msgn[c] ^ 3;
      break;
    }
  }
  double* o = rec.mjac + 6 * (size_t)m;
  o[0] = J[0]; o[1] = J[1]; o[2] = J[2]; o[3] = J[3]; o[4] = dp[0]; o[5] = dp[1];
  rec.msgn[m] = sg;
}

__global__ void k_seg_emit(Geom g, Seg s, State st, const Scal* sc, const int* offsets, int mcap, LinRec rec) {
  int c = blockIdx.x * kBlock + threadIdx.x;
  int base = sc->seg_base;
  int mlimit = s.kind == SEG_CORNER_MIX ? base : 0;
  MRec r;
  bool f = c < s.ncand && seg_eval(g, s, st, c, mlimit, &r);
  int tot;
  int pre = block_ballot_prefix(f, &tot);
  if (f) {
    int m = base + offsets[blockIdx.x] + pre;
    if (m < mcap) {
      int k = g.npfb + m;
      st.irel[k] = r.irel; st.vrel[k] = r.vrel; st.x[k] = r.x; st.y[k] = r.y;
      st.u[k] = r.u; st.v[k] = r.v; st.th[k] = r.th; st.dP[k] = r.dP;
      if (rec.mjac) seg_record(g, s, c, rec, m);
    }
  }
}

// mirrors of moving / heated solid particles keep the parent's value, :190-206
__global__ void k_mirror_fixup(Geom g, State st, const Scal* sc) {
  int m = blockIdx.x * kBlock + threadIdx.x;
  if (m >= sc->nmirror) return;
  int j = g.npfb + m, i = st.irel[j];
  if (i >= g.nbS) return;
  double uu = st.u[i] * st.u[i] + st.v[i] * st.v[i];
  if (uu != 0.0) { st.u[j] = st.u[i]; st.v[j] = st.v[i]; }
  double tt = st.th[i] * st.th[i];
  if (tt != 0.0) st.th[j] = st.th[i];
}

// ------------------------------------------------------------------------------------------------
// Phase 4: divide (neighbour_finding_mod.F90:56-127) as a counting sort, then
// neighboring_list_parallel (:261-365).
// ------------------------------------------------------------------------------------------------
__global__ void k_fill_int(int* a, int n, int v) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void k_cell_count(Geom g, State st, int np, int* cellkey, int* cellslot, int* nc, Scal* sc) {
  int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= np) return;
  if (st.inbin[k]) { cellkey[k] = -1; return; }
  double ddx = st.x[k] - g.xmint, ddy = st.y[k] - g.ymint;
  int icell = (int)(ddx * g.uno) + 1;
  int jcell = (int)(ddy * g.uno) + 1;
  int ii = icell + (jcell - 1) * g.ncx;
  if (ii > g.nct || ii < 1) { atomicOr(&sc->err, kErrOutOfGrid); cellkey[k] = -1; return; }
  int slot = atomicAdd(&nc[ii - 1], 1);
  if (slot + 1 > kNplink) atomicOr(&sc->err, kErrCellOverflow);
  cellkey[k] = ii;
  cellslot[k] = slot;
}

// Inclusive scan of v over a 1024-thread block (wavefront shuffles, then the 16 wave totals).
__device__ inline int block_inclusive_scan_1024(int v, int* total) {
  __shared__ int wsum[16];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  if (lane == 63) wsum[wid] = v;
  __syncthreads();
  int off = 0, tot = 0;
  for (int w = 0; w < 16; w++) {
    if (w < wid) off += wsum[w];
    tot += wsum[w];
  }
  __syncthreads();
  *total = tot;
  return v + off;
}

// Exclusive scan out[c] = sum of in[0..c-1], out[n] = total (the cell starts ist from the counts
// nc, and the cell-order ranks of the mixed solve): three passes over 1024-element blocks (block
// sums, one-block scan of the sums, block offsets added).
__global__ void k_cell_scan_blocks(const int* nc, int* ist, int* bsum, int nct) {
  int c = blockIdx.x * 1024 + threadIdx.x;
  int v = c < nct ? nc[c] : 0, tot;
  int inc = block_inclusive_scan_1024(v, &tot);
  if (c < nct) ist[c] = inc - v;                   // exclusive within the block
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ void k_cell_scan_top(int* bsum, int nblk, int* ist, int nct) {
  // one thread per block sum, in passes of 1024 block sums
  int carry = 0;
  for (int b0 = 0; b0 < nblk; b0 += 1024) {
    int b = b0 + threadIdx.x;
    int v = b < nblk ? bsum[b] : 0, tot;
    int inc = block_inclusive_scan_1024(v, &tot);
    if (b < nblk) bsum[b] = carry + inc - v;
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) ist[nct] = carry;
}

__global__ void k_cell_scan_add(int* ist, const int* bsum, int nct) {
  int c = blockIdx.x * 1024 + threadIdx.x;
  if (c < nct) ist[c] += bsum[blockIdx.x];
}

__global__ void k_cell_scatter(int np, const int* cellkey, const int* cellslot, const int* ist, int* ip) {
  int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= np) return;
  int key = cellkey[k];
  if (key < 1) return;
  ip[ist[key - 1] + cellslot[k]] = k;
}

// atomics give an arbitrary order inside a cell; the reference's scatter is stable (index order)
__global__ void k_cell_sort(const int* ist, int* ip, int nct) {
  int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= nct) return;
  int b = ist[c], e = ist[c + 1];
  for (int a = b + 1; a < e; a++) {
    int key = ip[a], q = a - 1;
    while (q >= b && ip[q] > key) { ip[q + 1] = ip[q]; q--; }
    ip[q + 1] = key;
  }
}

__global__ void k_neighbours(Geom g, State st, const int* cellkey, const int* ist, const int* ip,
                             int* cnt, int* link, Scal* sc) {
  const int i = xcd_row(g.npfb);
  if (i >= g.npfb) return;
  int n = 0;
  if (!st.inbin[i]) {
    const int ncx = g.ncx, ncy = g.ncy;
    int ic = cellkey[i];
    int icx = ic % ncx, icy = ic / ncx + 1;                // reference's row/column decode
    int cells[9], nce = 0;
    cells[nce++] = ic;
    if (icx + 1 <= ncx) {
      cells[nce++] = ic + 1;
      if (icy + 1 <= ncy) cells[nce++] = ic + ncx + 1;
      if (icy - 1 >= 1) cells[nce++] = ic - ncx + 1;
    }
    if (icx - 1 >= 1) {
      cells[nce++] = ic - 1;
      if (icy + 1 <= ncy) cells[nce++] = ic + ncx - 1;
      if (icy - 1 >= 1) cells[nce++] = ic - ncx - 1;
    }
    if (icy + 1 <= ncy) cells[nce++] = ic + ncx;
    if (icy - 1 >= 1) cells[nce++] = ic - ncx;
    const double xi = st.x[i], yi = st.y[i], sup2 = g.sup2;
    const int npfb = g.npfb;
    bool over = false;
    for (int q = 0; q < nce; q++) {
      int jc = cells[q] - 1;
      if (jc < 0 || jc >= g.nct) continue;
      int e = ist[jc + 1];
      for (int jj = ist[jc]; jj < e; jj++) {
        int j = ip[jj];
        double rx = xi - st.x[j], ry = yi - st.y[j];
        double rr2 = rx * rx + ry * ry;
        if (rr2 > 0.0 && rr2 <= sup2) {
          if (n < kNplink) link[n * npfb + i] = j; else over = true;
          n++;
        }
      }
    }
    if (over) { atomicOr(&sc->err, kErrCapacity); n = kNplink; }
  }
  cnt[i] = n;
}

// ------------------------------------------------------------------------------------------------
// Phase 5: kernel_calculation.F90:24-71 (weights, Morris Laplacian weight, KGC inverse).
// OPEN: the same pass also gathers what locate_free_surf_part needs from the pair (freesurf_mod.F90,
// calc_part_dist_vars :85-101, criterion :29-37, ppe_smooth :55-64), in slot order:
//   conc = dv (sum_k W(|r_ij|/h) + W(0)),  div r = -sum_k r_ij . w_G,  sn = -h sum_k w_G,
//   n_surf = (div r <= beta) with beta = 1.4 + 0.5 max(conc - 1, -0.2), and the PPE row scale:
//   1 on surface rows (their Dirichlet row is not smoothed) and where div r > 1.2 beta, else
//   0.5 (1 - cos((div r - beta) pi / (0.2 beta))).
// ------------------------------------------------------------------------------------------------
// (WT: the weight planes' type, float on the mixed path: the stores round to it)
template <bool OPEN, class WT>
__global__ void k_weights(Geom g, State st, const int* cnt, const int* link, WT* wg1, WT* wg2,
                          WT* wl, double* k11, double* k12, double* k21, double* k22, double* conc,
                          double* snx, double* sny, int* nsurf, double* pscale, Scal* sc) {
  const int i = xcd_row(g.npfb);
  int mx = 0;
  if (i < g.npfb) {
    const int npfb = g.npfb, n = cnt[i];
    mx = n;
    const double xi = st.x[i], yi = st.y[i], dv = g.dv;
    double a11 = 0.0, a12 = 0.0, a21 = 0.0, a22 = 0.0;
    double ct = 0.0, dr = 0.0, sx = 0.0, sy = 0.0;
    for_slots(link, i, n, npfb, [&](int j) { return make_double2(st.x[j], st.y[j]); },
              [&](int s, int, double2 pj) {
      double rx = xi - pj.x, ry = yi - pj.y;
      double rr2 = rx * rx + ry * ry;
      double rad = sqrt(rr2);
      double qq = rad * g.oneoverh;
      double f = fac(qq, g.ad7h);
      double den = rad > g.eta2 ? rad : g.eta2;
      double gx = rx * f / den, gy = ry * f / den;
      double w1 = gx * dv, w2 = gy * dv;
      double l = 2.0 * (rx * w1 + ry * w2) / (rr2 + g.eta_small);
      wg1[s] = w1; wg2[s] = w2; wl[s] = l;
      a11 = a11 - dv * gx * rx; a12 = a12 - dv * gy * rx;
      a21 = a21 - dv * gx * ry; a22 = a22 - dv * gy * ry;
      if (OPEN) {
        ct = ct + wab(rad / g.h, g.ad7);
        dr = dr - (rx * w1 + ry * w2);
        sx = sx - w1; sy = sy - w2;
      }
    });
    double t = 1.0 / ((a11 * a22 - a12 * a21) + 1.0e-15);
    k11[i] = t * a22; k12[i] = -t * a12; k21[i] = -t * a21; k22[i] = t * a11;
    if (OPEN) {
      const double c = ct * dv + dv * wab(0.0, g.ad7);
      const double beta = 1.4 + 0.5 * fmax(c - 1.0, -0.2);
      const int surf = dr > beta ? 0 : 1;
      double ps = 1.0;
      if (!(dr > 1.2 * beta) && !surf) ps = 0.5 * (1.0 - cos((dr - beta) * kPi / (0.2 * beta)));
      conc[i] = c; snx[i] = sx * g.h; sny[i] = sy * g.h; nsurf[i] = surf; pscale[i] = ps;
    }
  }
  // maxneighbours = maxval(ij_count)
  for (int o = 32; o > 0; o >>= 1) { int w = __shfl_down(mx, o, 64); mx = w > mx ? w : mx; }
  if ((threadIdx.x & 63) == 0 && mx > 0) atomicMax(&sc->maxneigh, mx);
}

// Phase 6: locate_free_surf_part (freesurf_mod.F90:39-45): closed domain n_surf = 0 everywhere; the
// open domain's rows 0..npfb-1 were flagged by k_weights<true>; mirrors take their parent's flag.
__global__ void k_freesurf(int* nsurf, const int* irel, int npfb, int np, int open) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= np) return;
  if (i >= npfb) nsurf[i] = open ? nsurf[irel[i]] : 0;
  else if (!open) nsurf[i] = 0;
}

// ------------------------------------------------------------------------------------------------
// Phase 7: shifting_forces + new_fick_shift (new_fick_shift.F90:15-92, 117-210, akinci off);
// ushift = dr_fs/dt and -u on walls (div_free.F90:65-78).
//   D: shift_coeff 0, 0.25 h^2 (:44); shift_coeff 1, 4 h dt (|u| + 0.2 exp(-|u| / (0.2 maxU)) maxU)
//      (:46-50) with the REAL(4) literal 0.2 widened, and maxU = max(1e-10, |u| of the last row),
//      the value of the reference's assignment inside its max reduction (:31-37) with one thread.
//   Open-domain surface rows (n_surf = 1) shift along the surface only: dr = -D (t . grad C) t, t the
//      rotated unit normal n = sn / max(|sn|, 1e-8 widened from REAL(4)) (:60-73), sn the surface
//      normal Shepard-filtered over the real neighbours, smooth_normals (freesurf_mod.F90:114-127):
//      dv (sum_{j real} W(|r_ij|/h) sn_j + W(0) sn_i), summed here in the same slot loop.
// ------------------------------------------------------------------------------------------------
template <class WT>
__global__ void k_shift(Geom g, State st, const int* cnt, const int* link, const WT* wg1,
                        const WT* wg2, const int* nsurf, const double* snx, const double* sny,
                        double* usx, double* usy, const Scal* sc) {
  const int i = xcd_row(g.npfb);
  if (i >= g.npfb) return;
  const int npfb = g.npfb, n = cnt[i];
  const double xi = st.x[i], yi = st.y[i], wab1 = sc->wab1, h = g.h;
  const bool surf = g.open && i >= g.nbS && nsurf[i] == 1;
  double sfx = 0.0, sfy = 0.0, stx = 0.0, sty = 0.0, nsx = 0.0, nsy = 0.0;
  for_slots(link, i, n, npfb, [&](int j) { return make_double2(st.x[j], st.y[j]); },
            [&](int s, int j, double2 pj) {
    double rx = xi - pj.x, ry = yi - pj.y;
    double rad = sqrt(rx * rx + ry * ry);
    double qq = rad / h;
    double w = wab(qq, g.ad7);
    // Rt (W/W1)^nt with the real exponent nt = 4 (:147): two squarings instead of pow (within an ulp
    // or two of the host's pow; pow's log/exp was the kernel's largest cost)
    const double wr = w / wab1, wr2 = wr * wr;
    double fab = 0.25 * (wr2 * wr2);
    double w1 = wg1[s], w2 = wg2[s];
    sfx = sfx + w1; sfy = sfy + w2;
    stx = stx + w1 * fab; sty = sty + w2 * fab;
    if (surf && j < npfb) { nsx = nsx + w * snx[j]; nsy = nsy + w * sny[j]; }
  });
  double dt = sc->dt;
  if (i < g.nbS) { usx[i] = -st.u[i]; usy[i] = -st.v[i]; return; }
  double D;
  if (g.shift_coeff == 1) {
    const double f02 = (double)0.2f;
    const double ul = st.u[npfb - 1], vl = st.v[npfb - 1];
    const double maxU = fmax(1.0e-10, sqrt(ul * ul + vl * vl));
    const double absU = sqrt(st.u[i] * st.u[i] + st.v[i] * st.v[i]);
    const double temp = f02 * exp(-absU / (f02 * maxU));
    D = 4.0 * h * dt * (absU + temp * maxU);
  } else {
    D = 0.25 * h * h;
  }
  const double gcx = sfx + stx, gcy = sfy + sty;
  double drx, dry;
  if (surf) {
    const double w0 = wab(0.0, g.ad7), dv = g.dv;
    const double sxs = nsx * dv + snx[i] * w0 * dv, sys = nsy * dv + sny[i] * w0 * dv;
    const double tn = sqrt(sxs * sxs + sys * sys);
    const double den = fmax(tn, (double)1.0e-8f);
    const double nx = sxs / den, ny = sys / den;
    const double tx = -ny, ty = nx;
    const double dCds = tx * gcx + ty * gcy;
    drx = -D * dCds * tx; dry = -D * dCds * ty;
  } else {
    drx = -D * gcx; dry = -D * gcy;
  }
  double t = sqrt(drx * drx + dry * dry);
  if (t > 0.1 * h) { drx = 0.1 * h * drx / t; dry = 0.1 * h * dry / t; }
  usx[i] = drx / dt; usy[i] = dry / dt;
}

// ------------------------------------------------------------------------------------------------
// Phase 8: prediction_step (predictor_mod.F90:27-139)
// ------------------------------------------------------------------------------------------------
// grad u, grad v (grad_operator, gradient_calculation.F90:61-71) fused with the per-particle
// log-conformation update (walls included, predictor_mod.F90:166)
template <class WT>
__global__ void k_grad_logconf(Geom g, State st, const int* cnt, const int* link, const WT* wg1,
                               const WT* wg2, const double* k11, const double* k12,
                               const double* k21, const double* k22, const double* kick, int use_kick,
                               double wi, double eps_ptt, const Scal* sc) {
  const int i = xcd_row(g.npfb);
  if (i >= g.npfb) return;
  const int npfb = g.npfb, n = cnt[i];
  const double ui = st.u[i], vi = st.v[i];
  double gux = 0.0, guy = 0.0, gvx = 0.0, gvy = 0.0;
  for_slots(link, i, n, npfb, [&](int j) { return make_double2(st.u[j], st.v[j]); },
            [&](int s, int, double2 uj) {
    double w1 = wg1[s], w2 = wg2[s];
    double tu = uj.x - ui, tv = uj.y - vi;
    gux = gux + tu * w1; guy = guy + tu * w2;
    gvx = gvx + tv * w1; gvy = gvy + tv * w2;
  });
  double K11 = k11[i], K12 = k12[i], K21 = k21[i], K22 = k22[i];
  double G[4] = {K11 * gux + K12 * guy, K21 * gux + K22 * guy, K11 * gvx + K12 * gvy, K21 * gvx + K22 * gvy};
  double tau[4] = {st.t11[i], st.t12[i], st.t21[i], st.t22[i]}, to[4], trc;
  logconf(tau, st.trc[i], G, sc->dt, wi, eps_ptt, use_kick ? kick[i] : 0.0, use_kick != 0, to, &trc);
  st.trc[i] = trc;
  st.t11[i] = to[0]; st.t12[i] = to[1]; st.t21[i] = to[2]; st.t22[i] = to[3];
}

// mirror_polymeric_stress, mirror_boundaries_mod.F90:586-653
__device__ inline void reflect_tau(const double* ti, double n1, double n2, double* tj) {
  double nt1 = -n2, nt2 = n1;
  double Q[4] = {1.0 * n1 + 0.0 * n2, 1.0 * nt1 + 0.0 * nt2, 0.0 * n1 + 1.0 * n2, 0.0 * nt1 + 1.0 * nt2};
  double QT[4] = {Q[0], Q[2], Q[1], Q[3]}, a[4], b[4];
  mm2(ti, Q, a);
  mm2(QT, a, b);
  b[0] = -b[0]; b[1] = -b[1];
  mm2(b, QT, a);
  mm2(Q, a, tj);
}

__global__ void k_mirror_stress(Geom g, State st, const Scal* sc) {
  int m = blockIdx.x * kBlock + threadIdx.x;
  if (m >= sc->nmirror) return;
  int j = g.npfb + m, i = st.irel[j], vr = st.vrel[j];
  double ti[4] = {st.t11[i], st.t12[i], st.t21[i], st.t22[i]}, tj[4];
  bool set = false;
  if (vr > 999) {
    double rx = st.x[i] - st.x[j], ry = st.y[i] - st.y[j];
    double d = sqrt(rx * rx + ry * ry);
    reflect_tau(ti, rx / d, ry / d, tj); set = true;
  }
  if (vr > 0 && vr < 999) {
    int t = g.btype[vr - 1];
    if (t == 1) {
      double rx = st.x[i] - st.x[j], ry = st.y[i] - st.y[j];
      double d = sqrt(rx * rx + ry * ry);
      reflect_tau(ti, rx / d, ry / d, tj); set = true;
    } else if (t == 2 || t == 4) {
      for (int q = 0; q < 4; q++) tj[q] = ti[q];
      set = true;
    } else if (t == 3) {
      for (int q = 0; q < 4; q++) tj[q] = 0.0;
      set = true;
    }
  }
  if (vr < 0 && vr > -g.npfb) {
    int ib = -vr - 1, ibm1 = (ib + g.nbp - 1) % g.nbp;
    int a = g.btype[ib], b = g.btype[ibm1];
    if (a == 1 && b == 1) { for (int q = 0; q < 4; q++) tj[q] = -ti[q]; set = true; }
    else if (a == 2 && b == 2) { for (int q = 0; q < 4; q++) tj[q] = ti[q]; set = true; }
    else if ((a == 1 && b == 2) || (a == 2 && b == 1)) {
      int w = a == 1 ? ib : ibm1;
      reflect_tau(ti, g.bnorm[w][0], g.bnorm[w][1], tj); set = true;
    }
  }
  if (set) { st.t11[j] = tj[0]; st.t12[j] = tj[1]; st.t21[j] = tj[2]; st.t22[j] = tj[3]; }
}

// RHS, buoyancy, u* and theta* (predictor_mod.F90:203-282), written to the alternate buffers
template <bool NEWTONIAN, class WT>
__global__ void k_rhs(Geom g, State st, const int* cnt, const int* link, const WT* wg1,
                      const WT* wg2, const WT* wl, const double* k11, const double* k12,
                      const double* k21, const double* k22, double cs, double cp, double ct,
                      double* un, double* vn, double* thn, const Scal* sc) {
  const int i = xcd_row(g.npfb);
  if (i >= g.npfb) return;
  const int npfb = g.npfb, n = cnt[i];
  const double ui = st.u[i], vi = st.v[i], thi = st.th[i];
  if (i < g.nbS) { un[i] = ui; vn[i] = vi; thn[i] = thi; return; }
  if (NEWTONIAN) {                               // predictor_mod.F90:58-77: solvent and thermal only
    double ax = 0.0, ay = 0.0, at = 0.0;
    for_slots(link, i, n, npfb, [&](int j) { return make_double4(st.u[j], st.v[j], st.th[j], 0.0); },
              [&](int s, int, double4 q) {
      const double l = wl[s];
      ax = ax + cs * (ui - q.x) * l;
      ay = ay + cs * (vi - q.y) * l;
      at = at + ct * (thi - q.z) * l;
    });
    ay = ay + thi;
    const double dt = sc->dt;
    un[i] = ui + ax * dt;
    vn[i] = vi + ay * dt;
    thn[i] = thi + at * dt;
    return;
  }
  const double T11i = st.t11[i], T12i = st.t12[i], T21i = st.t21[i], T22i = st.t22[i];
  const double K11 = k11[i], K12 = k12[i], K21 = k21[i], K22 = k22[i];
  double ax = 0.0, ay = 0.0, at = 0.0;
  struct Nb { double u, v, t11, t12, t21, t22, th; };
  for_slots(link, i, n, npfb,
            [&](int j) { return Nb{st.u[j], st.v[j], st.t11[j], st.t12[j], st.t21[j], st.t22[j], st.th[j]}; },
            [&](int s, int, const Nb& q) {
    double w1 = wg1[s], w2 = wg2[s], l = wl[s];
    double uij = ui - q.u, vij = vi - q.v;
    double T11 = q.t11 - T11i, T12 = q.t12 - T12i, T21 = q.t21 - T21i, T22 = q.t22 - T22i;
    double kw1 = K11 * w1 + K12 * w2, kw2 = K21 * w1 + K22 * w2;
    double tw1 = T11 * kw1 + T12 * kw2, tw2 = T21 * kw1 + T22 * kw2;
    ax = ax + cp * tw1 + cs * uij * l;
    ay = ay + cp * tw2 + cs * vij * l;
    at = at + ct * (thi - q.th) * l;
  });
  ay = ay + thi;
  double dt = sc->dt;
  un[i] = ui + ax * dt;
  vn[i] = vi + ay * dt;
  thn[i] = thi + at * dt;
}

// Phase 9: mirror_velocities, mirror_boundaries_mod.F90:509-584 (reads u*, writes the mirror rows
// of the alternate buffers; mirror temperatures are carried over unchanged)
__global__ void k_mirror_vel(Geom g, State st, double* un, double* vn, double* thn, const Scal* sc) {
  int m = blockIdx.x * kBlock + threadIdx.x;
  if (m >= sc->nmirror) return;
  int j = g.npfb + m, i = st.irel[j], vr = st.vrel[j];
  double ui = un[i], vi = vn[i];
  double uj = st.u[j], vj = st.v[j];
  if (vr > 999) {
    int ib = vr - 999 - 1;
    double rx = 0.5 * (st.x[i] - st.x[j]), ry = 0.5 * (st.y[i] - st.y[j]);
    double t = sqrt(rx * rx + ry * ry);
    double tx = ry / t, ty = -rx / t;
    uj = 2.0 * g.cvel[ib][0] + 2.0 * tx * g.crad[ib] * g.comg[ib] - 1.0 * ui;
    vj = 2.0 * g.cvel[ib][1] + 2.0 * ty * g.crad[ib] * g.comg[ib] - 1.0 * vi;
  }
  if (vr > 0 && vr < 999) {
    int t = g.btype[vr - 1];
    if (abs(t) == 1) {
      double rx = 0.5 * (st.x[i] - st.x[j]), ry = 0.5 * (st.y[i] - st.y[j]);
      double d = sqrt(rx * rx + ry * ry);
      double tx = ry / d, ty = -rx / d;
      uj = 2.0 * g.bvel[vr - 1] * tx - ui;
      vj = 2.0 * g.bvel[vr - 1] * ty - vi;
    } else if (t == 2) {
      uj = ui; vj = vi;
    }
  }
  if (vr < 0 && vr > -g.npfb) {
    int ib = -vr - 1, ibm1 = (ib + g.nbp - 1) % g.nbp;
    int a = g.btype[ib], b = g.btype[ibm1];
    if (a == 1 && b == 1) {
      uj = ui; vj = vi;
      if (g.bvel[ib] != 0.0) {
        double tmp = g.bvel[ib];
        uj = 2.0 * tmp * g.bnorm[ib][1] - ui; vj = 2.0 * tmp * (-g.bnorm[ib][0]) - vi;
      }
      if (g.bvel[ibm1] != 0.0) {
        double tmp = g.bvel[ibm1];
        uj = uj + 2.0 * tmp * g.bnorm[ibm1][1]; vj = vj + 2.0 * tmp * (-g.bnorm[ibm1][0]);
      }
    } else if (a == 2 && b == 2) {
      uj = ui; vj = vi;
    } else if ((a == 1 && b == 2) || (a == 2 && b == 1)) {
      double tmp, tx, ty;
      if (g.bvel[ib] != 0.0) { tx = g.bnorm[ib][1]; ty = -g.bnorm[ib][0]; tmp = g.bvel[ib]; }
      else if (g.bvel[ibm1] != 0.0) { tx = g.bnorm[ibm1][1]; ty = -g.bnorm[ibm1][0]; tmp = g.bvel[ibm1]; }
      else { tmp = 0.0; tx = 0.0; ty = 0.0; }
      uj = 2.0 * tmp * tx - ui;
      vj = 2.0 * tmp * ty - vi;
    }
  }
  if (i < g.nbS && (ui * ui + vi * vi) != 0.0) { uj = ui; vj = vi; }
  un[j] = uj; vn[j] = vj;
  thn[j] = st.th[j];
}

// ------------------------------------------------------------------------------------------------
// Phase 11: calc_pressure_gradient (gradient_calculation.F90:16-48) and the velocity update
// (div_free.F90:96-101); phase 12: final advection (:111-117), fused in one pass
// ------------------------------------------------------------------------------------------------
__global__ void k_mirror_p(Geom g, State st, const Scal* sc) {
  int m = blockIdx.x * kBlock + threadIdx.x;
  if (m >= sc->nmirror) return;
  int j = g.npfb + m;
  st.p[j] = st.p[st.irel[j]] + st.dP[j];
}

template <class WT>
__global__ void k_pressure_update(Geom g, State st, const int* cnt, const int* link, const WT* wg1,
                                  const WT* wg2, const double* k11, const double* k12,
                                  const double* k21, const double* k22, double* gpx, double* gpy,
                                  const Scal* sc) {
  const int i = xcd_row(g.npfb);
  if (i >= g.npfb) return;
  const int npfb = g.npfb, n = cnt[i];
  const double pi = st.p[i];
  double gx = 0.0, gy = 0.0;
  for_slots(link, i, n, npfb, [&](int j) { return st.p[j]; }, [&](int s, int, double pj) {
    double t = pj - pi;
    gx = gx + t * wg1[s]; gy = gy + t * wg2[s];
  });
  double px = k11[i] * gx + k12[i] * gy, py = k21[i] * gx + k22[i] * gy;
  gpx[i] = px; gpy[i] = py;
  if (i < g.nbS) return;
  double dt = sc->dt;
  double u = st.u[i] + dt * (g.bf[0] - px), v = st.v[i] + dt * (g.bf[1] - py);
  if (n <= 3) { u = st.u0[i] + dt * g.bf[0]; v = st.v0[i] + dt * g.bf[1]; }
  st.u[i] = u; st.v[i] = v;
}

__global__ void k_final(Geom g, State st, const double* usx, const double* usy, const Scal* sc) {
  int i = g.nbS + blockIdx.x * kBlock + threadIdx.x;
  if (i >= g.npfb) return;
  double dt = sc->dt;
  st.x[i] = st.x0[i] + 0.5 * (st.u[i] + st.u0[i]) * dt + usx[i] * dt;
  st.y[i] = st.y0[i] + 0.5 * (st.v[i] + st.v0[i]) * dt + usy[i] * dt;
}

// secondary_properties, div_free.F90:148-192: the vorticity of every row i <= npfb from the step's
// velocities, neighbour list and w_G (no KGC), vort_i = -sum_k [(v_i - v_j) w_G1 - (u_i - u_j) w_G2]
// in slot order (binned rows have no neighbours: 0), and the terms of sum_i u_i.u_i for the RMS
// velocity (:171-181).  The reference runs it between the velocity update and the final advection,
// which changes neither u nor the lists and weights, so it may run after the step.
template <class WT>
__global__ void k_vorticity(Geom g, State st, const int* cnt, const int* link, const WT* wg1,
                            const WT* wg2, double* vort, double* partials, double* red) {
  const int npfb = g.npfb;
  RowRange rr = block_rows(npfb, gridDim.x);
  Red<1> R(red, npfb);
  for (int i = rr.begin + threadIdx.x; i < rr.end; i += blockDim.x) {
    const double ui = st.u[i], vi = st.v[i];
    double w = 0.0;
    for_slots(link, i, cnt[i], npfb, [&](int j) { return make_double2(st.u[j], st.v[j]); },
              [&](int s, int, double2 uj) {
      const double t = (vi - uj.y) * wg1[s] - (ui - uj.x) * wg2[s];
      w = w - t;
    });
    vort[i] = w;
    R.add(0, i, ui * ui + vi * vi);                  // dot_product(u(i,:),u(i,:))
  }
  R.store(partials);
}

__global__ void k_fin_vorticity(const double* partials, int nblk, const double* red, int n, Scal* sc) {
  const double s = fold(partials, 0, nblk, red, n);
  if (threadIdx.x == 0) sc->u_rms = sqrt(s / (double)n);
}

// Run diagnostics the reference does not compute (defined here and identically in the oracle).
//
// Force of the fluid on circle ib: F = sum over its boundary particles b (wall rows whose distance
// from the centre is |R| to 1e-8 relative) of ds (sigma_b . n_b), with n_b = (r_b - c)/|r_b - c|,
// ds = 2 pi |R| / N_c, sigma = -p I + c_s (G + G^T) + c_p tau, and G the corrected velocity gradient
// K_b sum_k (u_j - u_b) w_G(b,k) (gradient_calculation.F90:62-71) over the step's neighbour list.
// One workgroup, one circle at a time: each thread sums its particles in index order, then the
// wavefronts and the workgroup fold in a fixed order.  out[3*ib + 0..2] = Fx, Fy, N_c.
constexpr int kDiagThreads = 1024;

__device__ inline bool on_circle(const Geom& g, const State& st, int i, int ib) {
  const double dx = st.x[i] - g.ccen[ib][0], dy = st.y[i] - g.ccen[ib][1];
  const double R = fabs(g.crad[ib]);
  return fabs(sqrt(dx * dx + dy * dy) - R) <= 1.0e-8 * R;
}

__device__ inline double block_sum_fixed(double v, double* sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  v = wave_sum(v);
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double s = sh[0];
  for (int w = 1; w < nw; w++) s += sh[w];
  __syncthreads();
  return s;
}

template <class WT>
__global__ void __launch_bounds__(kDiagThreads)
k_circle_forces(Geom g, State st, const int* cnt, const int* link, const WT* wg1, const WT* wg2,
                const double* k11, const double* k12, const double* k21, const double* k22, double cs,
                double cp, double* out) {
  __shared__ double sh[kDiagThreads / 64];
  const int npfb = g.npfb;
  for (int ib = 0; ib < g.ncirc; ib++) {
    double nloc = 0.0;
    for (int i = threadIdx.x; i < g.nbS; i += blockDim.x) nloc += on_circle(g, st, i, ib) ? 1.0 : 0.0;
    const double nc = block_sum_fixed(nloc, sh);
    const double ds = nc > 0.0 ? 2.0 * 3.14159265358979323846 * fabs(g.crad[ib]) / nc : 0.0;
    double fx = 0.0, fy = 0.0;
    for (int i = threadIdx.x; i < g.nbS; i += blockDim.x) {
      if (!on_circle(g, st, i, ib)) continue;
      const double ui = st.u[i], vi = st.v[i];
      double gux = 0.0, guy = 0.0, gvx = 0.0, gvy = 0.0;
      for (int k = 0; k < cnt[i]; k++) {
        const int s = k * npfb + i, j = link[s];
        const double tu = st.u[j] - ui, tv = st.v[j] - vi;
        gux = gux + tu * wg1[s]; guy = guy + tu * wg2[s];
        gvx = gvx + tv * wg1[s]; gvy = gvy + tv * wg2[s];
      }
      const double dudx = k11[i] * gux + k12[i] * guy, dudy = k21[i] * gux + k22[i] * guy;
      const double dvdx = k11[i] * gvx + k12[i] * gvy, dvdy = k21[i] * gvx + k22[i] * gvy;
      const double rx = st.x[i] - g.ccen[ib][0], ry = st.y[i] - g.ccen[ib][1];
      const double rn = sqrt(rx * rx + ry * ry), nx = rx / rn, ny = ry / rn;
      const double p = st.p[i];
      const double sxx = -p + cs * (2.0 * dudx) + cp * st.t11[i];
      const double sxy = cs * (dudy + dvdx) + cp * st.t12[i];
      const double syx = cs * (dvdx + dudy) + cp * st.t21[i];
      const double syy = -p + cs * (2.0 * dvdy) + cp * st.t22[i];
      fx = fx + ds * (sxx * nx + sxy * ny);
      fy = fy + ds * (syx * nx + syy * ny);
    }
    const double Fx = block_sum_fixed(fx, sh);
    const double Fy = block_sum_fixed(fy, sh);
    if (threadIdx.x == 0) { out[3 * ib] = Fx; out[3 * ib + 1] = Fy; out[3 * ib + 2] = nc; }
  }
}

// Nusselt number of the heated-from-below runs: the terms of sum v theta and of the particle count
// over the fluid rows not in the bin; Nu = 1 + scale * sum(v theta) / N with scale = sqrt(Ra Pr) H / dT.
__global__ void k_nusselt_terms(Geom g, State st, double* partials, double* red) {
  const int npfb = g.npfb;
  RowRange rr = block_rows(npfb, gridDim.x);
  Red<2> R(red, npfb);
  for (int i = rr.begin + threadIdx.x; i < rr.end; i += blockDim.x) {
    const bool fluid = i >= g.nbS && !st.inbin[i];
    R.add(0, i, fluid ? st.v[i] * st.th[i] : 0.0);
    R.add(1, i, fluid ? 1.0 : 0.0);
  }
  R.store(partials);
}

__global__ void k_fin_nusselt(const double* partials, int nblk, const double* red, int n, double scale,
                              double* out) {
  double f[2];
  fold_n<2>(partials, nblk, red, n, f);
  if (threadIdx.x == 0) out[0] = f[1] > 0.0 ? 1.0 + scale * (f[0] / f[1]) : 0.0;
}

// maxval(abs(u(1:npfb,1))), maxval(abs(u(1:npfb,2))), maxval(abs(P(1:npfb))) of output.F90:86
// (max is order-independent: per-block maxima, then one workgroup)
__global__ void k_field_max(State st, int npfb, double* partials) {
  double m[3] = {0.0, 0.0, 0.0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < npfb; i += gridDim.x * kBlock) {
    m[0] = fmax(m[0], fabs(st.u[i])); m[1] = fmax(m[1], fabs(st.v[i])); m[2] = fmax(m[2], fabs(st.p[i]));
  }
  __shared__ double sh[3][kBlock / 64];
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const double w = wave_max(m[q]);
    if ((threadIdx.x & 63) == 0) sh[q][threadIdx.x >> 6] = w;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double b = sh[threadIdx.x][0];
    for (int w = 1; w < kBlock / 64; w++) b = fmax(b, sh[threadIdx.x][w]);
    partials[threadIdx.x * gridDim.x + blockIdx.x] = b;
  }
}

__global__ void k_fin_field_max(const double* partials, int nblk, Scal* sc) {
  double m[3] = {0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nblk; b += 64)
    for (int q = 0; q < 3; q++) m[q] = fmax(m[q], partials[q * nblk + b]);
  for (int q = 0; q < 3; q++) {
    const double w = wave_max(m[q]);
    if (threadIdx.x == 0) sc->fmax[q] = w;
  }
}

// ================================================================================================
// launchers
// ================================================================================================
vb_status launch_set_time_step(vb_ctx* c) {
  int nb = grid_for(c->npfb - c->nbS);
  VB_LAUNCH(c, k_umax, nb, c->st, c->nbS, c->npfb, c->prm.dt_coef_cfl * c->prm.h, c->dt_visc, c->scal,
            c->partials);
  hipLaunchKernelGGL(k_fin_umax, dim3(1), dim3(64), 0, c->stream, c->partials, nb,
                     c->prm.dt_coef_cfl * c->prm.h, c->dt_visc, c->scal);
  VB_HIP(c, hipGetLastError());
  VB_LAUNCH(c, k_advect, grid_1d(c->npfb), c->st, c->nbS, c->npfb, c->scal);
  c->itime += 1;
  return VB_OK;
}

vb_status launch_put_in_bin(vb_ctx* c) {
  int nf = c->npfb - c->nbS;
  if (nf > 0) VB_LAUNCH(c, k_put_in_bin, grid_1d(nf), c->g, c->st, c->nbS, c->npfb, c->scal);
  VB_LAUNCH(c, k_check_all_binned, 1, nf, c->scal);
  return VB_OK;
}

static vb_status run_segment(vb_ctx* c, const Seg& s) {
  if (s.ncand <= 0) return VB_OK;
  int nb = grid_1d(s.ncand);
  if (nb > c->seg_blocks) return VB_ERR_CAPACITY;
  VB_LAUNCH(c, k_seg_count, nb, c->g, s, c->st, c->scal, c->seg_counts);
  hipLaunchKernelGGL(k_seg_scan, dim3(1), dim3(1024), 0, c->stream, c->seg_counts, nb, c->scal, c->mcap);
  VB_HIP(c, hipGetLastError());
  VB_LAUNCH(c, k_seg_emit, nb, c->g, s, c->st, c->scal, c->seg_counts, c->mcap, c->rec);
  return VB_OK;
}

static vb_status read_nmirror(vb_ctx* c) {
  VB_HIP(c, hipMemcpyAsync(c->scal_host, c->scal, sizeof(Scal), hipMemcpyDeviceToHost, c->stream));
  VB_HIP(c, hipStreamSynchronize(c->stream));
  c->nmirror_host = c->scal_host->nmirror;
  return VB_OK;
}

vb_status launch_mirrors(vb_ctx* c) {
  const Geom& g = c->g;
  int nf = c->npfb - c->nbS;
  if (nf > 0) VB_LAUNCH(c, k_escapees, grid_1d(nf), c->g, c->st, c->scal, c->rec);
  vb_status s;
  if (g.ncirc) {
    Seg sg{SEG_CIRCLE, 0, 0, nf * g.ncirc, 0.0};
    if ((s = run_segment(c, sg)) != VB_OK) return s;
  }
  for (int ib = 0; ib < g.nbp; ib++) {
    if (g.btype[ib] == 1) {
      Seg sg{SEG_WALL, ib, 0, c->npfb, 0.0};
      if ((s = run_segment(c, sg)) != VB_OK) return s;
    } else if (g.btype[ib] == 2) {
      Seg sg{SEG_PERIODIC, ib, g.bpp[ib], c->npfb, 0.0};
      if ((s = run_segment(c, sg)) != VB_OK) return s;
    }
  }
  // the fix-up reads the mirrors made so far; launch over the capacity bound, guarded on device
  VB_LAUNCH(c, k_mirror_fixup, grid_1d(c->mcap), c->g, c->st, c->scal);
  for (int ib = 0; ib < g.nbp; ib++) {
    int ibm1 = (ib + g.nbp - 1) % g.nbp;
    int a = g.btype[ib], b = g.btype[ibm1];
    if (a == 0 || b == 0) continue;
    double cross = g.bedge[ibm1][0] * g.bedge[ib][1] - g.bedge[ibm1][1] * g.bedge[ib][0];
    if (fabs(cross) < 1.0e-10) continue;
    if (a == 1 && b == 1 && g.bang[ib] > 0.0) {
      Seg sg{SEG_CORNER_WW, ib, 0, nf, cross};
      if ((s = run_segment(c, sg)) != VB_OK) return s;
    }
    if (a == 2 && b == 2) {
      Seg sg{SEG_CORNER_PP, ib, g.bpp[ib], nf, cross};
      if ((s = run_segment(c, sg)) != VB_OK) return s;
    }
    if (a != b && (a == 1 || b == 1)) {
      if ((s = read_nmirror(c)) != VB_OK) return s;
      Seg sg{SEG_CORNER_MIX, ib, b == 1 ? ibm1 : ib, c->nmirror_host, cross};
      if ((s = run_segment(c, sg)) != VB_OK) return s;
    }
  }
  if ((s = read_nmirror(c)) != VB_OK) return s;
  c->np_host = c->npfb + c->nmirror_host;
  if (c->scal_host->err & kErrCapacity) {
    c->err = "mirror capacity exceeded";
    return VB_ERR_CAPACITY;
  }
  return VB_OK;
}

vb_status launch_exclusive_scan(vb_ctx* c, const int* in, int* out, int n) {
  const int nsb = (n + 1023) / 1024;
  if (nsb < 1) return VB_OK;
  hipLaunchKernelGGL(k_cell_scan_blocks, dim3(nsb), dim3(1024), 0, c->stream, in, out, c->scan_bsum, n);
  VB_HIP(c, hipGetLastError());
  hipLaunchKernelGGL(k_cell_scan_top, dim3(1), dim3(1024), 0, c->stream, c->scan_bsum, nsb, out, n);
  VB_HIP(c, hipGetLastError());
  hipLaunchKernelGGL(k_cell_scan_add, dim3(nsb), dim3(1024), 0, c->stream, out, c->scan_bsum, n);
  VB_HIP(c, hipGetLastError());
  return VB_OK;
}

vb_status launch_neighbours(vb_ctx* c) {
  const Geom& g = c->g;
  c->op.valid = false; c->op.step_matrix = false;       // the stored PPE no longer matches the lists
  int np = c->np_host;
  VB_LAUNCH(c, k_fill_int, grid_1d(g.nct), c->nc, g.nct, 0);
  VB_LAUNCH(c, k_cell_count, grid_1d(np), c->g, c->st, np, c->cellkey, c->cellslot, c->nc, c->scal);
  vb_status s = launch_exclusive_scan(c, c->nc, c->ist, g.nct);
  if (s != VB_OK) return s;
  VB_LAUNCH(c, k_cell_scatter, grid_1d(np), np, c->cellkey, c->cellslot, c->ist, c->ip);
  VB_LAUNCH(c, k_cell_sort, grid_1d(g.nct), c->ist, c->ip, g.nct);
  VB_LAUNCH(c, k_neighbours, xcd_grid(c->npfb), c->g, c->st, c->cellkey, c->ist, c->ip, c->cnt, c->link,
            c->scal);
  return VB_OK;
}

vb_status launch_weights(vb_ctx* c) {
  return with_weights(c, [&](auto* w1, auto* w2, auto* wl) -> vb_status {
    using WT = weight_t<decltype(w1)>;
    if (c->g.open)
      VB_LAUNCH(c, (k_weights<true, WT>), xcd_grid(c->npfb), c->g, c->st, c->cnt, c->link, w1, w2, wl,
                c->k11, c->k12, c->k21, c->k22, c->conc, c->snx, c->sny, c->nsurf, c->pscale, c->scal);
    else
      VB_LAUNCH(c, (k_weights<false, WT>), xcd_grid(c->npfb), c->g, c->st, c->cnt, c->link, w1, w2, wl,
                c->k11, c->k12, c->k21, c->k22, nullptr, nullptr, nullptr, nullptr, nullptr, c->scal);
    return VB_OK;
  });
}

vb_status launch_freesurf(vb_ctx* c) {
  VB_LAUNCH(c, k_freesurf, grid_1d(c->np_host), c->nsurf, c->st.irel, c->npfb, c->np_host, c->g.open);
  return VB_OK;
}

vb_status launch_shift(vb_ctx* c) {
  return with_weights(c, [&](auto* w1, auto* w2, auto*) -> vb_status {
    VB_LAUNCH(c, k_shift<weight_t<decltype(w1)>>, xcd_grid(c->npfb), c->g, c->st, c->cnt, c->link, w1, w2,
              c->nsurf, c->snx, c->sny, c->usx, c->usy, c->scal);
    return VB_OK;
  });
}

vb_status launch_predict(vb_ctx* c) {
  const vb_params& p = c->prm;
  double wi = p.El * sqrt(p.Ra / p.Pr);
  int use_kick = p.reference_kick && c->itime != 2;
  const bool newtonian = p.fluid_model == 1;
  // the velocity gradient feeds only constitutive_eqn_evolve (and its mirror stresses): the
  // Newtonian build computes it and never reads it (predictor_mod.F90:35-46), so it is skipped
  double cs = (p.beta) * sqrt(p.Pr / p.Ra);
  double cp = (1.0 - p.beta) * p.Pr / (p.Ra * p.El);
  double ct = 1.0 / sqrt(p.Ra * p.Pr);
  return with_weights(c, [&](auto* w1, auto* w2, auto* wl) -> vb_status {
    using WT = weight_t<decltype(w1)>;
    if (!newtonian) {
      VB_LAUNCH(c, k_grad_logconf<WT>, xcd_grid(c->npfb), c->g, c->st, c->cnt, c->link, w1, w2, c->k11,
                c->k12, c->k21, c->k22, c->kick, use_kick, wi, p.ve_nonlinearity, c->scal);
      VB_LAUNCH(c, k_mirror_stress, grid_1d(c->nmirror_host), c->g, c->st, c->scal);
    }
    if (newtonian)
      VB_LAUNCH(c, (k_rhs<true, WT>), xcd_grid(c->npfb), c->g, c->st, c->cnt, c->link, w1, w2, wl, c->k11,
                c->k12, c->k21, c->k22, cs, cp, ct, c->u_alt, c->v_alt, c->th_alt, c->scal);
    else
      VB_LAUNCH(c, (k_rhs<false, WT>), xcd_grid(c->npfb), c->g, c->st, c->cnt, c->link, w1, w2, wl, c->k11,
                c->k12, c->k21, c->k22, cs, cp, ct, c->u_alt, c->v_alt, c->th_alt, c->scal);
    return VB_OK;
  });
}

vb_status launch_mirror_vel(vb_ctx* c) {
  if (c->nmirror_host > 0)
    VB_LAUNCH(c, k_mirror_vel, grid_1d(c->nmirror_host), c->g, c->st, c->u_alt, c->v_alt, c->th_alt,
              c->scal);
  double* t;
  t = c->st.u; c->st.u = c->u_alt; c->u_alt = t;
  t = c->st.v; c->st.v = c->v_alt; c->v_alt = t;
  t = c->st.th; c->st.th = c->th_alt; c->th_alt = t;
  return VB_OK;
}

vb_status launch_pressure_update(vb_ctx* c) {
  if (c->nmirror_host > 0) VB_LAUNCH(c, k_mirror_p, grid_1d(c->nmirror_host), c->g, c->st, c->scal);
  return with_weights(c, [&](auto* w1, auto* w2, auto*) -> vb_status {
    VB_LAUNCH(c, k_pressure_update<weight_t<decltype(w1)>>, xcd_grid(c->npfb), c->g, c->st, c->cnt, c->link,
              w1, w2, c->k11, c->k12, c->k21, c->k22, c->gpx, c->gpy, c->scal);
    return VB_OK;
  });
}

vb_status launch_mirror_p(vb_ctx* c) {
  if (c->nmirror_host > 0) VB_LAUNCH(c, k_mirror_p, grid_1d(c->nmirror_host), c->g, c->st, c->scal);
  return VB_OK;
}

vb_status launch_final(vb_ctx* c) {
  int nf = c->npfb - c->nbS;
  if (nf > 0) VB_LAUNCH(c, k_final, grid_1d(nf), c->g, c->st, c->usx, c->usy, c->scal);
  return VB_OK;
}

vb_status launch_secondary_properties(vb_ctx* c) {
  const int nb = spmv_grid(c->npfb);    // a neighbour sum: one row per thread
  vb_status s = with_weights(c, [&](auto* w1, auto* w2, auto*) -> vb_status {
    VB_LAUNCH_BS(c, k_vorticity<weight_t<decltype(w1)>>, nb, kSpmvBlock, c->g, c->st, c->cnt, c->link, w1, w2,
                 c->vort, c->partials, c->red);
    return VB_OK;
  });
  if (s != VB_OK) return s;
  hipLaunchKernelGGL(k_fin_vorticity, dim3(1), dim3(kBlock), 0, c->stream, c->partials, nb, c->red, c->npfb,
                     c->scal);
  VB_HIP(c, hipGetLastError());
  return VB_OK;
}

// device address of a double field of the context's device scalars
static double* scal_field(vb_ctx* c, size_t off) { return reinterpret_cast<double*>(reinterpret_cast<char*>(c->scal) + off); }

vb_status launch_circle_forces(vb_ctx* c) {
  const vb_params& p = c->prm;
  const double cs = (p.beta) * sqrt(p.Pr / p.Ra);            // the coefficients of launch_predict
  // the Newtonian build has no polymeric stress (its c_p may be 0/0 there)
  const double cp = p.fluid_model == 1 ? 0.0 : (1.0 - p.beta) * p.Pr / (p.Ra * p.El);
  return with_weights(c, [&](auto* w1, auto* w2, auto*) -> vb_status {
    hipLaunchKernelGGL(k_circle_forces<weight_t<decltype(w1)>>, dim3(1), dim3(kDiagThreads), 0, c->stream, c->g,
                       c->st, c->cnt, c->link, w1, w2, c->k11, c->k12, c->k21, c->k22, cs, cp,
                       scal_field(c, offsetof(Scal, circ)));
    VB_HIP(c, hipGetLastError());
    return VB_OK;
  });
}

vb_status launch_nusselt(vb_ctx* c, double scale) {
  const int nb = grid_for(c->npfb);
  VB_LAUNCH(c, k_nusselt_terms, nb, c->g, c->st, c->partials, c->red);
  hipLaunchKernelGGL(k_fin_nusselt, dim3(1), dim3(kBlock), 0, c->stream, c->partials, nb, c->red, c->npfb, scale,
                     scal_field(c, offsetof(Scal, nusselt)));
  VB_HIP(c, hipGetLastError());
  return VB_OK;
}

vb_status launch_field_maxima(vb_ctx* c) {
  const int nb = grid_for(c->npfb);
  VB_LAUNCH(c, k_field_max, nb, c->st, c->npfb, c->partials);
  hipLaunchKernelGGL(k_fin_field_max, dim3(1), dim3(64), 0, c->stream, c->partials, nb, c->scal);
  VB_HIP(c, hipGetLastError());
  return VB_OK;
}

}  // namespace vb
