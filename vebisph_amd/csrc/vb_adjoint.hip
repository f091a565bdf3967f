// The adjoint of the time step (the reference has none): reverse-mode derivatives of a run of
// div_free steps (div_free.F90:22-146) with respect to the start state and to Wi, beta and Ra, for a
// loss of the final state.
//
// Linearisation point.  Each step is differentiated with its neighbour lists, mirror set and branch
// choices frozen (as the forward pass took them); the PPE solve is differentiated as the exact solve
// of its assembled system, A p = b (the DC-projected BiCGSTAB leaves the constant in p free, and
// nothing downstream sees it: the corrected gradient of a constant is zero).  Everything else is the
// exact FP64 derivative of the kernels' own expressions: the kernel weights w_G, w_L and KGC
// matrices as functions of the positions (kernel_calculation.F90:40-71), the mirrors as affine maps
// of their parents (mirror_boundaries_mod.F90:106-324, their Jacobians recorded at emission), the
// Fick shift with its cap (new_fick_shift.F90:117-210), the predictor's right-hand side
// (predictor_mod.F90:79-133), the log-conformation map through forward-mode dual numbers
// (predictor_mod.F90:165-331), the PPE right-hand side, matrix and boundary terms
// (ppe_solve_in_house_simple_dirichlet.F90:27-122), the pressure gradient and velocity update
// (gradient_calculation.F90:16-48, div_free.F90:93-117), and dt = min(cfl h / max|u|, dt_visc(Ra))
// (div_free.F90:126-146).
//
// Scatter without atomics.  A neighbour sum over row i's slots touches its neighbours' values; the
// reverse of such a sum adds into the neighbours.  Each row kernel writes the neighbour-side terms to
// a slot-major buffer (same layout as the lists), then a gather kernel sums them per particle c:
// from its real neighbours i at the reverse slot (the slot of c in i's list; the relation is
// symmetric), and from every slot (i, k) that holds a mirror of c (a per-parent bucket, sorted by
// (i, k)), through that mirror's linear map.  Every sum has a fixed order, so a sweep gives the
// same bits run after run.
//
// Checkpoints.  The forward record keeps the state every `every` steps and the pressure of every
// step.  The reverse sweep recomputes each segment from its checkpoint with the stored pressures
// (no PPE solve: that is ~97% of a step at 1M particles, so storing 8 bytes per particle-step makes
// the recompute a few percent of a forward step), keeps the segment's states, and runs each step's
// adjoint from its own start state.  All device memory is allocated by the record, none in the sweep.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "vb_ctx.h"
#include "vb_logconf.h"

namespace vb {

// reduction quantities of one reverse step (partials at q * kPartialStride)
// (the right-hand side's slice also carries c_s, c_p, c_t; the log-conformation's Wi)
enum { AQ_DT_FINAL = 0, AQ_DT_PPE, AQ_DT_RHS, AQ_DT_LOGCONF, AQ_DT_ADVECT, AQ_LAM, AQ_N };

// device scalars of the sweep: accumulated parameter gradients
struct AdjScal {
  double g_wi, g_beta, g_ra;
  double dt_bar;                 // the last step's dL/d dt (diagnostic)
};

// Cotangent and work arrays (device; n = npfb rows unless noted).
struct Cot {
  double *x, *y, *u, *v, *th, *t11, *t12, *t21, *t22, *trc;   // state: end of step in, start of step out
  double *us, *vs;               // u* (after the predictor)
  double *u0, *v0;               // u0 (the step's start velocity, as the update and advection read it)
  double *x0, *y0;               // x0
  double *drx, *dry;             // the shift displacement (ushift dt)
  double *gpx, *gpy;             // K^T (dL / d grad p)
  double *k11, *k12, *k21, *k22; // dL / dK
  double *pb;                    // dL / dp (own-row part; the gather writes the total)
  double *db;                    // dL / d fb of the PPE right-hand side (fb = dt rhs)
  double *ax, *ay, *at;          // dt u*_bar, dt theta*_bar: the right-hand side's cotangent
  double *gux, *guy, *gvx, *gvy; // K^T (dL / d grad u), K^T (dL / d grad v)
  double *ua, *va, *tha;         // start u, v, theta (the predictor's reads)
  double *x1, *y1;               // positions at the kernel weights
  double *S1;                    // slot-major, 1 component
  double *S7;                    // slot-major, 7 components (component q at q * n * VB_NPLINK)
  size_t nw;                     // n * VB_NPLINK
};

// Forward values of the step the reverse step differentiates (after the re-run of its phases 1..10).
struct Fwd {
  const double *xs, *ys;         // x^1 (the weights' positions), mirrors included
  const double *us, *vs;         // u*, mirrors at phase 9
  const double *ua, *va, *tha;   // start u, v, theta with the phase-3 mirror values
  const double *t11, *t12, *t21, *t22;   // tau' (after the log-conformation), mirrors included
  const double *tau0[4], *trc0;  // start tau, trace
  const double *p, *dP;          // pressure (mirrors = parent + dP), mirror offsets
  const double *u0, *v0, *x0, *y0;
  const uint8_t* inbin;
  const int *irel, *vrel;
  const int *cnt, *link;
  const double *wg1, *wg2, *wl, *k11, *k12, *k21, *k22;
  const double* kick;
};

struct Coef { double cs, cp, ct, wi, pr, eps, dt_visc, cfl_h; int use_kick, newtonian; };

// velocity sign of the phase-9 mirror map u*_m = s u*_parent + c (mirror_velocities,
// mirror_boundaries_mod.F90:509-584)
__device__ inline double u9_sign(const Geom& g, int vr) {
  if (vr > 999) return -1.0;
  if (vr > 0 && vr < 999) {
    const int t = g.btype[vr - 1];
    if (abs(t) == 1) return -1.0;
    if (t == 2) return 1.0;
    return 0.0;
  }
  if (vr < 0 && vr > -g.npfb) {
    const int ib = -vr - 1, ibm1 = (ib + g.nbp - 1) % g.nbp;
    const int a = g.btype[ib], b = g.btype[ibm1];
    if (a == 1 && b == 1) return g.bvel[ib] != 0.0 ? -1.0 : 1.0;
    if (a == 2 && b == 2) return 1.0;
    if ((a == 1 && b == 2) || (a == 2 && b == 1)) return -1.0;
  }
  return 0.0;
}

// The stress mirror map tau_m = P tau_parent (mirror_polymeric_stress,
// mirror_boundaries_mod.F90:586-653); the reflection Q F Q^T is I - 2 n n^T.
__device__ inline void tau_map(const Geom& g, int vr, double* P) {
  P[0] = 0.0; P[1] = 0.0; P[2] = 0.0; P[3] = 0.0;
  auto refl = [&](double n1, double n2) {
    P[0] = 1.0 - 2.0 * n1 * n1; P[1] = -2.0 * n1 * n2; P[2] = -2.0 * n2 * n1; P[3] = 1.0 - 2.0 * n2 * n2;
  };
  if (vr > 0 && vr < 999) {
    const int t = g.btype[vr - 1];
    if (t == 1) refl(g.bnorm[vr - 1][0], g.bnorm[vr - 1][1]);
    else if (t == 2 || t == 4) { P[0] = 1.0; P[3] = 1.0; }
  } else if (vr < 0 && vr > -g.npfb) {
    const int ib = -vr - 1, ibm1 = (ib + g.nbp - 1) % g.nbp;
    const int a = g.btype[ib], b = g.btype[ibm1];
    if (a == 1 && b == 1) { P[0] = -1.0; P[3] = -1.0; }
    else if (a == 2 && b == 2) { P[0] = 1.0; P[3] = 1.0; }
    else if ((a == 1 && b == 2) || (a == 2 && b == 1)) {
      const int w = a == 1 ? ib : ibm1;
      refl(g.bnorm[w][0], g.bnorm[w][1]);
    }
  }
}

// d fac / dq (sphtools.F90:11-43 differentiated piecewise)
__device__ inline double dfac(double qq, double ad7h) {
  const int isel = (int)floor(qq);
  if (isel == 0) {
    const double q3 = 3.0 - qq, q2 = 2.0 - qq, q1 = 1.0 - qq;
    return ad7h * ((20.0 * q3 * q3 * q3 - 120.0 * q2 * q2 * q2) + 300.0 * q1 * q1 * q1);
  } else if (isel == 1) {
    const double q3 = 3.0 - qq, q2 = 2.0 - qq;
    return ad7h * (20.0 * q3 * q3 * q3 - 120.0 * q2 * q2 * q2);
  } else if (isel == 2) {
    const double q3 = 3.0 - qq;
    return ad7h * (20.0 * q3 * q3 * q3);
  }
  return 0.0;
}

// ------------------------------------------------------------------------------------------------
// Reverse of the velocity update, the pressure gradient and the final advection (div_free.F90:93-117,
// gradient_calculation.F90:16-48).  Per fluid row (walls and binned rows keep their values):
//   x_new = x0 + (u_new + u0) dt / 2 + dr,   u_new = u* + dt (f - K sum_k (p_j - p_i) w_G)   (n > 3)
//                                            u_new = u0 + dt f                                 (n <= 3)
// S1[s] = the p_j-side term of slot s; pb = the p_i-side terms.
// ------------------------------------------------------------------------------------------------
__global__ void k_adj_final(Geom g, Fwd F, Cot b, const Scal* sc, double* part) {
  const int n = g.npfb;
  RowRange rr = block_rows(n, gridDim.x);
  Red<1> R(nullptr, n);
  const double dt = sc->dt;
  for (int i = rr.begin + threadIdx.x; i < rr.end; i += blockDim.x) {
    const int m = F.cnt[i];
    double ubx = 0.0, uby = 0.0, u0x = 0.0, u0y = 0.0, x0x = 0.0, x0y = 0.0, gpx = 0.0, gpy = 0.0, pown = 0.0;
    double kb[4] = {0.0, 0.0, 0.0, 0.0};
    const bool fluid = i >= g.nbS && !F.inbin[i];
    if (fluid) {
      const double pi = F.p[i];
      double gx = 0.0, gy = 0.0;
      for_slots(F.link, i, m, n, [&](int j) { return F.p[j]; }, [&](int s, int, double pj) {
        const double t = pj - pi;
        gx = gx + t * F.wg1[s]; gy = gy + t * F.wg2[s];
      });
      const double K11 = F.k11[i], K12 = F.k12[i], K21 = F.k21[i], K22 = F.k22[i];
      const double px = K11 * gx + K12 * gy, py = K21 * gx + K22 * gy;
      const double Xb = b.x[i], Yb = b.y[i];
      const double u0 = F.u0[i], v0 = F.v0[i];
      double U, V;
      if (m > 3) { U = F.us[i] + dt * (g.bf[0] - px); V = F.vs[i] + dt * (g.bf[1] - py); }
      else { U = u0 + dt * g.bf[0]; V = v0 + dt * g.bf[1]; }
      x0x = Xb; x0y = Yb;
      const double Ut = b.u[i] + 0.5 * dt * Xb, Vt = b.v[i] + 0.5 * dt * Yb;
      u0x = 0.5 * dt * Xb; u0y = 0.5 * dt * Yb;
      double dtb = Xb * (0.5 * (U + u0)) + Yb * (0.5 * (V + v0));
      if (m > 3) {
        ubx = Ut; uby = Vt;
        dtb += Ut * (g.bf[0] - px) + Vt * (g.bf[1] - py);
        const double pbx = -dt * Ut, pby = -dt * Vt;            // dL / d grad p
        gpx = K11 * pbx + K21 * pby; gpy = K12 * pbx + K22 * pby;
        kb[0] = pbx * gx; kb[1] = pbx * gy; kb[2] = pby * gx; kb[3] = pby * gy;
      } else {
        u0x += Ut; u0y += Vt;
        dtb += Ut * g.bf[0] + Vt * g.bf[1];
      }
      R.add(0, i, dtb);
    }
    for (int k = 0; k < m; k++) {
      const size_t s = (size_t)k * n + i;
      const double t = gpx * F.wg1[s] + gpy * F.wg2[s];
      b.S1[s] = t;
      pown -= t;
    }
    b.us[i] = ubx; b.vs[i] = uby; b.u0[i] = u0x; b.v0[i] = u0y; b.x0[i] = x0x; b.y0[i] = x0y;
    b.drx[i] = x0x; b.dry[i] = x0y;
    b.gpx[i] = gpx; b.gpy[i] = gpy;
    b.k11[i] = kb[0]; b.k12[i] = kb[1]; b.k21[i] = kb[2]; b.k22[i] = kb[3];
    b.pb[i] = pown;
  }
  R.store(part + (size_t)AQ_DT_FINAL * kPartialStride);
}

// Gathers of the neighbour-side terms (see the file comment): real neighbours through the reverse
// slots, then the parent's bucket of mirror slots.
struct Nbr {
  const int *cnt, *link, *rslot, *boff, *bkey;
  int n;
};

// dL/dp of particle c (pressure mirrors are parent + dP: sign +1)
__global__ void k_adj_gather_p(Nbr N, const double* S1, double* pb) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= N.n) return;
  const int n = N.n;
  double acc = pb[c];
  for (int m = 0; m < N.cnt[c]; m++) {
    const size_t s = (size_t)m * n + c;
    const int r = N.rslot[s];
    if (r < 0) continue;
    acc = acc + S1[(size_t)r * n + N.link[s]];
  }
  for (int q = N.boff[c]; q < N.boff[c + 1]; q++) {
    const int key = N.bkey[q];
    acc = acc + S1[(size_t)(key % kNplink) * n + key / kNplink];
  }
  pb[c] = acc;
}

// ------------------------------------------------------------------------------------------------
// Reverse of the PPE right-hand side and its mean removal (ppe_solve_in_house_simple_dirichlet.F90:
// 27-53): rhs_i = fb_i / dt with fb_i = -sum_k (u*_i - u*_j) . K_i w_G, and b = rhs - mean(rhs) + the
// mirrors' dP w_L terms, so drhs = lambda - mean(lambda).  S7 components 0, 1: the u*_j-side terms.
// db = dL / d fb (the weights' pass uses it).
// ------------------------------------------------------------------------------------------------
__global__ void k_adj_ppe_rhs(Geom g, Fwd F, Cot b, const double* lam, const double* lam_part, int lam_nblk,
                              const Scal* sc, double* part) {
  const int n = g.npfb;
  const double lmean = g.open ? 0.0 : fold(lam_part, 0, lam_nblk, nullptr, n) / (double)n;
  RowRange rr = block_rows(n, gridDim.x);
  Red<1> R(nullptr, n);
  const double dt = sc->dt;
  for (int i = rr.begin + threadIdx.x; i < rr.end; i += blockDim.x) {
    const int m = F.cnt[i];
    const double ui = F.us[i], vi = F.vs[i];
    const double K11 = F.k11[i], K12 = F.k12[i], K21 = F.k21[i], K22 = F.k22[i];
    double fb = 0.0;
    for_slots(F.link, i, m, n, [&](int j) { return make_double2(F.us[j], F.vs[j]); },
              [&](int s, int, double2 uj) {
      const double w1 = F.wg1[s], w2 = F.wg2[s];
      const double uij = ui - uj.x, vij = vi - uj.y;
      const double kw1 = K11 * w1 + K12 * w2, kw2 = K21 * w1 + K22 * w2;
      fb = fb - (uij * kw1 + vij * kw2);
    });
    const double rb = lam[i] - lmean;                    // dL / d rhs_i
    const double fbb = rb / dt;                          // dL / d fb_i
    R.add(0, i, -rb * fb / (dt * dt));
    double dui = 0.0, dvi = 0.0, kb0 = 0.0, kb1 = 0.0, kb2 = 0.0, kb3 = 0.0;
    for (int k = 0; k < m; k++) {
      const size_t s = (size_t)k * n + i;
      const int j = F.link[s];
      const double w1 = F.wg1[s], w2 = F.wg2[s];
      const double uij = ui - F.us[j], vij = vi - F.vs[j];
      const double kw1 = K11 * w1 + K12 * w2, kw2 = K21 * w1 + K22 * w2;
      // fb -= uij kw1 + vij kw2
      const double gu = -fbb * kw1, gv = -fbb * kw2;     // dL / d uij, vij
      dui += gu; dvi += gv;
      b.S7[s] = -gu; b.S7[b.nw + s] = -gv;
      const double gk1 = -fbb * uij, gk2 = -fbb * vij;   // dL / d kw1, kw2
      kb0 += gk1 * w1; kb1 += gk1 * w2; kb2 += gk2 * w1; kb3 += gk2 * w2;
    }
    b.us[i] += dui; b.vs[i] += dvi;
    b.k11[i] += kb0; b.k12[i] += kb1; b.k21[i] += kb2; b.k22[i] += kb3;
    b.db[i] = fbb;
  }
  R.store(part + (size_t)AQ_DT_PPE * kPartialStride);
}

// dL/du* of particle c (phase-9 mirrors: u*_m = s u*_parent + c, s from the mirror's kind)
__global__ void k_adj_gather_us(Geom g, Nbr N, const int* vrel, const double* S7, size_t nw, double* us,
                                double* vs) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= N.n) return;
  const int n = N.n;
  double au = us[c], av = vs[c];
  for (int m = 0; m < N.cnt[c]; m++) {
    const size_t s = (size_t)m * n + c;
    const int r = N.rslot[s];
    if (r < 0) continue;
    const size_t t = (size_t)r * n + N.link[s];
    au = au + S7[t]; av = av + S7[nw + t];
  }
  for (int q = N.boff[c]; q < N.boff[c + 1]; q++) {
    const int key = N.bkey[q];
    const size_t t = (size_t)(key % kNplink) * n + key / kNplink;
    const double sg = u9_sign(g, vrel[N.link[t]]);
    au = au + sg * S7[t]; av = av + sg * S7[nw + t];
  }
  us[c] = au; vs[c] = av;
}

// ------------------------------------------------------------------------------------------------
// Reverse of the predictor's right-hand side and update (predictor_mod.F90:79-133), fluid rows:
//   u* = u + dt (sum_k [c_p (tau'_j - tau'_i) K_i w_G + c_s (u_i - u_j) w_L] + theta_i e_y),
//   theta* = theta + dt sum_k c_t (theta_i - theta_j) w_L.
// S7: components 0..3 the tau'_j-side terms, 4, 5 the u_j-side, 6 the theta_j-side.  Writes the
// start-state accumulators ua, va, tha (own-row terms), dL/dtau' of the own row into the state
// cotangent, and (ax, ay, at) = dt (u*_bar, theta*_bar) for the weights' pass.
// ------------------------------------------------------------------------------------------------
__global__ void k_adj_rhs(Geom g, Fwd F, Cot b, Coef cf, const Scal* sc, double* part) {
  const int n = g.npfb;
  RowRange rr = block_rows(n, gridDim.x);
  Red<4> R(nullptr, n);
  const double dt = sc->dt;
  const size_t nw = b.nw;
  for (int i = rr.begin + threadIdx.x; i < rr.end; i += blockDim.x) {
    const int m = F.cnt[i];
    if (i < g.nbS) {                                     // walls: u* = u, theta* = theta (constants)
      for (int k = 0; k < m; k++) {
        const size_t s = (size_t)k * n + i;
        for (int q = 0; q < 7; q++) b.S7[q * nw + s] = 0.0;
      }
      b.ua[i] = 0.0; b.va[i] = 0.0; b.tha[i] = 0.0; b.ax[i] = 0.0; b.ay[i] = 0.0; b.at[i] = 0.0;
      continue;
    }
    const double usb = b.us[i], vsb = b.vs[i], thsb = b.th[i];
    const double ax = dt * usb, ay = dt * vsb, at = dt * thsb;
    const double ui = F.ua[i], vi = F.va[i], thi = F.tha[i];
    const double T11i = F.t11[i], T12i = F.t12[i], T21i = F.t21[i], T22i = F.t22[i];
    const double K11 = F.k11[i], K12 = F.k12[i], K21 = F.k21[i], K22 = F.k22[i];
    const double cs = cf.cs, cp = cf.newtonian ? 0.0 : cf.cp, ct = cf.ct;
    double fx = 0.0, fy = 0.0, ft = 0.0;                // the right-hand side a_i (recomputed)
    double gcs = 0.0, gcp = 0.0, gct = 0.0;
    double dui = 0.0, dvi = 0.0, dthi = 0.0, dt11 = 0.0, dt12 = 0.0, dt21 = 0.0, dt22 = 0.0;
    double kb0 = 0.0, kb1 = 0.0, kb2 = 0.0, kb3 = 0.0;
    for (int k = 0; k < m; k++) {
      const size_t s = (size_t)k * n + i;
      const int j = F.link[s];
      const double w1 = F.wg1[s], w2 = F.wg2[s], l = F.wl[s];
      const double uij = ui - F.ua[j], vij = vi - F.va[j], thij = thi - F.tha[j];
      const double T11 = F.t11[j] - T11i, T12 = F.t12[j] - T12i, T21 = F.t21[j] - T21i, T22 = F.t22[j] - T22i;
      const double kw1 = K11 * w1 + K12 * w2, kw2 = K21 * w1 + K22 * w2;
      const double tw1 = T11 * kw1 + T12 * kw2, tw2 = T21 * kw1 + T22 * kw2;
      fx += cp * tw1 + cs * uij * l; fy += cp * tw2 + cs * vij * l; ft += ct * thij * l;
      gcp += ax * tw1 + ay * tw2;
      gcs += (ax * uij + ay * vij) * l;
      gct += at * thij * l;
      // tau'_j - tau'_i
      const double b11 = cp * ax * kw1, b12 = cp * ax * kw2, b21 = cp * ay * kw1, b22 = cp * ay * kw2;
      b.S7[s] = b11; b.S7[nw + s] = b12; b.S7[2 * nw + s] = b21; b.S7[3 * nw + s] = b22;
      dt11 -= b11; dt12 -= b12; dt21 -= b21; dt22 -= b22;
      // kw = K w_G
      const double gk1 = cp * (ax * T11 + ay * T21), gk2 = cp * (ax * T12 + ay * T22);
      kb0 += gk1 * w1; kb1 += gk1 * w2; kb2 += gk2 * w1; kb3 += gk2 * w2;
      // (u_i - u_j) w_L, (theta_i - theta_j) w_L
      const double eu = cs * ax * l, ev = cs * ay * l, et = ct * at * l;
      dui += eu; dvi += ev; dthi += et;
      b.S7[4 * nw + s] = -eu; b.S7[5 * nw + s] = -ev; b.S7[6 * nw + s] = -et;
    }
    fy += thi;                                           // buoyancy, predictor_mod.F90:115
    R.add(0, i, usb * fx + vsb * fy + thsb * ft);
    R.add(1, i, gcs); R.add(2, i, gcp); R.add(3, i, gct);
    b.ua[i] = usb + dui; b.va[i] = vsb + dvi; b.tha[i] = thsb + dthi + ay;
    b.t11[i] += dt11; b.t12[i] += dt12; b.t21[i] += dt21; b.t22[i] += dt22;
    b.k11[i] += kb0; b.k12[i] += kb1; b.k21[i] += kb2; b.k22[i] += kb3;
    b.ax[i] = ax; b.ay[i] = ay; b.at[i] = at;
  }
  R.store(part + (size_t)AQ_DT_RHS * kPartialStride);     // dt, c_s, c_p, c_t
}

// dL/dtau' and dL/dtheta of particle c from the right-hand side's neighbour-side terms (stress
// mirrors tau_m = P tau_parent: P^T; temperature mirrors: the recorded sign)
__global__ void k_adj_gather_rhs(Geom g, Nbr N, const int* vrel, const int* msgn, const double* S7, size_t nw,
                                 Cot b) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= N.n) return;
  const int n = N.n;
  double a0 = b.t11[c], a1 = b.t12[c], a2 = b.t21[c], a3 = b.t22[c], at = b.tha[c];
  for (int m = 0; m < N.cnt[c]; m++) {
    const size_t s = (size_t)m * n + c;
    const int r = N.rslot[s];
    if (r < 0) continue;
    const size_t t = (size_t)r * n + N.link[s];
    a0 = a0 + S7[t]; a1 = a1 + S7[nw + t]; a2 = a2 + S7[2 * nw + t]; a3 = a3 + S7[3 * nw + t];
    at = at + S7[6 * nw + t];
  }
  for (int q = N.boff[c]; q < N.boff[c + 1]; q++) {
    const int key = N.bkey[q];
    const size_t t = (size_t)(key % kNplink) * n + key / kNplink;
    const int j = N.link[t];
    double P[4];
    tau_map(g, vrel[j], P);
    const double m0 = S7[t], m1 = S7[nw + t], m2 = S7[2 * nw + t], m3 = S7[3 * nw + t];
    // tau_m = P tau_c  =>  tau_c_bar += P^T tau_m_bar
    a0 = a0 + (P[0] * m0 + P[2] * m2); a1 = a1 + (P[0] * m1 + P[2] * m3);
    a2 = a2 + (P[1] * m0 + P[3] * m2); a3 = a3 + (P[1] * m1 + P[3] * m3);
    const double sg = (msgn[j - n] & 2) ? -1.0 : 1.0;
    at = at + sg * S7[6 * nw + t];
  }
  b.t11[c] = a0; b.t12[c] = a1; b.t21[c] = a2; b.t22[c] = a3; b.tha[c] = at;
}

// ------------------------------------------------------------------------------------------------
// Reverse of the log-conformation update and of the velocity gradient it reads (predictor_mod.F90:
// 35-36, 141-331; walls included): J^T of the per-particle map by dual-number passes (one per
// input: tau, trace, grad u, Wi, dt), then grad u = K sum_k (u_j - u_i) w_G.  Writes dL/dtau and
// dL/dtrace of the start state in place, adds the u_j-side terms to S7 components 4, 5 and the own
// row's to ua, va, and keeps K^T (dL/d grad u), K^T (dL/d grad v) for the weights' pass.
// ------------------------------------------------------------------------------------------------
__global__ void k_adj_logconf(Geom g, Fwd F, Cot b, Coef cf, const Scal* sc, double* part) {
  const int n = g.npfb;
  RowRange rr = block_rows(n, gridDim.x);
  Red<2> R(nullptr, n);
  const double dt = sc->dt;
  const size_t nw = b.nw;
  for (int i = rr.begin + threadIdx.x; i < rr.end; i += blockDim.x) {
    const int m = F.cnt[i];
    const double ui = F.ua[i], vi = F.va[i];
    double gux = 0.0, guy = 0.0, gvx = 0.0, gvy = 0.0;
    for_slots(F.link, i, m, n, [&](int j) { return make_double2(F.ua[j], F.va[j]); },
              [&](int s, int, double2 uj) {
      const double w1 = F.wg1[s], w2 = F.wg2[s];
      const double tu = uj.x - ui, tv = uj.y - vi;
      gux = gux + tu * w1; guy = guy + tu * w2;
      gvx = gvx + tv * w1; gvy = gvy + tv * w2;
    });
    const double K11 = F.k11[i], K12 = F.k12[i], K21 = F.k21[i], K22 = F.k22[i];
    double Gb[4] = {0.0, 0.0, 0.0, 0.0}, wib = 0.0, dtb = 0.0;
    if (!cf.newtonian) {
      const double G[4] = {K11 * gux + K12 * guy, K21 * gux + K22 * guy, K11 * gvx + K12 * gvy,
                           K21 * gvx + K22 * gvy};
      const double x[11] = {F.tau0[0][i], F.tau0[1][i], F.tau0[2][i], F.tau0[3][i], F.trc0[i],
                            G[0], G[1], G[2], G[3], cf.wi, dt};
      const double ob[5] = {b.t11[i], b.t12[i], b.t21[i], b.t22[i], b.trc[i]};
      const double kick = cf.use_kick ? F.kick[i] : 0.0;
      double xb[11];
      for (int j = 0; j < 11; j++) {
        Dual in[11];
        for (int q = 0; q < 11; q++) in[q] = Dual(x[q], q == j ? 1.0 : 0.0);
        Dual o[4], tr;
        logconf(in, in[4], in + 5, in[10], in[9], cf.eps, kick, cf.use_kick != 0, o, &tr);
        xb[j] = (((ob[0] * o[0].d + ob[1] * o[1].d) + ob[2] * o[2].d) + ob[3] * o[3].d) + ob[4] * tr.d;
      }
      b.t11[i] = xb[0]; b.t12[i] = xb[1]; b.t21[i] = xb[2]; b.t22[i] = xb[3]; b.trc[i] = xb[4];
      Gb[0] = xb[5]; Gb[1] = xb[6]; Gb[2] = xb[7]; Gb[3] = xb[8];
      wib = xb[9]; dtb = xb[10];
    }
    // G = [K gu; K gv]
    const double gbux = K11 * Gb[0] + K21 * Gb[1], gbuy = K12 * Gb[0] + K22 * Gb[1];
    const double gbvx = K11 * Gb[2] + K21 * Gb[3], gbvy = K12 * Gb[2] + K22 * Gb[3];
    b.k11[i] += Gb[0] * gux + Gb[2] * gvx; b.k12[i] += Gb[0] * guy + Gb[2] * gvy;
    b.k21[i] += Gb[1] * gux + Gb[3] * gvx; b.k22[i] += Gb[1] * guy + Gb[3] * gvy;
    double dui = 0.0, dvi = 0.0;
    for (int k = 0; k < m; k++) {
      const size_t s = (size_t)k * n + i;
      const double w1 = F.wg1[s], w2 = F.wg2[s];
      const double eu = gbux * w1 + gbuy * w2, ev = gbvx * w1 + gbvy * w2;
      dui -= eu; dvi -= ev;
      b.S7[4 * nw + s] += eu; b.S7[5 * nw + s] += ev;       // the right-hand side wrote these slots
    }
    b.ua[i] += dui; b.va[i] += dvi;
    b.gux[i] = gbux; b.guy[i] = gbuy; b.gvx[i] = gbvx; b.gvy[i] = gbvy;
    R.add(0, i, dtb); R.add(1, i, wib);
  }
  R.store(part + (size_t)AQ_DT_LOGCONF * kPartialStride);
}

// dL/du of particle c (the start velocities the predictor reads; phase-3 mirrors: recorded sign)
__global__ void k_adj_gather_u(Nbr N, const int* msgn, const double* S7, size_t nw, double* ua, double* va) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= N.n) return;
  const int n = N.n;
  double au = ua[c], av = va[c];
  for (int m = 0; m < N.cnt[c]; m++) {
    const size_t s = (size_t)m * n + c;
    const int r = N.rslot[s];
    if (r < 0) continue;
    const size_t t = (size_t)r * n + N.link[s];
    au = au + S7[4 * nw + t]; av = av + S7[5 * nw + t];
  }
  for (int q = N.boff[c]; q < N.boff[c + 1]; q++) {
    const int key = N.bkey[q];
    const size_t t = (size_t)(key % kNplink) * n + key / kNplink;
    const double sg = (msgn[N.link[t] - n] & 1) ? -1.0 : 1.0;
    au = au + sg * S7[4 * nw + t]; av = av + sg * S7[5 * nw + t];
  }
  ua[c] = au; va[c] = av;
}

// ------------------------------------------------------------------------------------------------
// Reverse of the geometry: every use of w_G, w_L and K_i in the step, differentiated through the
// weights' formulas (kernel_calculation.F90:40-66) to the pair vectors r_ij = x_i - x_j:
//   w_G = dv fac(|r|/h) r / max(|r|, eta2),  w_L = 2 (r . w_G) / (|r|^2 + 0.01f h^2),
//   K = inverse(-dv sum_k g (x) r) with +1e-15 on the determinant;
// and the Fick shift (new_fick_shift.F90:130-154, fluid rows): dr = -D sum_k w_G (1 + 0.25 (W(q) /
// W(dx/h))^4), capped at 0.1 h.  The w_G, w_L cotangents of a slot collect the pressure gradient,
// the PPE right-hand side, matrix (A_ii = sum w_L, A_i,col(j) -= w_L, a self-mirror cancelling) and
// boundary terms (b_i += dP_j w_L), the predictor's right-hand side, the velocity gradient and the
// shift.  Own-row terms go to x1, the neighbour-side ones to S7 components 0, 1.
// ------------------------------------------------------------------------------------------------
__global__ void k_adj_geometry(Geom g, Fwd F, Cot b, Coef cf, const double* lam, const Scal* sc) {
  const int n = g.npfb;
  RowRange rr = block_rows(n, gridDim.x);
  const size_t nw = b.nw;
  const double dv = g.dv, wab1 = sc->wab1, h = g.h;
  for (int i = rr.begin + threadIdx.x; i < rr.end; i += blockDim.x) {
    const int m = F.cnt[i];
    if (m == 0) { b.x1[i] = 0.0; b.y1[i] = 0.0; continue; }
    const double xi = F.xs[i], yi = F.ys[i];
    const bool fluid = i >= g.nbS && !F.inbin[i];
    // pass 1: the moment sums and the shift's concentration gradient
    double a11 = 0.0, a12 = 0.0, a21 = 0.0, a22 = 0.0, sfx = 0.0, sfy = 0.0, stx = 0.0, sty = 0.0;
    for (int k = 0; k < m; k++) {
      const size_t s = (size_t)k * n + i;
      const int j = F.link[s];
      const double rx = xi - F.xs[j], ry = yi - F.ys[j];
      const double rr2 = rx * rx + ry * ry, rad = sqrt(rr2);
      const double f = fac(rad * g.oneoverh, g.ad7h);
      const double den = rad > g.eta2 ? rad : g.eta2;
      const double gx = rx * f / den, gy = ry * f / den;
      a11 = a11 - dv * gx * rx; a12 = a12 - dv * gy * rx;
      a21 = a21 - dv * gx * ry; a22 = a22 - dv * gy * ry;
      if (fluid) {
        const double w = wab(rad / h, g.ad7);
        const double wr = w / wab1, wr2 = wr * wr;
        const double fab = 0.25 * (wr2 * wr2);                       // as k_shift
        const double w1 = F.wg1[s], w2 = F.wg2[s];
        sfx = sfx + w1; sfy = sfy + w2; stx = stx + w1 * fab; sty = sty + w2 * fab;
      }
    }
    // K = t adj(M), t = 1 / (det M + 1e-15)
    const double det = a11 * a22 - a12 * a21, tt = 1.0 / (det + 1.0e-15);
    const double Kb11 = b.k11[i], Kb12 = b.k12[i], Kb21 = b.k21[i], Kb22 = b.k22[i];
    const double tb = Kb11 * a22 - Kb12 * a12 - Kb21 * a21 + Kb22 * a11;
    const double detb = -tt * tt * tb;
    const double ab11 = tt * Kb22 + detb * a22, ab22 = tt * Kb11 + detb * a11;
    const double ab12 = -tt * Kb12 - detb * a21, ab21 = -tt * Kb21 - detb * a12;
    // the shift: dr = -D gc, capped at 0.1 h  ->  dL/d gc
    double gcbx = 0.0, gcby = 0.0;
    if (fluid) {
      const double D = 0.25 * h * h;
      const double gcx = sfx + stx, gcy = sfy + sty;
      const double drx = -D * gcx, dry = -D * gcy;
      const double t = sqrt(drx * drx + dry * dry);
      double dbx = b.drx[i], dby = b.dry[i];
      if (t > 0.1 * h) {
        const double dot = drx * dbx + dry * dby;
        const double c1 = 0.1 * h / t, c2 = 0.1 * h * dot / (t * t * t);
        dbx = c1 * dbx - c2 * drx; dby = c1 * dby - c2 * dry;
      }
      gcbx = -D * dbx; gcby = -D * dby;
    }
    // row cotangents of the step's uses of this row's slots
    const double pi = F.p[i], lami = lam[i], dbi = b.db[i];
    const double gpx = b.gpx[i], gpy = b.gpy[i];
    const double K11 = F.k11[i], K12 = F.k12[i], K21 = F.k21[i], K22 = F.k22[i];
    const double usi = F.us[i], vsi = F.vs[i], ui = F.ua[i], vi = F.va[i], thi = F.tha[i];
    const double T11i = F.t11[i], T12i = F.t12[i], T21i = F.t21[i], T22i = F.t22[i];
    const double ax = b.ax[i], ay = b.ay[i], at = b.at[i];
    const double gux = b.gux[i], guy = b.guy[i], gvx = b.gvx[i], gvy = b.gvy[i];
    const double cp = cf.newtonian ? 0.0 : cf.cp;
    double xb = 0.0, yb = 0.0;
    for (int k = 0; k < m; k++) {
      const size_t s = (size_t)k * n + i;
      const int j = F.link[s];
      const bool mir = j >= n;
      const double rx = xi - F.xs[j], ry = yi - F.ys[j];
      const double rr2 = rx * rx + ry * ry, rad = sqrt(rr2);
      const double qq = rad * g.oneoverh;
      const double f = fac(qq, g.ad7h);
      const double den = rad > g.eta2 ? rad : g.eta2;
      const double phi = f / den;
      const double gx = rx * phi, gy = ry * phi;
      const double w1 = gx * dv, w2 = gy * dv;
      const double dn = rr2 + g.eta_small;
      const double l = 2.0 * (rx * w1 + ry * w2) / dn;
      // dL / d w_G
      const double pj = F.p[j];
      double wb1 = (pj - pi) * gpx, wb2 = (pj - pi) * gpy;                 // grad p
      wb1 += (F.ua[j] - ui) * gux + (F.va[j] - vi) * gvx;                  // grad u, grad v
      wb2 += (F.ua[j] - ui) * guy + (F.va[j] - vi) * gvy;
      // kw = K w_G in the PPE right-hand side and the polymeric term
      const double dus = usi - F.us[j], dvs = vsi - F.vs[j];
      const double T11 = F.t11[j] - T11i, T12 = F.t12[j] - T12i, T21 = F.t21[j] - T21i, T22 = F.t22[j] - T22i;
      const double kb1 = -dbi * dus + cp * (ax * T11 + ay * T21);
      const double kb2 = -dbi * dvs + cp * (ax * T12 + ay * T22);
      wb1 += K11 * kb1 + K21 * kb2; wb2 += K12 * kb1 + K22 * kb2;
      // dL / d w_L: the PPE matrix, boundary term, viscous and thermal terms
      const double pcol = mir ? F.p[F.irel[j]] : pj;
      double lb = lami * (pcol - pi);
      if (mir) lb += lami * F.dP[j];
      lb += cf.cs * (ax * (ui - F.ua[j]) + ay * (vi - F.va[j])) + cf.ct * at * (thi - F.tha[j]);
      double rbx = 0.0, rby = 0.0, radb = 0.0;
      if (fluid) {                                                          // the shift
        const double qs = rad / h;
        const double w = wab(qs, g.ad7);
        const double wr = w / wab1, wr2 = wr * wr;
        const double fab = 0.25 * (wr2 * wr2);                      // as k_shift
        wb1 += gcbx * (1.0 + fab); wb2 += gcby * (1.0 + fab);
        const double fabb = gcbx * w1 + gcby * w2;
        radb += fabb * (wr * wr * wr) * fac(qs, g.ad7) / (wab1 * h);      // d fab / d rad
      }
      // w_L = 2 N / dn, N = r . w_G
      const double Nb = 2.0 * lb / dn, dnb = -l * lb / dn;
      wb1 += Nb * rx; wb2 += Nb * ry;
      rbx += Nb * w1; rby += Nb * w2;
      double rr2b = dnb;
      // moment sums a_ab = -dv r_a g_b and w_G = dv g
      double gbx = dv * wb1 - dv * (ab11 * rx + ab21 * ry);
      double gby = dv * wb2 - dv * (ab12 * rx + ab22 * ry);
      rbx += -dv * (ab11 * gx + ab12 * gy);
      rby += -dv * (ab21 * gx + ab22 * gy);
      // g = r phi, phi = fac(q) / den
      rbx += phi * gbx; rby += phi * gby;
      const double phib = rx * gbx + ry * gby;
      const double fb = phib / den;
      if (rad > g.eta2) radb += -phib * f / (den * den);
      radb += fb * dfac(qq, g.ad7h) * g.oneoverh;
      rr2b += radb / (2.0 * rad);
      rbx += 2.0 * rx * rr2b; rby += 2.0 * ry * rr2b;
      xb += rbx; yb += rby;
      b.S7[s] = -rbx; b.S7[nw + s] = -rby;
    }
    b.x1[i] = xb; b.y1[i] = yb;
  }
}

// dL/dx^1 of particle c: its own row, its real neighbours' rows, and its mirrors (x_m = J x_c + c:
// J^T; their offsets dP_m, read by the pressure mirrors and the boundary terms: d dP_m / d x_c)
__global__ void k_adj_gather_x(Nbr N, LinRec rec, const double* S1, const double* S7, size_t nw,
                               const double* lam, const double* wl, double* x1, double* y1) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= N.n) return;
  const int n = N.n;
  double ax = x1[c], ay = y1[c];
  for (int m = 0; m < N.cnt[c]; m++) {
    const size_t s = (size_t)m * n + c;
    const int r = N.rslot[s];
    if (r < 0) continue;
    const size_t t = (size_t)r * n + N.link[s];
    ax = ax + S7[t]; ay = ay + S7[nw + t];
  }
  for (int q = N.boff[c]; q < N.boff[c + 1]; q++) {
    const int key = N.bkey[q];
    const int i = key / kNplink;
    const size_t t = (size_t)(key % kNplink) * n + i;
    const double* J = rec.mjac + 6 * (size_t)(N.link[t] - n);
    const double bx = S7[t], by = S7[nw + t];
    const double dpb = S1[t] + lam[i] * wl[t];
    ax = ax + ((J[0] * bx + J[2] * by) + J[4] * dpb);
    ay = ay + ((J[1] * bx + J[3] * by) + J[5] * dpb);
  }
  x1[c] = ax; y1[c] = ay;
}

// ------------------------------------------------------------------------------------------------
// Reverse of phase 1-3 for the particle itself: x^1 = x + u dt (fluid, div_free.F90:47-53), then
// replace_escapees (x^1 -> J_e x^1 + c, x0 = x^1 - u dt, mirror_boundaries_mod.F90:336-407), and
// put_in_bin (binned rows take constant positions and zero velocity: only x0 = x and u0 = u still
// read the start state).  Walls keep constant positions, velocities and temperatures.  Writes the
// start state's x, u, theta cotangents.
// ------------------------------------------------------------------------------------------------
__global__ void k_adj_advect(Geom g, Fwd F, Cot b, LinRec rec, const Scal* sc, double* part) {
  const int n = g.npfb;
  RowRange rr = block_rows(n, gridDim.x);
  Red<1> R(nullptr, n);
  const double dt = sc->dt;
  for (int i = rr.begin + threadIdx.x; i < rr.end; i += blockDim.x) {
    double xs = 0.0, ys = 0.0, us = 0.0, vs = 0.0, ths = 0.0, dtb = 0.0;
    if (i >= g.nbS) {
      ths = b.tha[i];
      if (F.inbin[i]) {
        xs = b.x0[i]; ys = b.y0[i]; us = b.u0[i]; vs = b.v0[i];
      } else {
        const double u = F.u0[i], v = F.v0[i];
        double x1b = b.x1[i], y1b = b.y1[i];
        us = b.ua[i] + b.u0[i]; vs = b.va[i] + b.v0[i];
        if (rec.eflag[i]) {                 // x0 = x^1' - u dt, x^1' = J_e x^1 + c
          const double x0b = b.x0[i], y0b = b.y0[i];
          x1b += x0b; y1b += y0b;
          us -= dt * x0b; vs -= dt * y0b;
          dtb -= u * x0b + v * y0b;
          const double* J = rec.ejac + 4 * (size_t)i;
          const double tx = J[0] * x1b + J[2] * y1b, ty = J[1] * x1b + J[3] * y1b;
          x1b = tx; y1b = ty;
        } else {
          xs = b.x0[i]; ys = b.y0[i];
        }
        xs += x1b; ys += y1b;
        us += dt * x1b; vs += dt * y1b;
        dtb += u * x1b + v * y1b;
      }
    }
    b.x[i] = xs; b.y[i] = ys; b.u[i] = us; b.v[i] = vs; b.th[i] = ths;
    R.add(0, i, dtb);
  }
  R.store(part + (size_t)AQ_DT_ADVECT * kPartialStride);
}

// One workgroup: fold the step's dt, c_s, c_p, c_t and Wi cotangents; dt = min(cfl h / umax,
// dt_visc) sends dL/d dt to the fastest fluid particle's velocity (CFL branch) or to Ra (viscous
// branch, dt_visc ~ sqrt(Ra)); accumulate dL/dWi, dL/dbeta, dL/dRa with El = Wi sqrt(Pr / Ra):
// c_s = beta sqrt(Pr/Ra), c_p = (1 - beta) sqrt(Pr/Ra) / Wi, c_t = 1 / sqrt(Ra Pr).
// Partials: quantity 0 of every slice is a dt term; the right-hand side's slice also holds c_s, c_p,
// c_t (quantities 1..3) and the log-conformation's Wi (quantity 1), at q * nblk.
__global__ void k_adj_scalars(Geom g, Coef cf, double ra, double beta, const double* part, int nblk,
                              const double* u, const double* v, const Scal* sc, Cot b, AdjScal* as) {
  double f1[1], r4[4], r2[2];
  double dtb = 0.0;
  fold_n<1>(part + (size_t)AQ_DT_FINAL * kPartialStride, nblk, nullptr, 0, f1); dtb += f1[0];
  fold_n<1>(part + (size_t)AQ_DT_PPE * kPartialStride, nblk, nullptr, 0, f1); dtb += f1[0];
  fold_n<4>(part + (size_t)AQ_DT_RHS * kPartialStride, nblk, nullptr, 0, r4); dtb += r4[0];
  fold_n<2>(part + (size_t)AQ_DT_LOGCONF * kPartialStride, nblk, nullptr, 0, r2); dtb += r2[0];
  fold_n<1>(part + (size_t)AQ_DT_ADVECT * kPartialStride, nblk, nullptr, 0, f1); dtb += f1[0];
  const double csb = r4[1], cpb = r4[2], ctb = r4[3], wib = r2[1];
  // CFL branch: the fastest fluid particle of the step's start velocities (lowest index among equals)
  const bool cfl = sc->dt != cf.dt_visc;
  __shared__ double smax[1024];
  __shared__ int sidx[1024];
  int am = -1;
  double um = 1.0e-10;
  if (cfl)
    for (int i = g.nbS + threadIdx.x; i < g.npfb; i += blockDim.x) {
      const double a = sqrt(u[i] * u[i] + v[i] * v[i]);
      if (a > um) { um = a; am = i; }
    }
  smax[threadIdx.x] = um; sidx[threadIdx.x] = am;
  __syncthreads();
  if (threadIdx.x != 0) return;
  double best = 1.0e-10;
  int bi = -1;
  for (int t = 0; t < (int)blockDim.x; t++) {
    if (sidx[t] < 0) continue;
    if (smax[t] > best || (smax[t] == best && sidx[t] < bi)) { best = smax[t]; bi = sidx[t]; }
  }
  double g_ra = 0.0;
  if (cfl) {
    if (bi >= 0) {                                     // dt = cfl h / |u_bi|
      const double dd = -cf.cfl_h / (best * best) * dtb;
      b.u[bi] += dd * u[bi] / best; b.v[bi] += dd * v[bi] / best;
    }
  } else {
    g_ra += dtb * 0.5 * cf.dt_visc / ra;
  }
  const double sq = sqrt(1.0 / ra) * sqrt(cf.pr);      // sqrt(Pr / Ra)
  g_ra += -0.5 * (csb * cf.cs + cpb * cf.cp + ctb * cf.ct) / ra;
  as->g_wi += wib - cpb * cf.cp / cf.wi;
  as->g_beta += csb * sq - cpb * sq / cf.wi;
  as->g_ra += g_ra;
  as->dt_bar = dtb;
  (void)beta;
}

// ------------------------------------------------------------------------------------------------
// Mirror buckets: for every particle c, the slots (i, k) whose neighbour is a mirror of c (its own
// mirrors included), sorted by (i, k).
// ------------------------------------------------------------------------------------------------
__global__ void k_adj_zero_int(int* a, int n) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) a[i] = 0;
}

__global__ void k_adj_bucket_count(int n, const int* cnt, const int* link, const int* irel, int* bcnt) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  for (int k = 0; k < cnt[i]; k++) {
    const int j = link[(size_t)k * n + i];
    if (j >= n) atomicAdd(&bcnt[irel[j]], 1);
  }
}

__global__ void k_adj_bucket_fill(int n, const int* cnt, const int* link, const int* irel, const int* boff,
                                  int* bfill, int* bkey) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  for (int k = 0; k < cnt[i]; k++) {
    const int j = link[(size_t)k * n + i];
    if (j >= n) {
      const int c = irel[j];
      bkey[boff[c] + atomicAdd(&bfill[c], 1)] = i * kNplink + k;
    }
  }
}

__global__ void k_adj_bucket_sort(int n, const int* boff, int* bkey) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= n) return;
  const int b = boff[c], e = boff[c + 1];
  for (int p = b + 1; p < e; p++) {
    const int key = bkey[p];
    int q = p - 1;
    while (q >= b && bkey[q] > key) { bkey[q + 1] = bkey[q]; q--; }
    bkey[q + 1] = key;
  }
}

__global__ void k_adj_lam_sum(const double* v, int n, double* partials) {
  RowRange rr = block_rows(n, gridDim.x);
  Red<1> R(nullptr, n);
  for (int i = rr.begin + threadIdx.x; i < rr.end; i += blockDim.x) R.add(0, i, v[i]);
  R.store(partials);
}

__global__ void k_adj_zero_scal(AdjScal* as) {
  if (threadIdx.x == 0 && blockIdx.x == 0) { as->g_wi = 0.0; as->g_beta = 0.0; as->g_ra = 0.0; as->dt_bar = 0.0; }
}

// ------------------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------------------
constexpr int kStateFields = 11;       // x, y, u, v, theta, tau11, tau12, tau21, tau22, trace, p
constexpr int kCotFields = 37;         // the n-vectors of Cot

}  // namespace vb

struct AdjointCtx {
  int n = 0, nsteps = 0, every = 0, nck = 0;
  double* ck = nullptr;   uint8_t* ck_in = nullptr;     // state at the start of steps 0, every, 2 every..
  double* pst = nullptr;                                // the pressure of every recorded step
  double* seg = nullptr;  uint8_t* seg_in = nullptr;    // the states of one segment, recomputed
  double* fin = nullptr;  uint8_t* fin_in = nullptr;    // the state after the last recorded step
  std::vector<double> ck_time, seg_time;
  std::vector<int> ck_itime, seg_itime;
  double fin_time = 0.0;
  int fin_itime = 0;
  vb::LinRec rec{};
  int *bcnt = nullptr, *boff = nullptr, *bfill = nullptr, *bkey = nullptr;
  vb::Cot b{};
  double* cot = nullptr;
  double* part = nullptr;
  vb::AdjScal* as = nullptr;
  vb::AdjScal* as_host = nullptr;
  std::vector<int> fwd_iters, adj_iters;
  std::vector<double> adj_resid;
  int adj_status = VB_OK;
  double* w64 = nullptr;   // mixed path: the step's FP32 weight planes widened (w_G x, w_G y, w_L)
};

namespace vb {

void free_adjoint(vb_ctx* c) {
  AdjointCtx* a = c->adj;
  if (!a) return;
  void* ptrs[] = {a->ck, a->ck_in, a->pst, a->seg, a->seg_in, a->fin, a->fin_in, a->rec.mjac, a->rec.msgn,
                  a->rec.ejac, a->rec.eflag, a->bcnt, a->boff, a->bfill, a->bkey, a->cot, a->b.S1, a->b.S7,
                  a->part, a->as, a->w64};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (a->as_host) (void)hipHostFree(a->as_host);
  delete a;
  c->adj = nullptr;
}

static double* field(vb_ctx* c, int q) {
  double* f[kStateFields] = {c->st.x, c->st.y, c->st.u, c->st.v, c->st.th, c->st.t11, c->st.t12, c->st.t21,
                             c->st.t22, c->st.trc, c->st.p};
  return f[q];
}

static vb_status save_state(vb_ctx* c, double* dst, uint8_t* dst_in) {
  const size_t n = c->npfb;
  for (int q = 0; q < kStateFields; q++)
    VB_HIP(c, hipMemcpyAsync(dst + q * n, field(c, q), n * 8, hipMemcpyDeviceToDevice, c->stream));
  VB_HIP(c, hipMemcpyAsync(dst_in, c->st.inbin, n, hipMemcpyDeviceToDevice, c->stream));
  return VB_OK;
}

static vb_status restore_state(vb_ctx* c, const double* src, const uint8_t* src_in, double time, int itime) {
  const size_t n = c->npfb;
  for (int q = 0; q < kStateFields; q++)
    VB_HIP(c, hipMemcpyAsync(field(c, q), src + q * n, n * 8, hipMemcpyDeviceToDevice, c->stream));
  VB_HIP(c, hipMemcpyAsync(c->st.inbin, src_in, n, hipMemcpyDeviceToDevice, c->stream));
  c->time = time;
  c->itime = itime;
  return VB_OK;
}

static bool warn_or_ok(vb_status s) { return s == VB_OK || s == VB_WARN_MAXITER || s == VB_WARN_BICGSTAB_BREAKDOWN; }

// div_free.F90:38-117 up to the velocity update's input: phases 1-9
static vb_status phases_1_9(vb_ctx* c) {
  vb_status s;
  if ((s = vb_set_time_step(c)) != VB_OK) return s;
  if ((s = vb_put_in_bin(c)) != VB_OK) return s;
  if ((s = vb_create_mirror_particles(c)) != VB_OK) return s;
  if ((s = vb_particle_allocation(c)) != VB_OK) return s;
  if ((s = vb_kernel_calculation(c)) != VB_OK) return s;
  if ((s = vb_locate_free_surf_part(c)) != VB_OK) return s;
  if ((s = vb_new_fick_shift(c)) != VB_OK) return s;
  if ((s = vb_prediction_step(c)) != VB_OK) return s;
  return vb_mirror_velocities(c);
}

// One step of the record (p_in null: the PPE solve, its pressure kept in p_out) or of the segment
// recompute (the stored pressure p_in instead of the solve).
static vb_status forward_step(vb_ctx* c, double tol, int itermax, const double* p_in, double* p_out, int* iters) {
  vb_status s = phases_1_9(c), ls = VB_OK;
  if (s != VB_OK) return s;
  const size_t n = c->npfb;
  if (p_in) {
    VB_HIP(c, hipMemcpyAsync(c->st.p, p_in, n * 8, hipMemcpyDeviceToDevice, c->stream));
  } else {
    ls = vb_ppe_solve(c, tol, itermax, iters, nullptr);
    if (!warn_or_ok(ls)) return ls;
    if (p_out) VB_HIP(c, hipMemcpyAsync(p_out, c->st.p, n * 8, hipMemcpyDeviceToDevice, c->stream));
  }
  if ((s = vb_pressure_update(c)) != VB_OK) return s;
  if ((s = vb_final_advection(c)) != VB_OK) return s;
  return ls;
}

// b = a widened (grid-stride)
__global__ void k_adj_widen(const float* a, double* b, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

static vb_status adjoint_alloc(vb_ctx* c, AdjointCtx* a) {
  const size_t n = c->npfb, nw = n * kNplink, mcap = c->mcap;
  const size_t sz = kStateFields * n;
  VB_HIP(c, hipMalloc((void**)&a->ck, (size_t)a->nck * sz * 8));
  VB_HIP(c, hipMalloc((void**)&a->ck_in, (size_t)a->nck * n));
  VB_HIP(c, hipMalloc((void**)&a->pst, (size_t)a->nsteps * n * 8));
  VB_HIP(c, hipMalloc((void**)&a->seg, (size_t)a->every * sz * 8));
  VB_HIP(c, hipMalloc((void**)&a->seg_in, (size_t)a->every * n));
  VB_HIP(c, hipMalloc((void**)&a->fin, sz * 8));
  VB_HIP(c, hipMalloc((void**)&a->fin_in, n));
  VB_HIP(c, hipMalloc((void**)&a->rec.mjac, 6 * mcap * 8));
  VB_HIP(c, hipMalloc((void**)&a->rec.msgn, mcap * 4));
  VB_HIP(c, hipMalloc((void**)&a->rec.ejac, 4 * n * 8));
  VB_HIP(c, hipMalloc((void**)&a->rec.eflag, n * 4));
  VB_HIP(c, hipMalloc((void**)&a->bcnt, (n + 1) * 4));
  VB_HIP(c, hipMalloc((void**)&a->boff, (n + 1) * 4));
  VB_HIP(c, hipMalloc((void**)&a->bfill, (n + 1) * 4));
  VB_HIP(c, hipMalloc((void**)&a->bkey, nw * 4));
  VB_HIP(c, hipMalloc((void**)&a->cot, (size_t)kCotFields * n * 8));
  VB_HIP(c, hipMalloc((void**)&a->b.S1, nw * 8));
  VB_HIP(c, hipMalloc((void**)&a->b.S7, 7 * nw * 8));
  VB_HIP(c, hipMalloc((void**)&a->part, (size_t)AQ_N * kPartialStride * 8));
  VB_HIP(c, hipMalloc((void**)&a->as, sizeof(AdjScal)));
  if (c->wf32) VB_HIP(c, hipMalloc((void**)&a->w64, 3 * nw * 8));
  VB_HIP(c, hipHostMalloc((void**)&a->as_host, sizeof(AdjScal)));
  double** f[kCotFields] = {&a->b.x, &a->b.y, &a->b.u, &a->b.v, &a->b.th, &a->b.t11, &a->b.t12, &a->b.t21,
                            &a->b.t22, &a->b.trc, &a->b.us, &a->b.vs, &a->b.u0, &a->b.v0, &a->b.x0, &a->b.y0,
                            &a->b.drx, &a->b.dry, &a->b.gpx, &a->b.gpy, &a->b.k11, &a->b.k12, &a->b.k21,
                            &a->b.k22, &a->b.db, &a->b.ax, &a->b.ay, &a->b.at, &a->b.gux, &a->b.guy,
                            &a->b.gvx, &a->b.gvy, &a->b.ua, &a->b.va, &a->b.tha, &a->b.x1, &a->b.y1};
  for (int q = 0; q < kCotFields; q++) *f[q] = a->cot + (size_t)q * n;
  a->b.nw = nw;
  // the operator's transposed map, buckets and adjoint-solve vectors, so the sweep allocates nothing
  vb_status s = alloc_operator(c, nw);
  return s;
}

static vb_status adjoint_record(vb_ctx* c, int nsteps, int every, double tol, int itermax, int* iters) {
  if (c->g.open || c->g.ncirc > 0 || c->g.shift_coeff != 0) {
    c->err = "adjoint: closed domains without circles and with i_shift_coeff = 0 only";
    return VB_ERR_ARG;
  }
  if (nsteps < 1 || every < 1) { c->err = "adjoint: nsteps and ckpt_every must be positive"; return VB_ERR_ARG; }
  free_adjoint(c);
  AdjointCtx* a = new AdjointCtx();
  c->adj = a;
  a->n = c->npfb; a->nsteps = nsteps; a->every = std::min(every, nsteps);
  a->nck = (nsteps + a->every - 1) / a->every;
  vb_status s = adjoint_alloc(c, a);
  if (s != VB_OK) { free_adjoint(c); return s; }
  const size_t n = c->npfb, sz = kStateFields * n;
  vb_status worst = VB_OK;
  a->fwd_iters.assign(nsteps, 0);
  for (int k = 0; k < nsteps; k++) {
    if (k % a->every == 0) {
      const int q = k / a->every;
      if ((s = save_state(c, a->ck + q * sz, a->ck_in + q * n)) != VB_OK) return s;
      a->ck_time.push_back(c->time); a->ck_itime.push_back(c->itime);
    }
    int it = 0;
    s = forward_step(c, tol, itermax, nullptr, a->pst + (size_t)k * n, &it);
    if (!warn_or_ok(s)) { free_adjoint(c); return s; }
    if (s != VB_OK) worst = s;
    a->fwd_iters[k] = it;
    if (iters) iters[k] = it;
  }
  if ((s = save_state(c, a->fin, a->fin_in)) != VB_OK) return s;
  a->fin_time = c->time; a->fin_itime = c->itime;
  VB_HIP(c, hipStreamSynchronize(c->stream));
  return worst;
}

// The adjoint of step k (its start state already in the context): re-run phases 1..10 with the
// Jacobians recorded and the stored pressure, then the reverse kernels.
static vb_status reverse_step(vb_ctx* c, int k, const double* tau0, double tol, int itermax) {
  AdjointCtx* a = c->adj;
  const int n = c->npfb;
  c->rec = a->rec;
  vb_status s = phases_1_9(c);
  c->rec = LinRec{};
  if (s != VB_OK) return s;
  if ((s = assemble_ppe(c)) != VB_OK) return s;
  VB_HIP(c, hipMemcpyAsync(c->st.p, a->pst + (size_t)k * n, (size_t)n * 8, hipMemcpyDeviceToDevice, c->stream));
  if ((s = launch_mirror_p(c)) != VB_OK) return s;
  if ((s = build_transpose(c)) != VB_OK) return s;
  // the mirror buckets
  VB_LAUNCH(c, k_adj_zero_int, grid_1d(n + 1), a->bcnt, n + 1);
  VB_LAUNCH(c, k_adj_zero_int, grid_1d(n + 1), a->bfill, n + 1);
  VB_LAUNCH(c, k_adj_bucket_count, grid_1d(n), n, c->cnt, c->link, c->st.irel, a->bcnt);
  if ((s = launch_exclusive_scan(c, a->bcnt, a->boff, n)) != VB_OK) return s;
  VB_LAUNCH(c, k_adj_bucket_fill, grid_1d(n), n, c->cnt, c->link, c->st.irel, a->boff, a->bfill, a->bkey);
  VB_LAUNCH(c, k_adj_bucket_sort, grid_1d(n), n, a->boff, a->bkey);

  const vb_params& p = c->prm;
  Coef cf;
  cf.newtonian = p.fluid_model == 1;
  cf.cs = p.beta * sqrt(p.Pr / p.Ra);
  cf.cp = cf.newtonian ? 0.0 : (1.0 - p.beta) * p.Pr / (p.Ra * p.El);
  cf.ct = 1.0 / sqrt(p.Ra * p.Pr);
  cf.wi = cf.newtonian ? 1.0 : p.El * sqrt(p.Ra / p.Pr);
  cf.pr = p.Pr;
  cf.eps = p.ve_nonlinearity;
  cf.dt_visc = c->dt_visc;
  cf.cfl_h = p.dt_coef_cfl * p.h;
  cf.use_kick = p.reference_kick && c->itime != 2;
  Fwd F;
  F.xs = c->st.x; F.ys = c->st.y; F.us = c->st.u; F.vs = c->st.v;
  F.ua = c->u_alt; F.va = c->v_alt; F.tha = c->th_alt;
  F.t11 = c->st.t11; F.t12 = c->st.t12; F.t21 = c->st.t21; F.t22 = c->st.t22;
  for (int q = 0; q < 4; q++) F.tau0[q] = tau0 + (size_t)(5 + q) * n;
  F.trc0 = tau0 + (size_t)9 * n;
  F.p = c->st.p; F.dP = c->st.dP;
  F.u0 = c->st.u0; F.v0 = c->st.v0; F.x0 = c->st.x0; F.y0 = c->st.y0;
  F.inbin = c->st.inbin; F.irel = c->st.irel; F.vrel = c->st.vrel;
  F.cnt = c->cnt; F.link = c->link; F.wg1 = c->wg1; F.wg2 = c->wg2; F.wl = c->wl;
  if (c->wf32) {          // the reverse kernels read FP64 weights: widen the mixed path's FP32 planes
    const size_t nw = (size_t)n * kNplink;
    const int nwb = (int)std::min<size_t>((nw + kBlock - 1) / kBlock, 65535);
    VB_LAUNCH(c, k_adj_widen, nwb, c->wg1f, a->w64, nw);
    VB_LAUNCH(c, k_adj_widen, nwb, c->wg2f, a->w64 + nw, nw);
    VB_LAUNCH(c, k_adj_widen, nwb, c->wlf, a->w64 + 2 * nw, nw);
    F.wg1 = a->w64; F.wg2 = a->w64 + nw; F.wl = a->w64 + 2 * nw;
  }
  F.k11 = c->k11; F.k12 = c->k12; F.k21 = c->k21; F.k22 = c->k22; F.kick = c->kick;
  Nbr N{c->cnt, c->link, c->op.rslot, a->boff, a->bkey, n};
  Cot b = a->b;
  b.pb = c->op.g;
  const int nb = grid_for(n), ng = grid_1d(n);
  double* part = a->part;

  VB_LAUNCH(c, k_adj_final, nb, c->g, F, b, c->scal, part);
  VB_LAUNCH(c, k_adj_gather_p, ng, N, b.S1, b.pb);
  int ait = 0, ast = VB_OK;
  double ares = 0.0;
  if ((s = adjoint_solve(c, tol, itermax, &ait, &ares, &ast)) != VB_OK) return s;   // lambda in op.lam
  a->adj_iters[k] = ait;
  a->adj_resid[k] = ares;
  if (ast != VB_OK) a->adj_status = ast;
  const double* lam = c->op.lam;
  VB_LAUNCH(c, k_adj_lam_sum, nb, lam, n, part + (size_t)AQ_LAM * kPartialStride);
  VB_LAUNCH(c, k_adj_ppe_rhs, nb, c->g, F, b, lam, part + (size_t)AQ_LAM * kPartialStride, nb, c->scal, part);
  VB_LAUNCH(c, k_adj_gather_us, ng, c->g, N, c->st.vrel, b.S7, b.nw, b.us, b.vs);
  VB_LAUNCH(c, k_adj_rhs, nb, c->g, F, b, cf, c->scal, part);
  VB_LAUNCH(c, k_adj_gather_rhs, ng, c->g, N, c->st.vrel, a->rec.msgn, b.S7, b.nw, b);
  VB_LAUNCH(c, k_adj_logconf, nb, c->g, F, b, cf, c->scal, part);
  VB_LAUNCH(c, k_adj_gather_u, ng, N, a->rec.msgn, b.S7, b.nw, b.ua, b.va);
  VB_LAUNCH(c, k_adj_geometry, nb, c->g, F, b, cf, lam, c->scal);
  VB_LAUNCH(c, k_adj_gather_x, ng, N, a->rec, b.S1, b.S7, b.nw, lam, F.wl, b.x1, b.y1);
  VB_LAUNCH(c, k_adj_advect, nb, c->g, F, b, a->rec, c->scal, part);
  hipLaunchKernelGGL(k_adj_scalars, dim3(1), dim3(1024), 0, c->stream, c->g, cf, p.Ra, p.beta, part, nb,
                     c->st.u0, c->st.v0, c->scal, b, a->as);
  VB_HIP(c, hipGetLastError());
  return VB_OK;
}

static vb_status adjoint_backward(vb_ctx* c, const double* const* fbar, double tol, int itermax, double* grads,
                                  double* const* ibar, int* adj_iters) {
  AdjointCtx* a = c->adj;
  if (!a) { c->err = "adjoint: no recorded forward run (call vb_adjoint_record first)"; return VB_ERR_ARG; }
  const size_t n = c->npfb, sz = kStateFields * n;
  double* cot_state[10] = {a->b.x, a->b.y, a->b.u, a->b.v, a->b.th, a->b.t11, a->b.t12, a->b.t21, a->b.t22,
                           a->b.trc};
  for (int q = 0; q < 10; q++) {
    if (fbar[q]) VB_HIP(c, hipMemcpyAsync(cot_state[q], fbar[q], n * 8, hipMemcpyHostToDevice, c->stream));
    else VB_HIP(c, hipMemsetAsync(cot_state[q], 0, n * 8, c->stream));
  }
  VB_LAUNCH(c, k_adj_zero_scal, 1, a->as);
  a->adj_status = VB_OK;
  a->adj_iters.assign(a->nsteps, 0);
  a->adj_resid.assign(a->nsteps, 0.0);
  a->seg_time.assign(a->every, 0.0);
  a->seg_itime.assign(a->every, 0);
  vb_status s;
  for (int q = a->nck - 1; q >= 0; q--) {
    const int c0 = q * a->every, c1 = std::min(c0 + a->every, a->nsteps);
    if ((s = restore_state(c, a->ck + q * sz, a->ck_in + q * n, a->ck_time[q], a->ck_itime[q])) != VB_OK) return s;
    for (int k = c0; k < c1; k++) {
      const int r = k - c0;
      if ((s = save_state(c, a->seg + r * sz, a->seg_in + r * n)) != VB_OK) return s;
      a->seg_time[r] = c->time; a->seg_itime[r] = c->itime;
      if (k + 1 < c1) {
        s = forward_step(c, 0.0, 0, a->pst + (size_t)k * n, nullptr, nullptr);
        if (!warn_or_ok(s)) return s;
      }
    }
    for (int k = c1 - 1; k >= c0; k--) {
      const int r = k - c0;
      if ((s = restore_state(c, a->seg + r * sz, a->seg_in + r * n, a->seg_time[r], a->seg_itime[r])) != VB_OK) return s;
      if ((s = reverse_step(c, k, a->seg + r * sz, tol, itermax)) != VB_OK) return s;
    }
  }
  // the gradients and the start state's cotangents; the context back at the recorded final state
  VB_HIP(c, hipMemcpyAsync(a->as_host, a->as, sizeof(AdjScal), hipMemcpyDeviceToHost, c->stream));
  for (int q = 0; q < 10; q++)
    if (ibar && ibar[q]) VB_HIP(c, hipMemcpyAsync(ibar[q], cot_state[q], n * 8, hipMemcpyDeviceToHost, c->stream));
  if ((s = restore_state(c, a->fin, a->fin_in, a->fin_time, a->fin_itime)) != VB_OK) return s;
  VB_HIP(c, hipStreamSynchronize(c->stream));
  grads[0] = a->as_host->g_wi; grads[1] = a->as_host->g_beta; grads[2] = a->as_host->g_ra;
  if (adj_iters) for (int k = 0; k < a->nsteps; k++) adj_iters[k] = a->adj_iters[k];
  return (vb_status)a->adj_status;
}

}  // namespace vb

using namespace vb;

extern "C" {

vb_status vb_adjoint_record(vb_ctx* c, int32_t nsteps, int32_t ckpt_every, double tol, int32_t itermax,
                            int32_t* ppe_iters) {
  if (!c) return VB_ERR_ARG;
  if (tol <= 0.0) tol = 1.0e-8;
  if (itermax <= 0) itermax = 1000;
  return adjoint_record(c, nsteps, ckpt_every, tol, itermax, ppe_iters);
}

vb_status vb_adjoint_backward(vb_ctx* c, int32_t n, const double* x_bar, const double* y_bar, const double* u_bar,
                              const double* v_bar, const double* theta_bar, const double* tau11_bar,
                              const double* tau12_bar, const double* tau21_bar, const double* tau22_bar,
                              const double* trace_bar, double tol, int32_t itermax, double* grads3,
                              double* x_bar0, double* y_bar0, double* u_bar0, double* v_bar0, double* theta_bar0,
                              double* tau11_bar0, double* tau12_bar0, double* tau21_bar0, double* tau22_bar0,
                              double* trace_bar0, int32_t* adj_iters) {
  if (!c || !grads3) return VB_ERR_ARG;
  if (n != c->npfb) { c->err = "vb_adjoint_backward: n must equal the context's npfb"; return VB_ERR_ARG; }
  if (tol <= 0.0) tol = 1.0e-8;
  if (itermax <= 0) itermax = 1000;
  const double* fbar[10] = {x_bar, y_bar, u_bar, v_bar, theta_bar, tau11_bar, tau12_bar, tau21_bar, tau22_bar,
                            trace_bar};
  double* ibar[10] = {x_bar0, y_bar0, u_bar0, v_bar0, theta_bar0, tau11_bar0, tau12_bar0, tau21_bar0, tau22_bar0,
                      trace_bar0};
  return adjoint_backward(c, fbar, tol, itermax, grads3, ibar, adj_iters);
}

}  // extern "C"
