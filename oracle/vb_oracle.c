/* CPU restatement of the reference's per-step path: jrcking/vebisph div_free.F90:22-146 and callees.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; never part of the product path).  Each function cites the
 * reference file:line it restates.  FP64, single thread, the reference's expression order; built
 * with -ffp-contract=off it matches oracle/_ref (the reference itself) bit for bit on the tests'
 * inputs.  Indices are 0-based: the reference's "i <= nbS" is "i < nbS" here, its vrelation codes
 * (patch ids, 999+circle id, -corner id) keep their 1-based meaning.
 */
#include "vb_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NPLINK 75                       /* common_parameter.F90:5 */
#define PI 3.14159265358979323846       /* common_parameter.F90:12 */
#define EPS_DBL 2.220446049250313e-16   /* Fortran epsilon(1d0), predictor_mod.F90:268 */

typedef struct { int n, cap; int* i; double* t; double* d; } plist;   /* near-patch lists */

static double* make_kick(int n);

struct vbo_ctx {
  /* parameters (input.F90:18-37) and derived constants (input.F90:164-172) */
  double h, dx, dv, Ra, Pr, El, beta, bf[2], cfl, visc, eps_ptt, time, dt, umax;
  int npfb, nbS, open_domain, shift_coeff, itime, cap;
  int const_mod;          /* the reference's build: 1 Newtonian, else viscoelastic */
  double sup, sup2, uno, ad7, ad7h, eta2;
  double xmint, xmaxt, ymint, ymaxt;
  int ncx, ncy, nct;
  /* boundaries (input.F90:51-81, 255-273) */
  int nbp, ncirc;
  double *bnode, *bedge, *bvel, *bth, *bang;
  int *btype, *bpp;
  double *ccen, *crad, *comg, *cvel, *cth;
  /* particle state, capacity cap rows */
  double *x, *y, *u, *v, *th, *t11, *t12, *t21, *t22, *trc, *p;
  unsigned char* inbin;
  int* nsurf;
  double *conc, *divr;    /* open domain: calc_part_dist_vars (freesurf_mod.F90:70-103) */
  /* per-step data */
  double *x0, *y0, *u0, *v0;
  int np, nmirror, nmirror_esti, n_inbin, maxneigh, ls_iters, status;
  double ls_resid;
  int *irel, *vrel;       /* indexed by mirror row (npfb..np-1) */
  double* dP;
  int *cellpart, *ist, *ip;
  int* cnt;
  int* link;              /* slot-major: link[k*npfb + i], 0-based j */
  double *wg1, *wg2, *wl; /* slot-major like link */
  double *k11, *k12, *k21, *k22;
  double *usx, *usy;
  double *snx, *sny, *psmooth;   /* open domain: surf_norm (smoothed), ppe_smooth */
  double* gvel;           /* (npfb,4): du/dx du/dy dv/dx dv/dy */
  double *ppe_a, *ppe_b; int* ppe_ija; int ppe_nnz;
  double *gpx, *gpy;
  double* kick;           /* per-particle B perturbation, see make_kick */
};

static void* xcalloc(size_t n, size_t s) { void* p = calloc(n ? n : 1, s); return p; }

/* ----------------------------------------------------------------------------------------------
 * sphtools.F90:11-76 — quintic spline and its gradient factor.  Integer powers are left-to-right
 * products (q^4 = ((q*q)*q)*q), the order the reference's compiler emits for q**4 and q**5.
 * -------------------------------------------------------------------------------------------- */
static inline double p4(double q) { return ((q * q) * q) * q; }
static inline double p5(double q) { return (((q * q) * q) * q) * q; }

static inline double fac(double qq, double ad7h) {          /* sphtools.F90:11-44 */
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

static inline double wab(double qq, double ad7) {           /* sphtools.F90:46-76 */
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

void vbo_fac_wab(int n, const double* q, double ad_7, double ad_7h, double* f, double* w) {
  for (int i = 0; i < n; i++) { f[i] = fac(q[i], ad_7h); w[i] = wab(q[i], ad_7); }
}

/* ---------------------------------------------------------------------------------------------- */
vbo_ctx* vbo_create(const int* iprm, const double* prm, int nbp, const double* b_node,
                    const int* b_type, const double* b_vel, const double* b_thermal,
                    const int* b_pp, int ncirc, const double* c_centre, const double* c_radius,
                    const double* c_omega, const double* c_vel, const double* c_thermal) {
  vbo_ctx* c = (vbo_ctx*)xcalloc(1, sizeof(vbo_ctx));
  c->h = prm[0]; c->dx = prm[1]; c->dv = prm[2]; c->Ra = prm[3]; c->Pr = prm[4]; c->El = prm[5];
  c->beta = prm[6]; c->bf[0] = prm[7]; c->bf[1] = prm[8]; c->cfl = prm[9]; c->visc = prm[10];
  c->eps_ptt = prm[11]; c->time = prm[12];
  c->npfb = iprm[0]; c->nbS = iprm[1]; c->open_domain = iprm[2]; c->shift_coeff = iprm[3];
  c->const_mod = iprm[6];
  c->itime = iprm[5];
  /* mirror rows need up to nmirror_esti extra rows (mirror_boundaries_mod.F90:53-54) */
  int esti = (c->npfb <= 500) ? 10 * c->npfb : 2 * c->npfb;
  c->cap = c->npfb + esti;
  if (iprm[4] > c->cap) c->cap = iprm[4];
  c->np = c->npfb;
  /* derived constants, input.F90:164-172 */
  double h = c->h, h2 = h * h;
  c->sup = 3.0 * h; c->sup2 = 9.0 * h2; c->uno = 1.0 / (3.0 * h);
  c->ad7 = 7.0 / 478.0 / PI / h / h; c->ad7h = c->ad7 / h; c->eta2 = h * h * 1.0e-10;
  /* boundaries */
  c->nbp = nbp; c->ncirc = ncirc;
  c->bnode = (double*)xcalloc(2 * nbp, 8); c->bedge = (double*)xcalloc(2 * nbp, 8);
  c->bvel = (double*)xcalloc(nbp, 8); c->bth = (double*)xcalloc(nbp, 8); c->bang = (double*)xcalloc(nbp, 8);
  c->btype = (int*)xcalloc(nbp, 4); c->bpp = (int*)xcalloc(nbp, 4);
  for (int i = 0; i < nbp; i++) {
    c->bnode[2 * i] = b_node[2 * i]; c->bnode[2 * i + 1] = b_node[2 * i + 1];
    c->bvel[i] = b_vel[i]; c->bth[i] = b_thermal[i]; c->btype[i] = b_type[i]; c->bpp[i] = b_pp[i];
  }
  for (int i = 0; i < nbp; i++) {   /* datclass.F90:730-744 */
    int ip1 = (i + 1) % nbp;
    c->bedge[2 * i] = c->bnode[2 * ip1] - c->bnode[2 * i];
    c->bedge[2 * i + 1] = c->bnode[2 * ip1 + 1] - c->bnode[2 * i + 1];
  }
  for (int i = 0; i < nbp; i++) {   /* input.F90:255-273 corner angles */
    int im1 = (i + nbp - 1) % nbp, ip1 = (i + 1) % nbp;
    const double* e = c->bedge + 2 * im1;
    double pd = e[1] * c->bnode[2 * ip1] - e[0] * c->bnode[2 * ip1 + 1] +
                c->bnode[2 * i] * c->bnode[2 * im1 + 1] - c->bnode[2 * i + 1] * c->bnode[2 * im1];
    pd = -1.0 * pd / sqrt(e[0] * e[0] + e[1] * e[1]);
    double t1 = sqrt(c->bedge[2 * i] * c->bedge[2 * i] + c->bedge[2 * i + 1] * c->bedge[2 * i + 1]);
    if (pd <= 0.0 && fabs(pd) > fabs(t1)) pd = -t1;
    else if (pd > 0.0 && fabs(pd) > fabs(t1)) pd = t1;
    c->bang[i] = asin(pd / t1);
  }
  if (ncirc) {
    c->ccen = (double*)xcalloc(2 * ncirc, 8); c->cvel = (double*)xcalloc(2 * ncirc, 8);
    c->crad = (double*)xcalloc(ncirc, 8); c->comg = (double*)xcalloc(ncirc, 8); c->cth = (double*)xcalloc(ncirc, 8);
    memcpy(c->ccen, c_centre, 16 * ncirc); memcpy(c->cvel, c_vel, 16 * ncirc);
    memcpy(c->crad, c_radius, 8 * ncirc); memcpy(c->comg, c_omega, 8 * ncirc); memcpy(c->cth, c_thermal, 8 * ncirc);
  }
  /* padded box and cells, sph2D_incom.F90:64-70 and cell_griddin.F90:13-24 */
  double xbmin = c->bnode[0], xbmax = c->bnode[0], ybmin = c->bnode[1], ybmax = c->bnode[1];
  for (int i = 1; i < nbp; i++) {
    if (c->bnode[2 * i] < xbmin) xbmin = c->bnode[2 * i];
    if (c->bnode[2 * i] > xbmax) xbmax = c->bnode[2 * i];
    if (c->bnode[2 * i + 1] < ybmin) ybmin = c->bnode[2 * i + 1];
    if (c->bnode[2 * i + 1] > ybmax) ybmax = c->bnode[2 * i + 1];
  }
  c->xmint = xbmin - c->sup * 1.5; c->xmaxt = xbmax + c->sup * 1.5;
  c->ymint = ybmin - c->sup * 1.5; c->ymaxt = ybmax + c->sup * 1.5;
  c->ncx = (int)((c->xmaxt - c->xmint) * c->uno) + 1;
  c->ncy = (int)((c->ymaxt - c->ymint) * c->uno) + 1;
  c->nct = c->ncx * c->ncy;
  /* particle arrays (input.F90:105-114); trace_conf starts at 2 */
  size_t n = (size_t)c->cap;
  c->x = (double*)xcalloc(n, 8); c->y = (double*)xcalloc(n, 8); c->u = (double*)xcalloc(n, 8);
  c->v = (double*)xcalloc(n, 8); c->th = (double*)xcalloc(n, 8); c->p = (double*)xcalloc(n, 8);
  c->t11 = (double*)xcalloc(n, 8); c->t12 = (double*)xcalloc(n, 8); c->t21 = (double*)xcalloc(n, 8);
  c->t22 = (double*)xcalloc(n, 8); c->trc = (double*)xcalloc(n, 8);
  for (size_t i = 0; i < n; i++) c->trc[i] = 2.0;
  c->inbin = (unsigned char*)xcalloc(n, 1); c->nsurf = (int*)xcalloc(n, 4);
  c->conc = (double*)xcalloc(n, 8); c->divr = (double*)xcalloc(n, 8);
  c->irel = (int*)xcalloc(n, 4); c->vrel = (int*)xcalloc(n, 4); c->dP = (double*)xcalloc(n, 8);
  c->kick = make_kick(c->npfb);
  return c;
}

static void free_step(vbo_ctx* c) {
  free(c->x0); free(c->y0); free(c->u0); free(c->v0); c->x0 = c->y0 = c->u0 = c->v0 = NULL;
  free(c->cellpart); free(c->ist); free(c->ip); c->cellpart = c->ist = c->ip = NULL;
  free(c->cnt); free(c->link); c->cnt = c->link = NULL;
  free(c->wg1); free(c->wg2); free(c->wl); c->wg1 = c->wg2 = c->wl = NULL;
  free(c->k11); free(c->k12); free(c->k21); free(c->k22); c->k11 = c->k12 = c->k21 = c->k22 = NULL;
  free(c->usx); free(c->usy); c->usx = c->usy = NULL;
  free(c->snx); free(c->sny); free(c->psmooth); c->snx = c->sny = c->psmooth = NULL;
  free(c->gvel); c->gvel = NULL;
  free(c->ppe_a); free(c->ppe_b); free(c->ppe_ija); c->ppe_a = c->ppe_b = NULL; c->ppe_ija = NULL;
  free(c->gpx); free(c->gpy); c->gpx = c->gpy = NULL;
}

void vbo_destroy(vbo_ctx* c) {
  if (!c) return;
  free_step(c);
  free(c->bnode); free(c->bedge); free(c->bvel); free(c->bth); free(c->bang); free(c->btype); free(c->bpp);
  free(c->ccen); free(c->crad); free(c->comg); free(c->cvel); free(c->cth);
  free(c->x); free(c->y); free(c->u); free(c->v); free(c->th); free(c->p);
  free(c->t11); free(c->t12); free(c->t21); free(c->t22); free(c->trc);
  free(c->inbin); free(c->nsurf); free(c->conc); free(c->divr); free(c->irel); free(c->vrel); free(c->dP); free(c->kick);
  free(c);
}

void vbo_set_particles(vbo_ctx* c, int n, const double* x, const double* y, const double* u,
                       const double* v, const double* th, const double* tau, const double* trc,
                       const double* p, const int* inbin) {
  for (int i = 0; i < n; i++) {
    c->x[i] = x[i]; c->y[i] = y[i]; c->u[i] = u[i]; c->v[i] = v[i]; c->th[i] = th[i];
    c->t11[i] = tau[4 * i]; c->t12[i] = tau[4 * i + 1]; c->t21[i] = tau[4 * i + 2]; c->t22[i] = tau[4 * i + 3];
    c->trc[i] = trc[i]; c->p[i] = p[i]; c->inbin[i] = inbin[i] ? 1 : 0;
  }
}

void vbo_get_particles(const vbo_ctx* c, int n, double* x, double* y, double* u, double* v,
                       double* th, double* tau, double* trc, double* p, int* inbin) {
  for (int i = 0; i < n; i++) {
    x[i] = c->x[i]; y[i] = c->y[i]; u[i] = c->u[i]; v[i] = c->v[i]; th[i] = c->th[i];
    tau[4 * i] = c->t11[i]; tau[4 * i + 1] = c->t12[i]; tau[4 * i + 2] = c->t21[i]; tau[4 * i + 3] = c->t22[i];
    trc[i] = c->trc[i]; p[i] = c->p[i]; inbin[i] = c->inbin[i];
  }
}

void vbo_get_scalars(const vbo_ctx* c, int* ints, double* dbls) {
  int iv[12] = {c->np, c->npfb, c->nbS, c->nmirror, c->maxneigh, c->ls_iters, c->itime,
                c->n_inbin, c->ncx, c->ncy, c->nct, c->status};
  double dv[8] = {c->dt, c->ls_resid, c->time, c->umax, c->xmint, c->xmaxt, c->ymint, c->ymaxt};
  memcpy(ints, iv, sizeof iv); memcpy(dbls, dv, sizeof dv);
}

/* ----------------------------------------------------------------------------------------------
 * Boundary geometry helpers, mirror_boundaries_mod.F90:447-491
 * -------------------------------------------------------------------------------------------- */
static double bound_dist(const vbo_ctx* c, double px, double py, int ib) {
  int ip1 = (ib + 1) % c->nbp;
  const double* e = c->bedge + 2 * ib;
  const double* n0 = c->bnode + 2 * ib;
  const double* n1 = c->bnode + 2 * ip1;
  double pd = e[1] * px - e[0] * py + n1[0] * n0[1] - n1[1] * n0[0];
  return -1.0 * pd / sqrt(e[0] * e[0] + e[1] * e[1]);
}

static void bound_norm(const vbo_ctx* c, int ib, double* nrm) {
  const double* e = c->bedge + 2 * ib;
  double t = sqrt(e[0] * e[0] + e[1] * e[1]);
  nrm[0] = -1.0 * e[1] / t;
  nrm[1] = e[0] / t;
}

static double pos_along(const vbo_ctx* c, double px, double py, int ib, double pd, const double* nb) {
  const double* e = c->bedge + 2 * ib;
  const double* n0 = c->bnode + 2 * ib;
  if (fabs(e[0]) > fabs(e[1])) return (px - n0[0] - pd * nb[0]) / e[0];
  return (py - n0[1] - pd * nb[1]) / e[1];
}

/* ----------------------------------------------------------------------------------------------
 * Phase 1: div_free.F90:39-53 with set_time_step div_free.F90:126-146
 * -------------------------------------------------------------------------------------------- */
static void phase_advect(vbo_ctx* c) {
  free_step(c);
  c->itime += 1;
  int n = c->npfb;
  c->x0 = (double*)xcalloc(n, 8); c->y0 = (double*)xcalloc(n, 8);
  c->u0 = (double*)xcalloc(n, 8); c->v0 = (double*)xcalloc(n, 8);
  for (int i = 0; i < n; i++) { c->u0[i] = c->u[i]; c->v0[i] = c->v[i]; c->x0[i] = c->x[i]; c->y0[i] = c->y[i]; }
  double umax = 1.0e-10;
  for (int i = c->nbS; i < n; i++) {
    double ua = sqrt(c->u[i] * c->u[i] + c->v[i] * c->v[i]);
    if (ua > umax) umax = ua;
  }
  c->umax = umax;
  double dt_cfl = c->cfl * c->h / umax;
  double a = sqrt(c->Ra / c->Pr), b = sqrt(c->Ra * c->Pr);
  double dt_visc = c->visc * c->h * c->h * (a < b ? a : b);
  c->dt = dt_cfl < dt_visc ? dt_cfl : dt_visc;
  for (int i = c->nbS; i < n; i++) {
    c->x[i] = c->x0[i] + c->u0[i] * c->dt;
    c->y[i] = c->y0[i] + c->v0[i] * c->dt;
  }
}

/* Phase 2: put_in_bin.F90:1-56 */
static int phase_bin(vbo_ctx* c) {
  for (int i = c->nbS; i < c->npfb; i++) {
    if (!c->inbin[i]) {
      if (c->x[i] < c->xmint || c->x[i] > c->xmaxt || c->y[i] < c->ymint || c->y[i] > c->ymaxt) c->inbin[i] = 1;
      if (isnan(c->x[i]) || isnan(c->y[i])) c->inbin[i] = 1;
    }
  }
  c->n_inbin = 0;
  for (int i = c->nbS; i < c->npfb; i++) {
    if (c->inbin[i]) {
      c->x[i] = 2.0 * c->xmint - c->xmaxt;
      c->y[i] = 2.0 * c->ymint - c->ymaxt;
      c->u[i] = 0.0; c->v[i] = 0.0;
      c->n_inbin++;
    }
  }
  if (c->n_inbin == c->npfb - c->nbS) return VBO_ERR_ALL_BINNED;
  return VBO_OK;
}

/* ----------------------------------------------------------------------------------------------
 * Phase 3: create_mirror_particles, mirror_boundaries_mod.F90:18-445
 * -------------------------------------------------------------------------------------------- */
static void replace_escapees(vbo_ctx* c) {            /* mirror_boundaries_mod.F90:336-407 */
  for (int ib = 0; ib < c->ncirc; ib++) {
    double R = c->crad[ib];
    for (int i = c->nbS; i < c->npfb; i++) {
      double rx = c->x[i] - c->ccen[2 * ib], ry = c->y[i] - c->ccen[2 * ib + 1];
      double rad = sqrt(rx * rx + ry * ry);
      if (rad < fabs(R) && R > 0.0) {
        c->x[i] = c->x[i] + (rx / rad) * 2.0 * (R - rad);
        c->y[i] = c->y[i] + (ry / rad) * 2.0 * (R - rad);
      }
      if (rad > fabs(R) && R < 0.0) {
        c->x[i] = c->x[i] - (rx / rad) * 2.0 * (rad + R);
        c->y[i] = c->y[i] - (ry / rad) * 2.0 * (rad + R);
      }
    }
  }
  for (int ib = 0; ib < c->nbp; ib++) {
    if (c->btype[ib] == 0) continue;
    if (c->btype[ib] != 1 && c->btype[ib] != 2) continue;
    double nb[2]; bound_norm(c, ib, nb);
    int ibp = c->bpp[ib] - 1;
    for (int i = c->nbS; i < c->npfb; i++) {
      if (c->inbin[i]) continue;
      double pd = bound_dist(c, c->x[i], c->y[i], ib);
      if (pd > 0.0) continue;
      if (fabs(pd) > c->sup + 1.0e-10) continue;
      double tmp = pos_along(c, c->x[i], c->y[i], ib, pd, nb);
      if (c->btype[ib] == 1) {
        if (tmp < 0.0 || tmp > 1.0) continue;
        c->x[i] = c->x[i] - 2.0 * pd * nb[0];
        c->y[i] = c->y[i] - 2.0 * pd * nb[1];
      } else {
        double np_[2]; bound_norm(c, ibp, np_);
        const double* e = c->bedge + 2 * ibp;
        double tx = -1.0 * tmp * e[0] - pd * np_[0];
        double ty = -1.0 * tmp * e[1] - pd * np_[1];
        c->x[i] = c->bnode[2 * ibp] + e[0] + tx;
        c->y[i] = c->bnode[2 * ibp + 1] + e[1] + ty;
      }
      c->x0[i] = c->x[i] - c->u[i] * c->dt;
      c->y0[i] = c->y[i] - c->v[i] * c->dt;
      c->u0[i] = c->u[i]; c->v0[i] = c->v[i];
    }
  }
}

static int new_mirror(vbo_ctx* c) {
  if (c->nmirror >= c->nmirror_esti || c->npfb + c->nmirror >= c->cap) return -1;
  return c->npfb + c->nmirror++;
}

static int phase_mirrors(vbo_ctx* c) {
  replace_escapees(c);
  const int nbp = c->nbp, npfb = c->npfb;
  /* create_boundary_boxes, mirror_boundaries_mod.F90:409-445 */
  plist* pl = (plist*)xcalloc(nbp, sizeof(plist));
  int** cl = (int**)xcalloc(nbp, sizeof(int*));
  int* ncl = (int*)xcalloc(nbp, 4);
  for (int ib = 0; ib < nbp; ib++) {
    pl[ib].i = (int*)xcalloc(npfb, 4); pl[ib].t = (double*)xcalloc(npfb, 8); pl[ib].d = (double*)xcalloc(npfb, 8);
    cl[ib] = (int*)xcalloc(npfb, 4);
    if (c->btype[ib] == 0) continue;
    double nb[2]; bound_norm(c, ib, nb);
    for (int i = 0; i < npfb; i++) {
      double pd = bound_dist(c, c->x[i], c->y[i], ib);
      if (fabs(pd) > c->sup + 1.0e-10) continue;
      double tmp = pos_along(c, c->x[i], c->y[i], ib, pd, nb);
      if (tmp < 0.0 || tmp > 1.0) continue;
      int k = pl[ib].n++;
      pl[ib].i[k] = i; pl[ib].t[k] = tmp; pl[ib].d[k] = pd;
    }
    for (int i = c->nbS; i < npfb; i++) {
      double ax = c->x[i] - c->bnode[2 * ib], ay = c->y[i] - c->bnode[2 * ib + 1];
      double pd = sqrt(ax * ax + ay * ay);
      if (pd > c->sup + 1.0e-10) continue;
      cl[ib][ncl[ib]++] = i;
    }
  }
  c->nmirror_esti = (npfb <= 500) ? 10 * npfb : 2 * npfb;
  c->nmirror = 0;
  int err = 0;
  const double bfx = c->bf[0], bfy = c->bf[1];
#define NEWM(k) do { k = new_mirror(c); if (k < 0) { err = 1; goto done; } } while (0)
  /* STEP 5a: circles, mirror_boundaries_mod.F90:61-96 */
  if (c->ncirc) {
    for (int i = c->nbS; i < npfb; i++) {
      for (int ib = 0; ib < c->ncirc; ib++) {
        double R = c->crad[ib];
        double cdir = R / fabs(R);
        double rx = c->x[i] - c->ccen[2 * ib], ry = c->y[i] - c->ccen[2 * ib + 1];
        double rad = sqrt(rx * rx + ry * ry);
        int k;
        if (cdir > 0.0) {
          if (rad <= R + c->sup) {
            rx = rx / rad; ry = ry / rad;
            NEWM(k);
            c->irel[k] = i; c->vrel[k] = 999 + (ib + 1);
            c->x[k] = c->x[i] - rx * 2.0 * (rad - R);
            c->y[k] = c->y[i] - ry * 2.0 * (rad - R);
            double tx = ry, ty = -rx;
            c->u[k] = 2.0 * c->cvel[2 * ib] + 2.0 * tx * R * c->comg[ib] - 1.0 * c->u[i];
            c->v[k] = 2.0 * c->cvel[2 * ib + 1] + 2.0 * ty * R * c->comg[ib] - 1.0 * c->v[i];
            c->th[k] = 2.0 * c->cth[ib] - 1.0 * c->th[i];
            c->dP[k] = bfx * (c->x[k] - c->x[i]) + bfy * (c->y[k] - c->y[i]);
          }
        } else if (cdir < 0.0) {
          if (rad >= fabs(R) - c->sup) {
            rx = rx / rad; ry = ry / rad;
            NEWM(k);
            c->irel[k] = i; c->vrel[k] = 999 + (ib + 1);
            c->x[k] = c->x[i] + rx * 2.0 * (fabs(R) - rad);
            c->y[k] = c->y[i] + ry * 2.0 * (fabs(R) - rad);
            double tx = ry, ty = -rx;
            c->u[k] = 2.0 * c->cvel[2 * ib] + 2.0 * tx * fabs(R) * c->comg[ib] - 1.0 * c->u[i];
            c->v[k] = 2.0 * c->cvel[2 * ib + 1] + 2.0 * ty * fabs(R) * c->comg[ib] - 1.0 * c->v[i];
            c->th[k] = 2.0 * c->cth[ib] - 1.0 * c->th[i];
            c->dP[k] = bfx * (c->x[k] - c->x[i]) + bfy * (c->y[k] - c->y[i]);
          }
        }
      }
    }
  }
  /* STEP 5: walls and periodic patches, mirror_boundaries_mod.F90:99-188 */
  for (int ib = 0; ib < nbp; ib++) {
    int ibp1 = (ib + 1) % nbp;
    if (c->btype[ib] == 0) continue;
    if (c->btype[ib] == 1) {
      double nb[2]; bound_norm(c, ib, nb);
      for (int jj = 0; jj < pl[ib].n; jj++) {
        int i = pl[ib].i[jj];
        double bp1 = pl[ib].t[jj], bp2 = pl[ib].d[jj];
        if (bp2 <= 1.0e-10) continue;
        if (c->bang[ib] < 0.0) {
          double ta = 0.5 * (PI + c->bang[ib]);
          const double* e = c->bedge + 2 * ib;
          double mbe = sqrt(e[0] * e[0] + e[1] * e[1]);
          if (bp2 + 0.1 * c->dx >= bp1 * tan(ta) * mbe) continue;
        }
        if (c->bang[ibp1] < 0.0) {
          double ta = 0.5 * (PI + c->bang[ibp1]);
          const double* e = c->bedge + 2 * ib;
          double mbe = sqrt(e[0] * e[0] + e[1] * e[1]);
          if (bp2 >= (1.0 - bp1) * tan(ta) * mbe) continue;
        }
        int k; NEWM(k);
        c->irel[k] = i; c->vrel[k] = ib + 1;
        c->x[k] = c->x[i] - 2.0 * bp2 * nb[0];
        c->y[k] = c->y[i] - 2.0 * bp2 * nb[1];
        double tx = nb[1], ty = -nb[0];
        c->u[k] = 2.0 * c->bvel[ib] * tx - c->u[i];
        c->v[k] = 2.0 * c->bvel[ib] * ty - c->v[i];
        c->th[k] = 2.0 * c->bth[ib] - c->th[i];
        c->dP[k] = bfx * (c->x[k] - c->x[i]) + bfy * (c->y[k] - c->y[i]);
      }
    }
    if (c->btype[ib] == 2) {
      int ibp = c->bpp[ib] - 1;
      double nb[2]; bound_norm(c, ib, nb);
      const double* e = c->bedge + 2 * ib;
      for (int jj = 0; jj < pl[ibp].n; jj++) {
        int i = pl[ibp].i[jj];
        double bp1 = pl[ibp].t[jj], bp2 = pl[ibp].d[jj];
        if (bp2 <= 0.0) continue;
        int k; NEWM(k);
        c->irel[k] = i; c->vrel[k] = ib + 1;
        double tx = -1.0 * bp1 * e[0] - bp2 * nb[0];
        double ty = -1.0 * bp1 * e[1] - bp2 * nb[1];
        c->x[k] = c->bnode[2 * ib] + e[0] + tx;
        c->y[k] = c->bnode[2 * ib + 1] + e[1] + ty;
        c->u[k] = c->u[i]; c->v[k] = c->v[i];
        c->th[k] = c->th[i];
        c->dP[k] = 0.0;
      }
    }
  }
  /* mirrors of moving / heated solid particles, mirror_boundaries_mod.F90:190-206 */
  for (int j = npfb; j < npfb + c->nmirror; j++) {
    int i = c->irel[j];
    double tmp = c->u[i] * c->u[i] + c->v[i] * c->v[i];
    if (i < c->nbS && tmp != 0.0) { c->u[j] = c->u[i]; c->v[j] = c->v[i]; }
  }
  for (int j = npfb; j < npfb + c->nmirror; j++) {
    int i = c->irel[j];
    double tmp = c->th[i] * c->th[i];
    if (i < c->nbS && tmp != 0.0) c->th[j] = c->th[i];
  }
  /* STEP 6: corners, mirror_boundaries_mod.F90:209-324 */
  for (int ib = 0; ib < nbp; ib++) {
    int ibm1 = (ib + nbp - 1) % nbp;
    if (c->btype[ib] == 0 || c->btype[ibm1] == 0) continue;
    const double* em = c->bedge + 2 * ibm1;
    const double* e = c->bedge + 2 * ib;
    double tmp = em[0] * e[1] - em[1] * e[0];
    if (fabs(tmp) < 1.0e-10) continue;
    if (c->btype[ib] == 1 && c->btype[ibm1] == 1) {
      if (c->bang[ib] <= 0.0) continue;
      for (int jj = 0; jj < ncl[ib]; jj++) {
        int i = cl[ib][jj];
        double tbx = c->x[i] - c->bnode[2 * ib], tby = c->y[i] - c->bnode[2 * ib + 1];
        int k; NEWM(k);
        c->irel[k] = i; c->vrel[k] = -(ib + 1);
        c->x[k] = c->x[i] - 2.0 * tbx;
        c->y[k] = c->y[i] - 2.0 * tby;
        double nb[2]; bound_norm(c, ib, nb);
        double tx = nb[1], ty = -nb[0];
        c->u[k] = 2.0 * tmp * tx - c->u[i];      /* tmp: the cross product above (reference uses it here) */
        c->v[k] = 2.0 * tmp * ty - c->v[i];
        bound_norm(c, ibm1, nb);
        tx = nb[1]; ty = -nb[0];
        c->u[k] = c->u[k] + 2.0 * tmp * tx;
        c->v[k] = c->v[k] + 2.0 * tmp * ty;
        c->th[k] = c->th[i] + 2.0 * c->bth[ib];
        c->dP[k] = bfx * (c->x[k] - c->x[i]) + bfy * (c->y[k] - c->y[i]);
      }
    }
    if (c->btype[ib] == 2 && c->btype[ibm1] == 2) {
      int ibp = c->bpp[ib] - 1;
      for (int jj = 0; jj < ncl[ibp]; jj++) {
        int i = cl[ibp][jj];
        int k; NEWM(k);
        c->irel[k] = i; c->vrel[k] = -(ib + 1);
        double tx = c->bnode[2 * ib] - c->bnode[2 * ibp], ty = c->bnode[2 * ib + 1] - c->bnode[2 * ibp + 1];
        c->x[k] = c->x[i] + tx; c->y[k] = c->y[i] + ty;
        c->u[k] = c->u[i]; c->v[k] = c->v[i]; c->th[k] = c->th[i];
        c->dP[k] = 0.0;
      }
    }
    if (c->btype[ib] != c->btype[ibm1]) {
      if (c->btype[ib] == 1 || c->btype[ibm1] == 1) {
        int ibb = (c->btype[ibm1] == 1) ? ibm1 : ib;
        int imp_temp = c->nmirror;
        for (int j = npfb; j < npfb + imp_temp; j++) {
          double pd = bound_dist(c, c->x[j], c->y[j], ibb);
          if (pd > c->sup + 1.0e-10) continue;
          if (pd < 1.0e-10) continue;
          double nb[2]; bound_norm(c, ibb, nb);
          double t = pos_along(c, c->x[j], c->y[j], ibb, pd, nb);
          if ((t < 0.0 && ibb == ib) || (t > 1.0 && ibb == ibm1)) {
            int k; NEWM(k);
            c->irel[k] = c->irel[j]; c->vrel[k] = -(ib + 1);
            c->x[k] = c->x[j] - 2.0 * pd * nb[0];
            c->y[k] = c->y[j] - 2.0 * pd * nb[1];
            double tx = nb[1], ty = -nb[0];
            c->u[k] = 2.0 * c->bvel[ibb] * tx - c->u[j];
            c->v[k] = 2.0 * c->bvel[ibb] * ty - c->v[j];
            c->th[k] = 2.0 * c->bth[ibb] - c->th[j];
            c->dP[k] = c->dP[j];
            c->dP[k] = c->dP[k] + (bfx * (c->x[k] - c->x[j]) + bfy * (c->y[k] - c->y[j]));
          }
        }
      }
    }
  }
#undef NEWM
done:
  for (int ib = 0; ib < nbp; ib++) { free(pl[ib].i); free(pl[ib].t); free(pl[ib].d); free(cl[ib]); }
  free(pl); free(cl); free(ncl);
  c->np = npfb + c->nmirror;
  return err ? VBO_ERR_CAPACITY : VBO_OK;
}

/* ----------------------------------------------------------------------------------------------
 * Phase 4: particle_allocation -> divide + neighboring_list_parallel,
 * neighbour_finding_mod.F90:15-127, 261-365.  Cells are 1-based as in the reference.
 * -------------------------------------------------------------------------------------------- */
static int phase_neighbours(vbo_ctx* c) {
  const int np = c->np, npfb = c->npfb, ncx = c->ncx, ncy = c->ncy, nct = c->nct;
  int* nc = (int*)xcalloc(nct + ncx + ncy + 3, 4);
  c->ist = (int*)xcalloc(nct + ncx + ncy + 4, 4);
  c->ip = (int*)xcalloc(np + 1, 4);
  c->cellpart = (int*)xcalloc(np + 1, 4);
  int ncmax = 0;
  for (int k = 0; k < np; k++) {                           /* neighbour_finding_mod.F90:72-108 */
    if (c->inbin[k]) continue;
    double ddx = c->x[k] - c->xmint, ddy = c->y[k] - c->ymint;
    int icell = (int)(ddx * c->uno) + 1;
    int jcell = (int)(ddy * c->uno) + 1;
    int ii = icell + (jcell - 1) * ncx;
    if (ii > nct || ii < 1) { free(nc); return VBO_ERR_OUT_OF_GRID; }
    nc[ii]++;
    if (nc[ii] > ncmax) ncmax = nc[ii];
    if (ncmax > NPLINK) { free(nc); return VBO_ERR_CELL_OVERFLOW; }
    c->cellpart[k] = ii;
  }
  c->ist[1] = 1;                                            /* :112-116 */
  for (int ii = 1; ii <= nct; ii++) { c->ist[ii + 1] = c->ist[ii] + nc[ii]; nc[ii] = 0; }
  for (int k = 0; k < np; k++) {                            /* :121-125 (ip holds 0-based k) */
    if (c->inbin[k]) continue;
    int cp = c->cellpart[k];
    c->ip[c->ist[cp] + nc[cp]] = k;
    nc[cp]++;
  }
  c->cnt = (int*)xcalloc(npfb, 4);
  c->link = (int*)xcalloc((size_t)npfb * NPLINK, 4);
  int err = 0;
  for (int i = 0; i < npfb; i++) {                          /* :271-337 */
    if (c->inbin[i]) continue;
    int ic = c->cellpart[i];
    int icx = ic % ncx, icy = ic / ncx + 1;
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
    double xi = c->x[i], yi = c->y[i];
    for (int q = 0; q < nce; q++) {                         /* neighbour_particle_cell :341-365 */
      int jc = cells[q];
      if (nc[jc] == 0) continue;
      for (int jj = c->ist[jc]; jj <= c->ist[jc + 1] - 1; jj++) {
        int j = c->ip[jj];
        double rx = xi - c->x[j], ry = yi - c->y[j];
        double rr2 = rx * rx + ry * ry;
        if (rr2 > 0.0 && rr2 <= c->sup2) {
          if (c->cnt[i] >= NPLINK) { err = 1; continue; }
          c->link[(size_t)c->cnt[i] * npfb + i] = j;
          c->cnt[i]++;
        }
      }
    }
  }
  free(nc);
  return err ? VBO_ERR_CAPACITY : VBO_OK;
}

/* Phase 5: kernel_calculation.F90:2-75 (KGC on) */
static void phase_weights(vbo_ctx* c) {
  const int npfb = c->npfb;
  size_t nw = (size_t)npfb * NPLINK;
  c->wg1 = (double*)xcalloc(nw, 8); c->wg2 = (double*)xcalloc(nw, 8); c->wl = (double*)xcalloc(nw, 8);
  c->k11 = (double*)xcalloc(npfb, 8); c->k12 = (double*)xcalloc(npfb, 8);
  c->k21 = (double*)xcalloc(npfb, 8); c->k22 = (double*)xcalloc(npfb, 8);
  int mx = 0;
  for (int i = 0; i < npfb; i++) if (c->cnt[i] > mx) mx = c->cnt[i];
  c->maxneigh = mx;
  const double eta_small = (double)0.01f * c->h * c->h;
  const double oneoverh = 1.0 / c->h, dv = c->dv;
  for (int i = 0; i < npfb; i++) {
    double a11 = 0.0, a12 = 0.0, a21 = 0.0, a22 = 0.0;
    for (int k = 0; k < c->cnt[i]; k++) {
      size_t s = (size_t)k * npfb + i;
      int j = c->link[s];
      double rx = c->x[i] - c->x[j], ry = c->y[i] - c->y[j];
      double rr2 = rx * rx + ry * ry;
      double rad = sqrt(rr2);
      double qq = rad * oneoverh;
      double f = fac(qq, c->ad7h);
      double den = rad > c->eta2 ? rad : c->eta2;
      double gx = rx * f / den, gy = ry * f / den;
      double w1 = gx * dv, w2 = gy * dv;
      c->wg1[s] = w1; c->wg2[s] = w2;
      c->wl[s] = 2.0 * (rx * w1 + ry * w2) / (rr2 + eta_small);
      a11 = a11 - dv * gx * rx; a12 = a12 - dv * gy * rx;
      a21 = a21 - dv * gx * ry; a22 = a22 - dv * gy * ry;
    }
    double t = 1.0 / ((a11 * a22 - a12 * a21) + 1.0e-15);
    c->k11[i] = t * a22; c->k12[i] = -t * a12; c->k21[i] = -t * a21; c->k22[i] = t * a11;
  }
}

/* Phase 6: locate_free_surf_part (freesurf_mod.F90:17-68).  Closed domain: n_surf = 0.  Open
 * domain: calc_part_dist_vars (:70-103) gives the concentration, div r and the surface normal
 * -h sum w_G; a particle is on the surface when div r <= 1.4 + 0.5 max(conc - 1, -0.2) (:27-38);
 * mirrors take their parent's flag (:43-45); smooth_normals (:105-137) Shepard-filters the normals
 * over the real neighbours; ppe_smooth (:54-64) ramps the PPE rows of the particles just inside the
 * surface. */
static void phase_freesurf(vbo_ctx* c) {
  const int npfb = c->npfb;
  const int open = c->open_domain == 1;
  double* bdr = NULL;
  if (open) {
    c->snx = (double*)xcalloc(npfb, 8); c->sny = (double*)xcalloc(npfb, 8);
    c->psmooth = (double*)xcalloc(npfb, 8);
    bdr = (double*)xcalloc(npfb, 8);
    for (int i = 0; i < npfb; i++) {                             /* :85-97 */
      double ct = 0.0, dr = 0.0, sx = 0.0, sy = 0.0;
      for (int k = 0; k < c->cnt[i]; k++) {
        const size_t s = (size_t)k * npfb + i;
        const int j = c->link[s];
        const double rx = c->x[i] - c->x[j], ry = c->y[i] - c->y[j];
        const double rad = sqrt(rx * rx + ry * ry);
        const double qq = rad / c->h;
        ct = ct + wab(qq, c->ad7);
        dr = dr - (rx * c->wg1[s] + ry * c->wg2[s]);
        sx = sx - c->wg1[s]; sy = sy - c->wg2[s];
      }
      c->conc[i] = ct * c->dv; c->divr[i] = dr; c->snx[i] = sx; c->sny[i] = sy;
    }
    const double w0 = wab(0.0, c->ad7);                          /* :99-101 */
    for (int i = 0; i < npfb; i++) c->conc[i] = c->conc[i] + c->dv * w0;
    for (int i = 0; i < npfb; i++) { c->snx[i] = c->snx[i] * c->h; c->sny[i] = c->sny[i] * c->h; }
    for (int i = 0; i < npfb; i++) {                             /* :29-37 */
      bdr[i] = 1.4 + 0.5 * fmax(c->conc[i] - 1.0, -0.2);
      c->nsurf[i] = (c->divr[i] > bdr[i]) ? 0 : 1;
    }
  } else {
    for (int i = 0; i < npfb; i++) c->nsurf[i] = 0;
  }
  for (int j = npfb; j < c->np; j++) c->nsurf[j] = c->nsurf[c->irel[j]];
  if (!open) return;
  {                                                              /* smooth_normals, :114-135 */
    const double w0 = wab(0.0, c->ad7);
    double* tx = (double*)xcalloc(npfb, 8);
    double* ty = (double*)xcalloc(npfb, 8);
    for (int i = 0; i < npfb; i++) {
      double sx = 0.0, sy = 0.0;
      for (int k = 0; k < c->cnt[i]; k++) {
        const int j = c->link[(size_t)k * npfb + i];
        if (j >= npfb) continue;
        const double rx = c->x[i] - c->x[j], ry = c->y[i] - c->y[j];
        const double rad = sqrt(rx * rx + ry * ry);
        const double w = wab(rad / c->h, c->ad7);
        sx = sx + w * c->snx[j]; sy = sy + w * c->sny[j];
      }
      tx[i] = sx * c->dv + c->snx[i] * w0 * c->dv;
      ty[i] = sy * c->dv + c->sny[i] * w0 * c->dv;
    }
    memcpy(c->snx, tx, 8 * (size_t)npfb); memcpy(c->sny, ty, 8 * (size_t)npfb);
    free(tx); free(ty);
  }
  for (int i = 0; i < npfb; i++) {                               /* ppe_smooth, :55-64 */
    double ps = 1.0;
    if (!(c->divr[i] > 1.2 * bdr[i])) {
      if (c->nsurf[i] == 1) ps = 0.0;
      else ps = 0.5 * (1.0 - cos((c->divr[i] - bdr[i]) * PI / (0.2 * bdr[i])));
    }
    c->psmooth[i] = ps;
  }
  free(bdr);
}

/* Phase 7: new_fick_shift + shifting_forces (new_fick_shift.F90:15-92, 117-210; akinci off) and
 * ushift as div_free.F90:65-78.  Shifting coefficient D: i_shift_coeff 0, 0.25 h^2 (:44);
 * i_shift_coeff 1, 4 h dt (|u| + 0.2 exp(-|u|/(0.2 maxU)) maxU) (:46-50), whose REAL(4) literals 0.2
 * are the float 0.2 widened.  maxU (:31-37) assigns rather than maximises inside an OpenMP max
 * reduction, so it is max(1e-10, the speed of each thread's last particle): with one thread
 * max(1e-10, |u(npfb)|), which is what is computed here.  Open-domain surface particles shift along
 * the surface only, dr = -D (t . grad C) t with t the rotated unit smoothed normal
 * n = sn / max(|sn|, 1e-8) (REAL(4) 1e-8 widened) (:60-73). */
static void phase_shift(vbo_ctx* c) {
  const int npfb = c->npfb;
  const double nt = 4.0, Rt = 0.25;
  const double wab1 = wab(c->dx / c->h, c->ad7);
  const double f02 = (double)0.2f, small_n = (double)1.0e-8f;
  c->usx = (double*)xcalloc(npfb, 8); c->usy = (double*)xcalloc(npfb, 8);
  double maxU = 1.0e-10;
  if (npfb > 0) {
    const int l = npfb - 1;
    const double t = sqrt(c->u[l] * c->u[l] + c->v[l] * c->v[l]);
    if (t > maxU) maxU = t;
  }
  const int open = c->open_domain == 1;
  for (int i = 0; i < npfb; i++) {
    double sfx = 0.0, sfy = 0.0, stx = 0.0, sty = 0.0;
    for (int k = 0; k < c->cnt[i]; k++) {
      size_t s = (size_t)k * npfb + i;
      int j = c->link[s];
      double rx = c->x[i] - c->x[j], ry = c->y[i] - c->y[j];
      double rad = sqrt(rx * rx + ry * ry);
      double qq = rad / c->h;
      double fab = Rt * pow(wab(qq, c->ad7) / wab1, nt);
      sfx = sfx + c->wg1[s]; sfy = sfy + c->wg2[s];
      stx = stx + c->wg1[s] * fab; sty = sty + c->wg2[s] * fab;
    }
    double gcx = sfx + stx, gcy = sfy + sty;
    double drx = 0.0, dry = 0.0;
    if (i >= c->nbS) {
      double D;
      if (c->shift_coeff == 1) {
        const double absU = sqrt(c->u[i] * c->u[i] + c->v[i] * c->v[i]);
        const double temp = f02 * exp(-absU / (f02 * maxU));
        D = 4.0 * c->h * c->dt * (absU + temp * maxU);
      } else {
        D = 0.25 * c->h * c->h;
      }
      if (open && c->nsurf[i] == 1) {
        const double tn = sqrt(c->snx[i] * c->snx[i] + c->sny[i] * c->sny[i]);
        const double den = tn > small_n ? tn : small_n;
        const double nx = c->snx[i] / den, ny = c->sny[i] / den;
        const double tx = -ny, ty = nx;
        const double dCds = tx * gcx + ty * gcy;
        drx = -D * dCds * tx; dry = -D * dCds * ty;
      } else {
        drx = -D * gcx; dry = -D * gcy;
      }
      double t = sqrt(drx * drx + dry * dry);
      if (t > 0.1 * c->h) { drx = 0.1 * c->h * drx / t; dry = 0.1 * c->h * dry / t; }
    }
    c->usx[i] = drx / c->dt; c->usy[i] = dry / c->dt;
  }
  for (int i = 0; i < c->nbS; i++) { c->usx[i] = -c->u[i]; c->usy[i] = -c->v[i]; }
}

/* ----------------------------------------------------------------------------------------------
 * Phase 8: prediction_step (predictor_mod.F90:27-139) with constitutive_eqn_evolve (:141-251)
 * -------------------------------------------------------------------------------------------- */
/* calc_eigens, predictor_mod.F90:253-290.  A and R are (a11,a12,a21,a22). */
static void calc_eigens(const double* A, double* L, double* R) {
  double tr = A[0] + A[3];
  double det = A[0] * A[3] - A[2] * A[1];
  double tmp = 0.25 * tr * tr - det;
  if (tmp <= EPS_DBL) tmp = 0.0;
  L[0] = 0.5 * tr + sqrt(tmp);
  L[1] = 0.5 * tr - sqrt(tmp);
  if (A[2] != 0.0) {
    double r11 = L[0] - A[3], r21 = A[2];
    double nrm = sqrt(r11 * r11 + r21 * r21);
    r11 = r11 / nrm; r21 = r21 / nrm;
    R[0] = r11; R[2] = r21; R[1] = -r21; R[3] = r11;
  } else {
    L[0] = A[0]; L[1] = A[3];
    R[0] = 1.0; R[3] = 1.0; R[1] = 0.0; R[2] = 0.0;
  }
}

void vbo_calc_eigens(int n, const double* a, double* lam, double* r) {
  for (int i = 0; i < n; i++) calc_eigens(a + 4 * i, lam + 2 * i, r + 4 * i);
}

/* 2x2 products in the order matmul sums them: C(a,b) = A(a,1)B(1,b) + A(a,2)B(2,b) */
static inline void mm(const double* A, const double* B, double* Cm) {
  double c11 = A[0] * B[0] + A[1] * B[2], c12 = A[0] * B[1] + A[1] * B[3];
  double c21 = A[2] * B[0] + A[3] * B[2], c22 = A[2] * B[1] + A[3] * B[3];
  Cm[0] = c11; Cm[1] = c12; Cm[2] = c21; Cm[3] = c22;
}
static inline void tr2(const double* A, double* T) { T[0] = A[0]; T[1] = A[2]; T[2] = A[1]; T[3] = A[3]; }
/* R diag(d1,d2) R^T as matmul(R, matmul(D, transpose(R))) with D diagonal (zero products drop out) */
static inline void rdrt(const double* R, double d1, double d2, double* out) {
  double m11 = d1 * R[0], m12 = d1 * R[2], m21 = d2 * R[1], m22 = d2 * R[3];
  out[0] = R[0] * m11 + R[1] * m21; out[1] = R[0] * m12 + R[1] * m22;
  out[2] = R[2] * m11 + R[3] * m21; out[3] = R[2] * m12 + R[3] * m22;
}

/* The velocity-gradient "kick" of predictor_mod.F90:369-371.  Its guard
 *   if(itime.eq.2.and.lagrangian.eqv..false.)
 * parses as ((itime==2 .and. lagrangian) .eqv. .false.), so for a Lagrangian run it fires on every
 * step except the second.  random_seed() (:357) resets the runtime generator to its default state
 * each step, so particle i (1-based, single thread) always draws the i-th number of the same
 * stream: Park-Miller "minstd" x <- 48271 x mod (2^31-1) from x = 1, two draws per real(8):
 * f = (a << 30 | ((b - 1) & (2^30-1))) >> 7, r = f * 2^-54.  B gets 1e13*epsilon(1d0)*(r - 0.5)
 * added to all four entries. */
static double* make_kick(int n) {
  double* k = (double*)xcalloc(n, 8);
  uint64_t x = 1;
  const uint64_t m = 2147483647ull;
  for (int i = 0; i < n; i++) {
    x = (x * 48271ull) % m; uint64_t a = x;
    x = (x * 48271ull) % m; uint64_t b = (x - 1) & ((1ull << 30) - 1);
    uint64_t f = ((a << 30) | b) >> 7;
    double r = ldexp((double)f, -54);
    k[i] = (1.0e13 * EPS_DBL) * (r - 0.5);
  }
  return k;
}

/* One particle's log-conformation update, predictor_mod.F90:165-245 (lagrangian, no advection of
 * Psi); kick is the B perturbation above (0 on the step where the guard is false). */
static void logconf_one(const double* tau, double trc_prev, const double* G, double dt, double wi,
                        double eps_ptt, double kick, int use_kick, double* tau_out, double* trc_out) {
  double A[4] = {tau[0] / 1.0 + 1.0, tau[1] / 1.0 + 0.0, tau[2] / 1.0 + 0.0, tau[3] / 1.0 + 1.0};
  double L[2], R[4];
  calc_eigens(A, L, R);
  double Psi[4], Psi0[4];
  rdrt(R, log(L[0]), log(L[1]), Psi);
  memcpy(Psi0, Psi, sizeof Psi);
  double RT[4], GR[4], M[4];
  tr2(R, RT);
  mm(G, R, GR);
  mm(RT, GR, M);
  double om = (L[1] * M[1] + L[0] * M[2]) / (L[1] - L[0] + 1.0e-10);
  double B[4];
  rdrt(R, M[0], M[3], B);
  if (use_kick) for (int q = 0; q < 4; q++) B[q] = B[q] + kick;
  /* Omega = R [[0,om],[-om,0]] R^T */
  double w11 = om * R[1], w12 = om * R[3], w21 = -om * R[0], w22 = -om * R[2];
  double Om[4] = {R[0] * w11 + R[1] * w21, R[0] * w12 + R[1] * w22,
                  R[2] * w11 + R[3] * w21, R[2] * w12 + R[3] * w22};
  double OP[4], PO[4];
  mm(Om, Psi0, OP);
  mm(Psi0, Om, PO);
  for (int q = 0; q < 4; q++) Psi[q] = Psi[q] + dt * ((2.0 * B[q] + OP[q]) - PO[q]);
  calc_eigens(Psi, L, R);
  L[0] = exp(L[0]); L[1] = exp(L[1]);
  double Aa[4];
  rdrt(R, L[0], L[1], Aa);
  /* f_relaxation_integrate, predictor_mod.F90:318-331 */
  double aR = 1.0 + eps_ptt * (trc_prev - 2.0);
  double t1 = exp(-aR * 1.0 * dt / wi), t2 = (1.0 - t1) / 1.0;
  Aa[0] = Aa[0] * t1 + t2; Aa[1] = Aa[1] * t1; Aa[2] = Aa[2] * t1; Aa[3] = Aa[3] * t1 + t2;
  double trc = Aa[0] + Aa[3];
  *trc_out = trc;
  tau_out[0] = 1.0 * (1.0 * Aa[0] - 1.0); tau_out[1] = 1.0 * (1.0 * Aa[1] - 0.0);
  tau_out[2] = 1.0 * (1.0 * Aa[2] - 0.0); tau_out[3] = 1.0 * (1.0 * Aa[3] - 1.0);
}

void vbo_logconf(int n, const double* tau, const double* trc, const double* gv, double dt,
                 double wi, double eps_ptt, int use_kick, double* tau_out, double* trc_out) {
  double* kick = make_kick(n);
  for (int i = 0; i < n; i++)
    logconf_one(tau + 4 * i, trc[i], gv + 4 * i, dt, wi, eps_ptt, kick[i], use_kick, tau_out + 4 * i, trc_out + i);
  free(kick);
}

/* Corrected gradient of phi at i, gradient_calculation.F90:61-71 */
static void grad_one(const vbo_ctx* c, const double* phi, int i, double* g) {
  double gx = 0.0, gy = 0.0;
  for (int k = 0; k < c->cnt[i]; k++) {
    size_t s = (size_t)k * c->npfb + i;
    double t = phi[c->link[s]] - phi[i];
    gx = gx + t * c->wg1[s]; gy = gy + t * c->wg2[s];
  }
  g[0] = c->k11[i] * gx + c->k12[i] * gy;
  g[1] = c->k21[i] * gx + c->k22[i] * gy;
}

/* mirror_polymeric_stress, mirror_boundaries_mod.F90:586-653: tau_j = Q F Q^T tau_i Q Q^T, F flips row 1 */
static void reflect_tau(const double* ti, double n1, double n2, double* tj) {
  double nt1 = -n2, nt2 = n1;
  double Q[4] = {1.0 * n1 + 0.0 * n2, 1.0 * nt1 + 0.0 * nt2, 0.0 * n1 + 1.0 * n2, 0.0 * nt1 + 1.0 * nt2};
  double QT[4], a[4], b[4];
  tr2(Q, QT);
  mm(ti, Q, a);
  mm(QT, a, b);
  b[0] = -b[0]; b[1] = -b[1];
  mm(b, QT, a);
  mm(Q, a, tj);
}

static void mirror_stress(vbo_ctx* c) {
  const int nbp = c->nbp;
  for (int j = c->npfb; j < c->np; j++) {
    int i = c->irel[j], vr = c->vrel[j];
    double ti[4] = {c->t11[i], c->t12[i], c->t21[i], c->t22[i]}, tj[4];
    int set = 0;
    if (vr > 999) {
      double rx = c->x[i] - c->x[j], ry = c->y[i] - c->y[j];
      double d = sqrt(rx * rx + ry * ry);
      reflect_tau(ti, rx / d, ry / d, tj); set = 1;
    }
    if (vr > 0 && vr < 999) {
      int t = c->btype[vr - 1];
      if (t == 1) {
        double rx = c->x[i] - c->x[j], ry = c->y[i] - c->y[j];
        double d = sqrt(rx * rx + ry * ry);
        reflect_tau(ti, rx / d, ry / d, tj); set = 1;
      } else if (t == 2 || t == 4) {
        memcpy(tj, ti, sizeof tj); set = 1;
      } else if (t == 3) {
        tj[0] = tj[1] = tj[2] = tj[3] = 0.0; set = 1;
      }
    }
    if (vr < 0 && vr > -c->npfb) {
      int ib = -vr - 1, ibm1 = (ib + nbp - 1) % nbp;
      int a = c->btype[ib], b = c->btype[ibm1];
      if (a == 1 && b == 1) { for (int q = 0; q < 4; q++) tj[q] = -ti[q]; set = 1; }
      else if (a == 2 && b == 2) { memcpy(tj, ti, sizeof tj); set = 1; }
      else if ((a == 1 && b == 2) || (a == 2 && b == 1)) {
        double nb[2];
        bound_norm(c, a == 1 ? ib : ibm1, nb);
        reflect_tau(ti, nb[0], nb[1], tj); set = 1;
      }
    }
    if (set) { c->t11[j] = tj[0]; c->t12[j] = tj[1]; c->t21[j] = tj[2]; c->t22[j] = tj[3]; }
  }
}

static void phase_predict(vbo_ctx* c) {
  const int npfb = c->npfb;
  c->gvel = (double*)xcalloc((size_t)npfb * 4, 8);
  for (int i = 0; i < npfb; i++) {                   /* predictor_mod.F90:35-36 */
    grad_one(c, c->u, i, c->gvel + 4 * i);
    grad_one(c, c->v, i, c->gvel + 4 * i + 2);
  }
  const int newtonian = c->const_mod == 1;
  /* constitutive_eqn_evolve (viscoelastic build only, :43-46): all of 1..npfb including walls */
  const double wi = c->El * sqrt(c->Ra / c->Pr);
  for (int i = 0; i < npfb && !newtonian; i++) {
    double ti[4] = {c->t11[i], c->t12[i], c->t21[i], c->t22[i]}, to[4], trc;
    logconf_one(ti, c->trc[i], c->gvel + 4 * i, c->dt, wi, c->eps_ptt, c->kick[i], c->itime != 2, to, &trc);
    c->trc[i] = trc;
    c->t11[i] = to[0]; c->t12[i] = to[1]; c->t21[i] = to[2]; c->t22[i] = to[3];
  }
  if (!newtonian) mirror_stress(c);
  /* RHS: viscoelastic :79-100, Newtonian :58-77 (solvent and thermal terms only) */
  const double cs = (c->beta) * sqrt(c->Pr / c->Ra);
  const double cp = (1.0 - c->beta) * c->Pr / (c->Ra * c->El);
  const double ct = 1.0 / sqrt(c->Ra * c->Pr);
  double* rx = (double*)xcalloc(npfb, 8);
  double* ry = (double*)xcalloc(npfb, 8);
  double* rt = (double*)xcalloc(npfb, 8);
  for (int i = 0; i < npfb; i++) {
    double ax = 0.0, ay = 0.0, at = 0.0;
    for (int k = 0; k < c->cnt[i]; k++) {
      size_t s = (size_t)k * npfb + i;
      int j = c->link[s];
      double uij = c->u[i] - c->u[j], vij = c->v[i] - c->v[j];
      if (newtonian) {
        ax = ax + cs * uij * c->wl[s];
        ay = ay + cs * vij * c->wl[s];
        at = at + ct * (c->th[i] - c->th[j]) * c->wl[s];
        continue;
      }
      double T11 = c->t11[j] - c->t11[i], T12 = c->t12[j] - c->t12[i];
      double T21 = c->t21[j] - c->t21[i], T22 = c->t22[j] - c->t22[i];
      double kw1 = c->k11[i] * c->wg1[s] + c->k12[i] * c->wg2[s];
      double kw2 = c->k21[i] * c->wg1[s] + c->k22[i] * c->wg2[s];
      double tw1 = T11 * kw1 + T12 * kw2, tw2 = T21 * kw1 + T22 * kw2;
      ax = ax + cp * tw1 + cs * uij * c->wl[s];
      ay = ay + cp * tw2 + cs * vij * c->wl[s];
      double tij = c->th[i] - c->th[j];
      at = at + ct * tij * c->wl[s];
    }
    rx[i] = ax; ry[i] = ay; rt[i] = at;
  }
  for (int i = 0; i < npfb; i++) ry[i] = ry[i] + c->th[i];      /* :111-117 buoyancy */
  for (int i = c->nbS; i < npfb; i++) {                         /* :120-133 */
    c->u[i] = c->u[i] + rx[i] * c->dt;
    c->v[i] = c->v[i] + ry[i] * c->dt;
  }
  for (int i = c->nbS; i < npfb; i++) c->th[i] = c->th[i] + rt[i] * c->dt;
  free(rx); free(ry); free(rt);
}

/* Phase 9: mirror_velocities, mirror_boundaries_mod.F90:509-584 */
static void phase_mirror_vel(vbo_ctx* c) {
  const int nbp = c->nbp;
  for (int j = c->npfb; j < c->np; j++) {
    int i = c->irel[j], vr = c->vrel[j];
    if (vr > 999) {
      int ib = vr - 999 - 1;
      c->u[j] = -c->u[i]; c->v[j] = -c->v[i];
      double rx = 0.5 * (c->x[i] - c->x[j]), ry = 0.5 * (c->y[i] - c->y[j]);
      double t = sqrt(rx * rx + ry * ry);
      double tx = ry / t, ty = -rx / t;
      c->u[j] = 2.0 * c->cvel[2 * ib] + 2.0 * tx * c->crad[ib] * c->comg[ib] - 1.0 * c->u[i];
      c->v[j] = 2.0 * c->cvel[2 * ib + 1] + 2.0 * ty * c->crad[ib] * c->comg[ib] - 1.0 * c->v[i];
    }
    if (vr > 0 && vr < 999) {
      int t = c->btype[vr - 1];
      if (abs(t) == 1) {
        double rx = 0.5 * (c->x[i] - c->x[j]), ry = 0.5 * (c->y[i] - c->y[j]);
        double d = sqrt(rx * rx + ry * ry);
        double tx = ry / d, ty = -rx / d;
        c->u[j] = 2.0 * c->bvel[vr - 1] * tx - c->u[i];
        c->v[j] = 2.0 * c->bvel[vr - 1] * ty - c->v[i];
      } else if (t == 2) {
        c->u[j] = c->u[i]; c->v[j] = c->v[i];
      }
    }
    if (vr < 0 && vr > -c->npfb) {
      int ib = -vr - 1, ibm1 = (ib + nbp - 1) % nbp;
      int a = c->btype[ib], b = c->btype[ibm1];
      double nb[2];
      if (a == 1 && b == 1) {
        c->u[j] = c->u[i]; c->v[j] = c->v[i];
        if (c->bvel[ib] != 0.0) {
          bound_norm(c, ib, nb);
          double tmp = c->bvel[ib];
          c->u[j] = 2.0 * tmp * nb[1] - c->u[i]; c->v[j] = 2.0 * tmp * (-nb[0]) - c->v[i];
        }
        if (c->bvel[ibm1] != 0.0) {
          bound_norm(c, ibm1, nb);
          double tmp = c->bvel[ibm1];
          c->u[j] = c->u[j] + 2.0 * tmp * nb[1]; c->v[j] = c->v[j] + 2.0 * tmp * (-nb[0]);
        }
      } else if (a == 2 && b == 2) {
        c->u[j] = c->u[i]; c->v[j] = c->v[i];
      } else if ((a == 1 && b == 2) || (a == 2 && b == 1)) {
        double tmp, tx, ty;
        if (c->bvel[ib] != 0.0) { bound_norm(c, ib, nb); tx = nb[1]; ty = -nb[0]; tmp = c->bvel[ib]; }
        else if (c->bvel[ibm1] != 0.0) { bound_norm(c, ibm1, nb); tx = nb[1]; ty = -nb[0]; tmp = c->bvel[ibm1]; }
        else { tmp = 0.0; tx = 0.0; ty = 0.0; }
        c->u[j] = 2.0 * tmp * tx - c->u[i];
        c->v[j] = 2.0 * tmp * ty - c->v[i];
      }
    }
  }
  for (int j = c->npfb; j < c->np; j++) {
    int i = c->irel[j];
    double tmp = c->u[i] * c->u[i] + c->v[i] * c->v[i];
    if (i < c->nbS && tmp != 0.0) { c->u[j] = c->u[i]; c->v[j] = c->v[i]; }
  }
}

/* ----------------------------------------------------------------------------------------------
 * Linear solver, linear_solver.F90:6-59, 231-425.  NR row-indexed storage, 1-based ija.
 * -------------------------------------------------------------------------------------------- */
void vbo_spmv_nr(int n, const double* A, const int* ija, const double* x, double* b) {  /* :254-273 */
  for (int i = 0; i < n; i++) {
    double bb = A[i] * x[i];
    for (int k = ija[i]; k <= ija[i + 1] - 1; k++) bb = bb + A[k - 1] * x[ija[k - 1] - 1];
    b[i] = bb;
  }
}

static double dotp(int n, const double* a, const double* b) {
  double t = 0.0;
  for (int i = 0; i < n; i++) t = t + a[i] * b[i];
  return t;
}

static void remove_dc(int n, double* a) {                 /* :231-250 */
  double t = 0.0;
  for (int i = 0; i < n; i++) t = t + a[i];
  t = t / (double)n;
  for (int i = 0; i < n; i++) a[i] = a[i] - t;
}

/* The reference's two BiCGSTABs, linear_solver.F90:82-227 (open) and :275-425 (nonull).  They differ
 * in three places: SPARSE_PBICGSTAB preconditions with Jacobi, M = A(i) (:63-75), and removes no
 * mean; SPARSE_PBICGSTAB_nonull uses M = 1 (:318) and projects the constant out of p and s
 * (remove_dc, :351, :383); and the first-iteration norm thresholds are 10-6 -- an integer expression
 * equal to 4 -- in the former (:171, :207) and 1.0d-6 in the latter (:367, :404). */
static int bicgstab(int n, const double* A, const int* ija, double* x, const double* B, int itermax,
                    double tol, int* iters, double* resid, int open) {
  double *r = (double*)xcalloc(n, 8), *rb = (double*)xcalloc(n, 8), *v = (double*)xcalloc(n, 8);
  double *ax = (double*)xcalloc(n, 8), *p = (double*)xcalloc(n, 8), *pinv = (double*)xcalloc(n, 8);
  double *s = (double*)xcalloc(n, 8), *sinv = (double*)xcalloc(n, 8), *t = (double*)xcalloc(n, 8);
  double* M = (double*)xcalloc(n, 8);
  for (int i = 0; i < n; i++) M[i] = open ? A[i] : 1.0;
  const double small = open ? (double)(10 - 6) : 1.0e-6;
  double temp = 0.0, rho0 = 1.0, alpha = 1.0, omega = 1.0, rho = 0.0, beta;
  int status = VBO_OK;
  vbo_spmv_nr(n, A, ija, x, ax);
  for (int i = 0; i < n; i++) { r[i] = B[i] - ax[i]; rb[i] = r[i]; }
  int iter = 1;
  double n2r0 = 1.0, n2s0 = 1.0, n2s, n2t, n2r;
  while (iter <= itermax) {
    rho = dotp(n, rb, r);
    if (rho == 0.0) { status = VBO_WARN_BREAKDOWN; break; }
    if (iter == 1) {
      for (int i = 0; i < n; i++) p[i] = r[i];
    } else {
      beta = rho / (rho0 + 1.0e-15) * alpha / (omega + 1.0e-15);
      for (int i = 0; i < n; i++) p[i] = r[i] + beta * (p[i] - omega * v[i]);
    }
    if (!open) remove_dc(n, p);
    for (int i = 0; i < n; i++) pinv[i] = p[i] / (M[i] + 1.0e-15);
    vbo_spmv_nr(n, A, ija, pinv, v);
    temp = dotp(n, rb, v);
    alpha = rho / (temp + 1.0e-15);
    for (int i = 0; i < n; i++) s[i] = r[i] - alpha * v[i];
    n2s = dotp(n, s, s);
    if (iter == 1) n2s0 = (n2s <= small) ? 1.0 : n2s;
    if (n2s <= tol * tol * n2s0) {
      for (int i = 0; i < n; i++) x[i] = x[i] + alpha * pinv[i];
      *resid = sqrt(n2s / n2s0);
      break;
    }
    if (!open) remove_dc(n, s);
    for (int i = 0; i < n; i++) sinv[i] = s[i] / (M[i] + 1.0e-15);
    vbo_spmv_nr(n, A, ija, sinv, t);
    temp = dotp(n, t, s);
    n2t = dotp(n, t, t);
    omega = temp / (n2t + 1.0e-15);
    for (int i = 0; i < n; i++) x[i] = x[i] + alpha * pinv[i] + omega * sinv[i];
    for (int i = 0; i < n; i++) r[i] = s[i] - omega * t[i];
    n2r = dotp(n, r, r);
    if (iter == 1) n2r0 = (n2r <= small) ? 1.0 : n2r;
    *resid = sqrt(n2r / n2r0);
    if (n2r <= tol * tol * n2r0) break;
    iter = iter + 1;
    rho0 = rho;
  }
  *iters = iter;
  free(r); free(rb); free(v); free(ax); free(p); free(pinv); free(s); free(sinv); free(t); free(M);
  return status;
}

int vbo_bicgstab_nonull(int n, int nnz, const double* A, const int* ija, double* x,
                        const double* B, int itermax, double tol, int* iters, double* resid) {
  (void)nnz;
  return bicgstab(n, A, ija, x, B, itermax, tol, iters, resid, 0);
}

int vbo_bicgstab_open(int n, int nnz, const double* A, const int* ija, double* x,
                      const double* B, int itermax, double tol, int* iters, double* resid) {
  (void)nnz;
  return bicgstab(n, A, ija, x, B, itermax, tol, iters, resid, 1);
}

/* Phase 10: ppe_solve_in_house_simple_dirichlet.F90:1-158 */
static int phase_ppe(vbo_ctx* c) {
  const int npfb = c->npfb;
  double* B = (double*)xcalloc(npfb, 8);
  double temp = 0.0;
  for (int i = 0; i < npfb; i++) {                          /* :27-43 */
    double fb = 0.0;
    for (int k = 0; k < c->cnt[i]; k++) {
      size_t s = (size_t)k * npfb + i;
      int j = c->link[s];
      double uij = c->u[i] - c->u[j], vij = c->v[i] - c->v[j];
      double kw1 = c->k11[i] * c->wg1[s] + c->k12[i] * c->wg2[s];
      double kw2 = c->k21[i] * c->wg1[s] + c->k22[i] * c->wg2[s];
      fb = fb - (uij * kw1 + vij * kw2);
    }
    B[i] = fb / c->dt;
    temp = temp + B[i];
  }
  if (c->open_domain == 0) {                                /* :46-53 */
    temp = temp / (double)npfb;
    for (int i = 0; i < npfb; i++) B[i] = B[i] - temp;
  }
  /* assembly :60-122; arrays are 1-based in the reference, stored here at [idx-1] */
  int* ia = (int*)xcalloc(npfb + 1, 4);
  ia[0] = npfb + 2;
  for (int i = 0; i < npfb; i++) ia[i + 1] = ia[i] + c->cnt[i];
  int nnz = ia[npfb];
  double* A = (double*)xcalloc(nnz, 8);
  int* ija = (int*)xcalloc(nnz, 4);
  int* ijnum = (int*)xcalloc(npfb + 1, 4);
  for (int i = 0; i < npfb; i++) {
    ija[i] = ia[i];
    int num = ija[i] - 1;
    for (int k = 0; k < c->cnt[i]; k++) {
      int j = c->link[(size_t)k * npfb + i];
      int col = (j < npfb) ? j : c->irel[j];
      num = num + 1;
      ijnum[col] = num;
      ija[num - 1] = col + 1;
    }
    double adiag = 0.0;
    for (int k = 0; k < c->cnt[i]; k++) {
      size_t s = (size_t)k * npfb + i;
      int j = c->link[s];
      double t = -c->wl[s];
      if (i != j) {
        adiag = adiag - t;
        if (j >= npfb) {
          int ij = c->irel[j];
          if (i != ij) { int nt = ijnum[ij]; A[nt - 1] = A[nt - 1] + t; }
          else adiag = adiag + t;
          B[i] = B[i] - c->dP[j] * t;
        } else {
          int nt = ijnum[j]; A[nt - 1] = A[nt - 1] + t;
        }
      }
    }
    if (c->nsurf[i] != 0) adiag = 1.0e30;
    A[i] = A[i] + adiag;
  }
  ija[npfb] = ia[npfb];
  free(ia); free(ijnum);
  if (c->open_domain == 1) {                                /* PPE smoothing, :126-138 */
    for (int i = 0; i < npfb; i++) {
      if (c->nsurf[i] == 1) continue;
      if (c->psmooth[i] == 1.0) continue;
      B[i] = B[i] * c->psmooth[i];
      for (int e = ija[i]; e <= ija[i + 1] - 1; e++) A[e - 1] = A[e - 1] * c->psmooth[i];
    }
  }
  c->ppe_a = (double*)xcalloc(nnz, 8); c->ppe_ija = (int*)xcalloc(nnz, 4); c->ppe_b = (double*)xcalloc(npfb, 8);
  memcpy(c->ppe_a, A, 8 * (size_t)nnz); memcpy(c->ppe_ija, ija, 4 * (size_t)nnz); memcpy(c->ppe_b, B, 8 * (size_t)npfb);
  c->ppe_nnz = nnz;
  for (int i = 0; i < npfb; i++) c->p[i] = 0.0;             /* :143-152 */
  int st = bicgstab(npfb, A, ija, c->p, B, 1000, 1.0e-8, &c->ls_iters, &c->ls_resid,   /* :148-152 */
                    c->open_domain == 1);
  free(A); free(ija); free(B);
  return st;
}

/* Phase 11: calc_pressure_gradient (gradient_calculation.F90:16-48) and div_free.F90:93-102 */
static void phase_pressure(vbo_ctx* c) {
  const int npfb = c->npfb;
  for (int i = npfb; i < c->np; i++) c->p[i] = c->p[c->irel[i]] + c->dP[i];
  c->gpx = (double*)xcalloc(npfb, 8); c->gpy = (double*)xcalloc(npfb, 8);
  for (int i = 0; i < npfb; i++) {
    double g[2];
    grad_one(c, c->p, i, g);
    c->gpx[i] = g[0]; c->gpy[i] = g[1];
  }
  for (int i = c->nbS; i < npfb; i++) {
    c->u[i] = c->u[i] + c->dt * (c->bf[0] - c->gpx[i]);
    c->v[i] = c->v[i] + c->dt * (c->bf[1] - c->gpy[i]);
    if (c->cnt[i] <= 3) { c->u[i] = c->u0[i] + c->dt * c->bf[0]; c->v[i] = c->v0[i] + c->dt * c->bf[1]; }
  }
}

/* Phase 12: div_free.F90:111-117 final advection, and the step's time update */
static void phase_final(vbo_ctx* c) {
  for (int i = c->nbS; i < c->npfb; i++) {
    c->x[i] = c->x0[i] + 0.5 * (c->u[i] + c->u0[i]) * c->dt + c->usx[i] * c->dt;
    c->y[i] = c->y0[i] + 0.5 * (c->v[i] + c->v0[i]) * c->dt + c->usy[i] * c->dt;
  }
  c->time = c->time + c->dt;
}

int vbo_phase(vbo_ctx* c, int k) {
  int st = VBO_OK;
  switch (k) {
    case 1: phase_advect(c); break;
    case 2: st = phase_bin(c); break;
    case 3: st = phase_mirrors(c); break;
    case 4: st = phase_neighbours(c); break;
    case 5: phase_weights(c); break;
    case 6: phase_freesurf(c); break;
    case 7: phase_shift(c); break;
    case 8: phase_predict(c); break;
    case 9: phase_mirror_vel(c); break;
    case 10: st = phase_ppe(c); break;
    case 11: phase_pressure(c); break;
    case 12: phase_final(c); break;
    default: break;
  }
  c->status = st;
  return st;
}

int vbo_step(vbo_ctx* c) {
  int worst = VBO_OK;
  for (int k = 1; k <= 12; k++) {
    int st = vbo_phase(c, k);
    if (st != VBO_OK && st != VBO_WARN_BREAKDOWN) return st;
    if (st) worst = st;
  }
  return worst;
}

/* ---------------------------------------------------------------------------------------------- */
void vbo_get_mirrors(const vbo_ctx* c, int m, int* irel, int* vrel, double* dP) {
  for (int j = 0; j < m; j++) {
    irel[j] = c->irel[c->npfb + j] + 1; vrel[j] = c->vrel[c->npfb + j]; dP[j] = c->dP[c->npfb + j];
  }
}

void vbo_get_neighbours(const vbo_ctx* c, int n, int w, int* cnt, int* link) {
  for (int i = 0; i < n; i++) {
    cnt[i] = c->cnt[i];
    for (int k = 0; k < w; k++) link[(size_t)k * n + i] = (k < c->cnt[i]) ? c->link[(size_t)k * c->npfb + i] + 1 : 0;
  }
}

void vbo_get_cells(const vbo_ctx* c, int np, int* cellpart, int nct1, int* ist, int* ip) {
  for (int k = 0; k < np; k++) { cellpart[k] = c->cellpart[k]; ip[k] = c->ip[k + 1] + 1; }
  for (int ii = 0; ii < nct1; ii++) ist[ii] = c->ist[ii + 1];
}

void vbo_get_weights(const vbo_ctx* c, int n, int w, double* wg, double* wl, double* kg) {
  for (int i = 0; i < n; i++) {
    for (int k = 0; k < w; k++) {
      size_t s = (size_t)k * c->npfb + i, o = (size_t)k * n + i;
      int ok = k < c->cnt[i];
      wg[o] = ok ? c->wg1[s] : 0.0;
      wg[(size_t)n * w + o] = ok ? c->wg2[s] : 0.0;
      wl[o] = ok ? c->wl[s] : 0.0;
    }
    kg[i] = c->k11[i]; kg[n + i] = c->k12[i]; kg[2 * n + i] = c->k21[i]; kg[3 * n + i] = c->k22[i];
  }
}

void vbo_get_ushift(const vbo_ctx* c, int n, double* us) {
  for (int i = 0; i < n; i++) { us[i] = c->usx[i]; us[n + i] = c->usy[i]; }
}

void vbo_get_grad_vel(const vbo_ctx* c, int n, double* g) { memcpy(g, c->gvel, 32 * (size_t)n); }

int vbo_get_ppe_nnz(const vbo_ctx* c) { return c->ppe_nnz; }

void vbo_get_ppe(const vbo_ctx* c, int nnz, double* a, int* ija, double* b) {
  memcpy(a, c->ppe_a, 8 * (size_t)nnz); memcpy(ija, c->ppe_ija, 4 * (size_t)nnz);
  memcpy(b, c->ppe_b, 8 * (size_t)c->npfb);
}

void vbo_get_surface(const vbo_ctx* c, int n, int* nsurf, double* conc, double* divr, double* sn,
                     double* psmooth) {
  for (int i = 0; i < n; i++) {
    const int in = i < c->npfb;
    nsurf[i] = c->nsurf[i]; conc[i] = c->conc[i]; divr[i] = c->divr[i];
    sn[i] = (in && c->snx) ? c->snx[i] : 0.0; sn[n + i] = (in && c->sny) ? c->sny[i] : 0.0;
    psmooth[i] = (in && c->psmooth) ? c->psmooth[i] : 0.0;
  }
}

void vbo_get_grad_p(const vbo_ctx* c, int n, double* gp) {
  for (int i = 0; i < n; i++) { gp[i] = c->gpx[i]; gp[n + i] = c->gpy[i]; }
}

/* secondary_properties, div_free.F90:148-192 (STEP 11: after phase 11, before phase 12): the
 * vorticity of rows 1..npfb from the neighbour lists, w_G and the current velocities, and the RMS
 * velocity sqrt(sum u.u / npfb) that the reference writes to unit 192 (its reduction in index order,
 * as with one OpenMP thread). */
void vbo_secondary_properties(const vbo_ctx* c, int n, double* vort, double* u_rms) {
  const int npfb = c->npfb;
  for (int i = 0; i < n && i < npfb; i++) {                   /* :158-167 */
    double w = 0.0;
    for (int k = 0; k < c->cnt[i]; k++) {
      const int s = k * npfb + i, j = c->link[s];
      const double tmp = (c->v[i] - c->v[j]) * c->wg1[s] - (c->u[i] - c->u[j]) * c->wg2[s];
      w = w - tmp;
    }
    vort[i] = w;
  }
  double a = 0.0;                                             /* :171-181 */
  for (int i = 0; i < npfb; i++) a = a + (c->u[i] * c->u[i] + c->v[i] * c->v[i]);
  *u_rms = sqrt(a / (double)npfb);
}

/* Run diagnostics the reference does not compute (the GPU library defines them identically).
 * Force of the fluid on circle ib: sum over its boundary particles b (wall rows at distance |R| from
 * the centre, 1e-8 relative) of ds (sigma_b . n_b), sigma = -p I + c_s (G + G^T) + c_p tau,
 * n_b = (r_b - c)/|r_b - c|, ds = 2 pi |R| / N_c, G the corrected velocity gradient of
 * gradient_calculation.F90:62-71 over the step's neighbour list, c_s and c_p as predictor_mod.F90:54-56.
 * Valid after phase 11 (needs the lists and weights).  force is (nc,2) row-major. */
void vbo_circle_forces(const vbo_ctx* c, int nc, double* force, int* counts) {
  const int npfb = c->npfb;
  const double cs = c->beta * sqrt(c->Pr / c->Ra);
  /* the Newtonian build has no polymeric stress (its c_p may be 0/0 there) */
  const double cp = c->const_mod == 1 ? 0.0 : (1.0 - c->beta) * c->Pr / (c->Ra * c->El);
  for (int ib = 0; ib < nc && ib < c->ncirc; ib++) {
    const double cx = c->ccen[2 * ib], cy = c->ccen[2 * ib + 1], R = fabs(c->crad[ib]);
    int n = 0;
    for (int i = 0; i < c->nbS; i++) {
      const double dx = c->x[i] - cx, dy = c->y[i] - cy;
      if (fabs(sqrt(dx * dx + dy * dy) - R) <= 1.0e-8 * R) n++;
    }
    const double ds = n > 0 ? 2.0 * 3.14159265358979323846 * R / (double)n : 0.0;
    double fx = 0.0, fy = 0.0;
    for (int i = 0; i < c->nbS; i++) {
      const double rx = c->x[i] - cx, ry = c->y[i] - cy;
      if (!(fabs(sqrt(rx * rx + ry * ry) - R) <= 1.0e-8 * R)) continue;
      double gux = 0.0, guy = 0.0, gvx = 0.0, gvy = 0.0;
      for (int k = 0; k < c->cnt[i]; k++) {
        const int s = k * npfb + i, j = c->link[s];
        const double tu = c->u[j] - c->u[i], tv = c->v[j] - c->v[i];
        gux = gux + tu * c->wg1[s]; guy = guy + tu * c->wg2[s];
        gvx = gvx + tv * c->wg1[s]; gvy = gvy + tv * c->wg2[s];
      }
      const double dudx = c->k11[i] * gux + c->k12[i] * guy, dudy = c->k21[i] * gux + c->k22[i] * guy;
      const double dvdx = c->k11[i] * gvx + c->k12[i] * gvy, dvdy = c->k21[i] * gvx + c->k22[i] * gvy;
      const double rn = sqrt(rx * rx + ry * ry), nx = rx / rn, ny = ry / rn;
      const double p = c->p[i];
      const double sxx = -p + cs * (2.0 * dudx) + cp * c->t11[i];
      const double sxy = cs * (dudy + dvdx) + cp * c->t12[i];
      const double syx = cs * (dvdx + dudy) + cp * c->t21[i];
      const double syy = -p + cs * (2.0 * dvdy) + cp * c->t22[i];
      fx = fx + ds * (sxx * nx + sxy * ny);
      fy = fy + ds * (syx * nx + syy * ny);
    }
    force[2 * ib] = fx; force[2 * ib + 1] = fy; counts[ib] = n;
  }
}

/* Nusselt number, Nu = 1 + sqrt(Ra Pr) <v theta> H / dT over the fluid rows not in the bin, H the y
 * extent of the boundary nodes, dT the largest minus the smallest wall-patch temperature. */
double vbo_nusselt(const vbo_ctx* c) {
  double ylo = c->bnode[1], yhi = ylo, tlo = 0.0, thi = 0.0;
  int wall = 0;
  for (int i = 0; i < c->nbp; i++) {
    const double y = c->bnode[2 * i + 1];
    ylo = y < ylo ? y : ylo; yhi = y > yhi ? y : yhi;
    if (c->btype[i] != 1) continue;
    if (!wall) { tlo = thi = c->bth[i]; wall = 1; }
    tlo = c->bth[i] < tlo ? c->bth[i] : tlo; thi = c->bth[i] > thi ? c->bth[i] : thi;
  }
  if (!wall || thi - tlo <= 0.0) return 0.0;
  double s = 0.0, n = 0.0;
  for (int i = c->nbS; i < c->npfb; i++) {
    if (c->inbin[i]) continue;
    s = s + c->v[i] * c->th[i];
    n = n + 1.0;
  }
  return n > 0.0 ? 1.0 + sqrt(c->Ra * c->Pr) * (yhi - ylo) / (thi - tlo) * (s / n) : 0.0;
}

/* nnf_models_mod.F90:8-110 (no callers in the reference; restated for completeness) */
void vbo_nnf(int model, int n, const double* sr, double mu, double* out) {
  for (int i = 0; i < n; i++) {
    switch (model) {
      case 0: out[i] = sr[i]; break;                                     /* :8-21, mu array copy */
      case 1: out[i] = mu * pow(sr[i] + 1.0, 1.5 - 1.0); break;         /* :23-39 */
      case 2: {                                                          /* :41-61 */
        double ty = 1.0e-2, al = 1.0e2;
        out[i] = (sr[i] <= ty / (mu * al)) ? mu * al : mu + ty / sr[i];
      } break;
      case 3: {                                                          /* :63-82 */
        double a = 0.05, m0 = mu, mi = a * m0, kk = 1.0e-3, mm_ = 3.0;
        out[i] = mi + (m0 - mi) / (1.0 + pow(kk * sr[i], mm_));
      } break;
      case 4: {                                                          /* :84-108 */
        double m0 = mu, mi = 1.0e-3, kk = 0.14, mm_ = 0.1, aa = 1.5;
        out[i] = mi + (m0 - mi) * pow(1.0 + pow(kk * sr[i], aa), (mm_ - 1.0) / aa);
      } break;
      default: out[i] = 0.0;
    }
  }
}
