/* CPU restatement of the reference's per-step path (jrcking/vebisph, div_free.F90 and callees).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, by the smoke check and by bench.py's CPU-baseline leg,
 * always as the checker.  The product path (vebisph_amd/) never links or calls it.
 *
 * FP64 throughout, single-threaded, expression order as in the reference so that, built with
 * -ffp-contract=off, it reproduces the reference (compiled in oracle/_ref) bit for bit; the tests
 * check exactly that.  Particle indices are 0-based here; where an array is handed out for
 * comparison with the reference (neighbour lists, mirror parents) the getters say so.
 */
#ifndef VB_ORACLE_H
#define VB_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct vbo_ctx vbo_ctx;

enum {
  VBO_OK = 0,
  VBO_ERR_OUT_OF_GRID = 1,     /* neighbour_finding_mod.F90:82-88 */
  VBO_ERR_CELL_OVERFLOW = 2,   /* neighbour_finding_mod.F90:93-103 */
  VBO_ERR_ALL_BINNED = 3,      /* put_in_bin.F90:51-54 */
  VBO_ERR_CAPACITY = 4,        /* mirror or neighbour capacity exceeded */
  VBO_WARN_BREAKDOWN = 5,      /* linear_solver.F90:332-335 */
};

/* iprm: npfb, nbS, i_open_domain, i_shift_coeff, capacity(rows), itime, const_mod (the reference's
 *       build: 1 Newtonian, 2 viscoelastic)
 * prm:  h, dx, dv, Ra, Pr, El, beta, bf_x, bf_y, dt_coef_cfl, dt_coef_visc, ve_nonlinearity, time
 * b_node is (nbp,2) row-major; c_centre, c_vel are (ncirc,2) row-major. */
vbo_ctx* vbo_create(const int* iprm, const double* prm, int nbp, const double* b_node,
                    const int* b_type, const double* b_vel, const double* b_thermal,
                    const int* b_periodic_parent, int ncirc, const double* c_centre,
                    const double* c_radius, const double* c_omega, const double* c_vel,
                    const double* c_thermal);
void vbo_destroy(vbo_ctx* c);

/* tau is (n,4) row-major: tau11, tau12, tau21, tau22 per particle */
void vbo_set_particles(vbo_ctx* c, int n, const double* x, const double* y, const double* u,
                       const double* v, const double* th, const double* tau, const double* trc,
                       const double* p, const int* inbin);
void vbo_get_particles(const vbo_ctx* c, int n, double* x, double* y, double* u, double* v,
                       double* th, double* tau, double* trc, double* p, int* inbin);
/* ints: np, npfb, nbS, nmirror, maxneighbours, LS_iters, itime, n_inbin, ncx, ncy, nct, status
 * dbls: dt, LS_residual, time, umax, x_mint, x_maxt, y_mint, y_maxt */
void vbo_get_scalars(const vbo_ctx* c, int* ints, double* dbls);

/* One phase of div_free (1..12, same split as the reference harness) or the whole step. */
int vbo_phase(vbo_ctx* c, int k);
int vbo_step(vbo_ctx* c);

/* Intermediate results (valid after the phase that makes them, until phase 12). */
void vbo_get_mirrors(const vbo_ctx* c, int m, int* irel_1based, int* vrel, double* dP);
void vbo_get_neighbours(const vbo_ctx* c, int n, int w, int* cnt, int* link_1based_slotmajor);
void vbo_get_cells(const vbo_ctx* c, int np, int* cellpart_1based, int nct1, int* ist_1based, int* ip_1based);
void vbo_get_weights(const vbo_ctx* c, int n, int w, double* wg_nw2, double* wl_nw, double* kg_n4);
void vbo_get_ushift(const vbo_ctx* c, int n, double* us_n2);
void vbo_get_grad_vel(const vbo_ctx* c, int n, double* g_n4);
/* PPE as assembled (before the solve): NR row-indexed arrays, 1-based ija, length nnz. */
int  vbo_get_ppe_nnz(const vbo_ctx* c);
void vbo_get_ppe(const vbo_ctx* c, int nnz, double* a, int* ija, double* b);
void vbo_get_grad_p(const vbo_ctx* c, int n, double* gp_n2);
/* secondary_properties (div_free.F90:148-192), valid after phase 11: vorticity of rows 1..n and
 * the RMS velocity */
void vbo_secondary_properties(const vbo_ctx* c, int n, double* vort, double* u_rms);
/* Run diagnostics (no reference counterpart; the GPU library defines them identically): force of the
 * fluid on each circle (valid after phase 11; force (nc,2) row-major) and the Nusselt number. */
void vbo_circle_forces(const vbo_ctx* c, int nc, double* force, int* counts);
double vbo_nusselt(const vbo_ctx* c);
/* Free-surface quantities of rows 1..n (locate_free_surf_part, freesurf_mod.F90:17-68), valid from
 * phase 6: n_surf, conc, div_r; in an open domain also the smoothed surface normal (n,2) and
 * ppe_smooth until phase 12 (zeros otherwise). */
void vbo_get_surface(const vbo_ctx* c, int n, int* nsurf, double* conc, double* divr, double* sn,
                     double* psmooth);

/* Stand-alone pieces. */
void vbo_fac_wab(int n, const double* q, double ad_7, double ad_7h, double* fac, double* wab);
void vbo_calc_eigens(int n, const double* a_n4, double* lam_n2, double* r_n4);
void vbo_logconf(int n, const double* tau_n4, const double* trc, const double* gradvel_n4,
                 double dt, double wi, double eps_ptt, int use_kick, double* tau_out_n4, double* trc_out);
int  vbo_bicgstab_nonull(int n, int nnz, const double* a, const int* ija, double* x,
                         const double* b, int itermax, double tol, int* iters, double* resid);
/* SPARSE_PBICGSTAB (linear_solver.F90:82-227), the open-domain solve: Jacobi, no DC projection */
int  vbo_bicgstab_open(int n, int nnz, const double* a, const int* ija, double* x,
                       const double* b, int itermax, double tol, int* iters, double* resid);
void vbo_spmv_nr(int n, const double* a, const int* ija, const double* x, double* y);
/* nnf_models_mod.F90: model 0 newtonian(mu array), 1 power law, 2 bingham, 3 cross, 4 carreau */
void vbo_nnf(int model, int n, const double* sr, double mu, double* mu_eff);

#ifdef __cplusplus
}
#endif
#endif
