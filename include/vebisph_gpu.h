/* vebisph_gpu.h — C ABI of the MI355X (gfx950) implementation of one VEBISPH time step.
 *
 * The reference (jrcking/vebisph, Fortran 90 + OpenMP) runs one step as div_free
 * (source/base/div_free.F90:22-124), a sequence of argument-less subroutines working on module
 * globals.  Each entry point below replaces one of those calls; a context owns the particle state in
 * HBM between calls, so a Fortran driver keeps the reference's call sequence and calls these through
 * iso_c_binding (see vebisph_amd/fortran/vebisph_gpu.F90).
 *
 * Conventions
 *   - plain C types only: double, float, int32_t, pointers and sizes; no framework types;
 *   - host arrays are per-component columns, i.e. the reference's column-major (npar,2) arrays are
 *     passed as two pointers r(:,1), r(:,2); tau_p(npar,2,2) as four: tau(:,1,1), (:,1,2), (:,2,1),
 *     (:,2,2);
 *   - particle indices handed out (neighbour lists, mirror parents) are 1-based, like the
 *     reference's; neighbour lists are slot-major ("ij_link(npfb,nplink)", leading dimension ld);
 *   - every call returns a vb_status; VB_OK is 0.  vb_last_error() gives a message;
 *   - one host thread per context; all work goes to the context's HIP stream, and a call returns
 *     once its kernels are enqueued, except where it hands back host data (then it synchronises).
 */
#ifndef VEBISPH_GPU_H
#define VEBISPH_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  VB_OK = 0,
  VB_ERR_OUT_OF_GRID = 1,        /* particle outside the cell grid: neighbour_finding_mod.F90:82-88 STOP */
  VB_ERR_CELL_OVERFLOW = 2,      /* more than nplink particles in a cell: neighbour_finding_mod.F90:93-103 STOP */
  VB_ERR_ALL_BINNED = 3,         /* every fluid particle in the bin: put_in_bin.F90:51-54 STOP */
  VB_ERR_CAPACITY = 4,           /* mirror (2*npfb) or neighbour (nplink) capacity exceeded */
  VB_WARN_BICGSTAB_BREAKDOWN = 5,/* rho == 0: linear_solver.F90:332-335 prints and returns */
  VB_WARN_MAXITER = 6,           /* itermax reached without convergence (ppe_...:144-149) */
  VB_ERR_HIP = 7,                /* a HIP runtime call failed */
  VB_ERR_ARG = 8,                /* invalid argument */
} vb_status;

enum { VB_NPLINK = 75 };         /* common_parameter.F90:5, neighbour / cell cap */
enum { VB_MAX_PATCHES = 16, VB_MAX_CIRCLES = 8 };

typedef enum { VB_PREC_FP64 = 0, VB_PREC_MIXED = 1 } vb_precision;

/* Run parameters: the INDAT values of input.F90:18-37 that the step uses. */
typedef struct {
  double h, dx, dv;              /* smoothing length, particle spacing, particle volume (dx*dx) */
  double Ra, Pr, El, beta;       /* dimensionless groups; Wi = El*sqrt(Ra/Pr) */
  double body_force[2];
  double dt_coef_cfl, dt_coef_visc;
  double ve_nonlinearity;        /* eps of linear PTT; 0 = Oldroyd-B */
  double time;                   /* start time */
  int32_t itime;                 /* steps already taken */
  int32_t i_open_domain;         /* 0 closed domain; 1 open domain with a free surface
                                    (freesurf_mod.F90, Dirichlet surface rows, Jacobi BiCGSTAB) */
  int32_t i_shift_coeff;         /* 0: D = 0.25 h^2; 1: D = 4 h dt (|u| + 0.2 exp(-|u|/(0.2 maxU)) maxU)
                                    (new_fick_shift.F90:43-51) */
  int32_t precision;             /* vb_precision */
  int32_t reference_kick;        /* 1: reproduce predictor_mod.F90:369-371 as the reference runs it */
  int32_t serial_reductions;     /* 1: global sums in index order on one thread, as the reference's
                                    single-thread loops (bit-for-bit parity runs; slow at scale) */
  int32_t fluid_model;           /* the reference's compile-time fluid model (its -Dconst_mod): 0 or 2
                                    viscoelastic (log-conformation linear PTT / Oldroyd-B), 1 Newtonian
                                    (no polymeric stress; predictor_mod.F90:58-77) */
} vb_params;

/* Boundary patches and circles: IBOUND, input.F90:51-81.  b_node is (nb_patches,2) row-major;
 * b_edge(ib) = b_node(ib+1) - b_node(ib).  Circle arrays are (nb_circles[,2]) row-major. */
typedef struct {
  int32_t nb_patches;
  const double* b_node;
  const int32_t* b_type;         /* 0 none, 1 wall, 2 periodic (3/4 inflow/outflow unsupported) */
  const double* b_vel;
  const double* b_thermal;
  const int32_t* b_periodic_parent; /* 1-based */
  int32_t nb_circles;
  const double* c_centre;
  const double* c_radius;        /* > 0: solid disc; < 0: fluid inside */
  const double* c_omega;
  const double* c_vel;
  const double* c_thermal;
} vb_boundary;

typedef struct vb_ctx vb_ctx;

/* Context: replaces the module globals of common_2d.F90:20-80 and the allocations of
 * input.F90:105-114.  capacity rows = npfb + mirror capacity (2*npfb, 10*npfb when npfb <= 500,
 * mirror_boundaries_mod.F90:53-54).  Neighbour slots are addressed with 32-bit indices, so npfb is
 * at most (2^31 - 1) / VB_NPLINK = 28,633,115: larger returns VB_ERR_CAPACITY before any device call. */
vb_status vb_create(const vb_params* prm, const vb_boundary* bnd, int32_t npfb, int32_t nbS, vb_ctx** out);
vb_status vb_destroy(vb_ctx* ctx);
const char* vb_last_error(const vb_ctx* ctx);

/* State transfer (explicit; never done inside a step).  Rows 1..n of the reference arrays r, u,
 * theta, tau_p, trace_conf, p, inbin (common_2d.F90:49-61).  download may ask for n up to np to
 * include the mirror rows of the last step. */
vb_status vb_upload_state(vb_ctx* ctx, int32_t n, const double* x, const double* y, const double* u,
                          const double* v, const double* theta, const double* tau11, const double* tau12,
                          const double* tau21, const double* tau22, const double* trace_conf,
                          const double* p, const int32_t* inbin);
vb_status vb_download_state(vb_ctx* ctx, int32_t n, double* x, double* y, double* u, double* v,
                            double* theta, double* tau11, double* tau12, double* tau21, double* tau22,
                            double* trace_conf, double* p, int32_t* inbin);

/* ---- one time step, phase by phase (the call sequence of div_free.F90:38-122) ---- */
/* div_free.F90:39-53 with set_time_step (:126-146): r0,u0 copies, dt, r = r0 + u0 dt */
vb_status vb_set_time_step(vb_ctx* ctx);
/* put_in_bin.F90:1-56 (x_mint.. are the padded box of sph2D_incom.F90:64-67, held by the context) */
vb_status vb_put_in_bin(vb_ctx* ctx);
/* create_mirror_particles, mirror_boundaries_mod.F90:18-334 */
vb_status vb_create_mirror_particles(vb_ctx* ctx);
/* particle_allocation: divide + neighboring_list_parallel, neighbour_finding_mod.F90:15-127,261-365 */
vb_status vb_particle_allocation(vb_ctx* ctx);
/* kernel_calculation.F90:2-75; in an open domain the same pass also takes calc_part_dist_vars
 * (freesurf_mod.F90:70-103), the free-surface criterion (:27-38) and ppe_smooth (:54-64) */
vb_status vb_kernel_calculation(vb_ctx* ctx);
/* locate_free_surf_part, freesurf_mod.F90:17-68: n_surf = 0 (closed domain), or the mirrors take
 * their parent's flag (:43-45); smooth_normals (:105-137) is folded into vb_new_fick_shift, the only
 * user of the smoothed normals */
vb_status vb_locate_free_surf_part(vb_ctx* ctx);
/* new_fick_shift.F90:15-92,117-210 and ushift, div_free.F90:65-78 (open-domain surface particles
 * shift along the surface only, :60-73) */
vb_status vb_new_fick_shift(vb_ctx* ctx);
/* prediction_step, predictor_mod.F90:27-139 (log-conformation :141-331, mirror stress
 * mirror_boundaries_mod.F90:586-653) */
vb_status vb_prediction_step(vb_ctx* ctx);
/* mirror_velocities, mirror_boundaries_mod.F90:509-584 */
vb_status vb_mirror_velocities(vb_ctx* ctx);
/* ppe_solve_in_house_simple_dirichlet.F90:1-158 with SPARSE_PBICGSTAB_nonull
 * (linear_solver.F90:275-425), or in an open domain (i_open_domain = 1) with the Dirichlet surface
 * rows, ppe_smooth-scaled rows near the surface and SPARSE_PBICGSTAB (:82-227).
 * tol <= 0 / itermax <= 0 select the reference's 1e-8 / 1000.
 * iters and residual may be NULL; when given, the call synchronises to return them. */
vb_status vb_ppe_solve(vb_ctx* ctx, double tol, int32_t itermax, int32_t* iters, double* residual);
/* SPARSE_PBICGSTAB_nonull (linear_solver.F90:275-285) on a caller's system, same argument meaning:
 * A(nmax), ija(nmax) in Numerical-Recipes row-indexed storage (1-based, A(1:n) the diagonal,
 * ija(1) = n+2), x(n) initial guess in / solution out, B(n).  Rows may hold at most VB_NPLINK
 * off-diagonal slots.  Uses the context's PPE buffers (n must equal the context's npfb) and its
 * reduction mode; returns VB_WARN_* like vb_ppe_solve. */
vb_status vb_sparse_pbicgstab_nonull(vb_ctx* ctx, const double* A, const int32_t* ija, double* x,
                                     const double* B, int32_t n, int32_t nmax, int32_t itermax,
                                     double tol, int32_t* iter, double* residual);
/* SPARSE_PBICGSTAB (linear_solver.F90:82-227), the open-domain solver, on a caller's system with the
 * arguments of vb_sparse_pbicgstab_nonull: Jacobi preconditioner M = A(1:n), no mean removal, and
 * the reference's first-iteration norm threshold 10-6 (an integer expression, 4). */
vb_status vb_sparse_pbicgstab(vb_ctx* ctx, const double* A, const int32_t* ija, double* x, const double* B,
                              int32_t n, int32_t nmax, int32_t itermax, double tol, int32_t* iter,
                              double* residual);
/* calc_pressure_gradient (gradient_calculation.F90:16-48) and div_free.F90:93-102 */
vb_status vb_pressure_update(vb_ctx* ctx);
/* div_free.F90:111-117 final advection; advances time */
vb_status vb_final_advection(vb_ctx* ctx);
/* the whole step: the calls above in order (one synchronisation for the step's scalars) */
vb_status vb_div_free(vb_ctx* ctx);

/* ---- output quantities of the reference's driver (these synchronise) ---- */
/* secondary_properties (div_free.F90:148-192, STEP 11 of an output step): the vorticity
 * vort(1..n), n <= npfb, from the last step's velocities, neighbour lists and w_G, and
 * u_rms = sqrt(sum_i u_i.u_i / npfb) (the value the reference writes to unit 192).  The reference
 * computes them between the velocity update and the final advection, which changes neither, so
 * call this after vb_pressure_update or vb_div_free.  vort and u_rms may be NULL. */
vb_status vb_secondary_properties(vb_ctx* ctx, int32_t n, double* vort, double* u_rms);
/* The assembled PPE as an operator, for an adjoint solve (no counterpart in the reference, which has
 * no adjoint): y = A x (transpose = 0) or y = A^T x (transpose = 1), x and y host arrays of n = npfb
 * rows, A the matrix the last vb_ppe_solve assembled from the step's lists
 * (ppe_solve_in_house_simple_dirichlet.F90:55-138, mirrors folded onto their parents, Dirichlet and
 * ramp rows included).  A^T comes from a transposed map built on the first transpose call after each
 * assembly, with a fixed summation order.  transpose = 2 / 3 apply A^T / A through the solvers' own
 * SpMV rows instead (the adjoint solve's compact A^T rows and map rows; the forward FP64 solve's
 * compact rows in column order): the same operator, summed in the solvers' order.  VB_ERR_ARG once
 * the lists are rebuilt by the next step's vb_particle_allocation, or after vb_sparse_pbicgstab*
 * replaced the stored matrix. */
vb_status vb_ppe_operator(vb_ctx* ctx, int32_t transpose, int32_t n, const double* x, double* y);
/* The SPH gradient of calc_pressure_gradient (gradient_calculation.F90:16-48) as a linear operator, and
 * its transpose, for an adjoint step (no counterpart in the reference): row i of G x is
 * K_i sum_k (x_parent(j_k) - x_i) grad W_ik with the step's lists, kernel gradients and KGC matrices,
 * mirrors taking their parent's value.  transpose = 0: x has n = npfb values, y = [gx; gy] 2n;
 * transpose = 1: x = [gx; gy] 2n, y = G^T x, n values.  Host arrays; valid, as vb_ppe_operator, from
 * the step's vb_ppe_solve until the next vb_particle_allocation. */
vb_status vb_grad_operator(vb_ctx* ctx, int32_t transpose, int32_t n, const double* x, double* y);
/* The adjoint PPE solve A^T lambda = g (no counterpart in the reference), A as for vb_ppe_operator:
 * FP64 iterative refinement around the reference's own BiCGSTAB with A^T as its matrix --
 * SPARSE_PBICGSTAB_nonull in a closed domain, where the mean of g and of each refinement residual is
 * removed (A^T maps into the mean-zero vectors; this projection is the adjoint of the forward solve's
 * removal of the mean of its right-hand side), SPARSE_PBICGSTAB in an open one.  Each round solves
 * A^T d = r / |r| and adds |r| d to lambda (from lambda = 0), then recomputes r = g - A^T lambda in
 * FP64, until |r| <= tol |g| (BiCGSTAB alone stalls near 3e-9 on A^T, which is not symmetric in
 * wall-bounded domains; on a VB_PREC_MIXED context the inner solve reads the compact rows' A^T values
 * in FP32, which are exact there, while the residual stays FP64); tol <= 0 / itermax <= 0 select
 * 1e-8 / 1000, itermax bounding the inner iterations in total.  g, lambda host arrays of n = npfb;
 * iters (inner iterations) and residual
 * (|r| / |g| at the end) may be NULL; returns VB_WARN_* like vb_ppe_solve. */
vb_status vb_ppe_adjoint_solve(vb_ctx* ctx, int32_t n, const double* g, double* lambda, double tol,
                               int32_t itermax, int32_t* iters, double* residual);
/* ---- the adjoint of a run of steps (no counterpart in the reference, which has no adjoint) ---- */
/* Records a forward run for vb_adjoint_backward: nsteps steps of vb_div_free from the current state
 * (each PPE solved to tol / itermax; <= 0 selects the reference's 1e-8 / 1000), keeping in device
 * memory the state at the start of every ckpt_every-th step and the pressure of every step, plus the
 * sweep's work arrays -- all allocated here, none during the steps or the sweep.  The context ends at
 * the run's final state.  ppe_iters (nsteps) may be NULL.  Closed domains without circles and with
 * i_shift_coeff = 0 only (VB_ERR_ARG otherwise).  Returns VB_WARN_MAXITER or
 * VB_WARN_BICGSTAB_BREAKDOWN when a solve ended so. */
vb_status vb_adjoint_record(vb_ctx* ctx, int32_t nsteps, int32_t ckpt_every, double tol, int32_t itermax,
                            int32_t* ppe_iters);
/* The reverse sweep over the recorded run, for a loss L of its final state: given dL/d(final state)
 * (host arrays of n = npfb, each may be NULL for zero: positions, velocities, temperature, the four
 * stress components, the conformation trace), returns grads3 = (dL/dWi, dL/dbeta, dL/dRa) with Pr
 * fixed and El = Wi sqrt(Pr/Ra) (so dL/dRa includes dt's viscous limit, div_free.F90:142), and, when
 * the *_bar0 arrays are given, dL/d(start state).  Each step is differentiated with its neighbour
 * lists and mirror set frozen and its PPE solve taken as exact; the adjoint PPE solves A^T lambda = g
 * with the reference's BiCGSTAB to tol / itermax.  Segments between checkpoints are recomputed from
 * their checkpoint with the stored pressures.  adj_iters (nsteps) may be NULL.  The context is left at
 * the recorded final state; the sweep may be run again with other cotangents. */
vb_status vb_adjoint_backward(vb_ctx* ctx, int32_t n, const double* x_bar, const double* y_bar,
                              const double* u_bar, const double* v_bar, const double* theta_bar,
                              const double* tau11_bar, const double* tau12_bar, const double* tau21_bar,
                              const double* tau22_bar, const double* trace_bar, double tol, int32_t itermax,
                              double* grads3, double* x_bar0, double* y_bar0, double* u_bar0, double* v_bar0,
                              double* theta_bar0, double* tau11_bar0, double* tau12_bar0, double* tau21_bar0,
                              double* tau22_bar0, double* trace_bar0, int32_t* adj_iters);
/* The columns conc and n_surf of the PART files (output.F90:47-49) for rows 1..n, n <= np: the
 * free-surface flag and, in an open domain, the concentration (calc_part_dist_vars,
 * freesurf_mod.F90:95-100) of the last step (rows past npfb: mirror flags, conc 0; conc is 0 in a
 * closed domain, where the reference never sets it).  Either pointer may be NULL. */
vb_status vb_free_surface(vb_ctx* ctx, int32_t n, int32_t* n_surf, double* conc);
/* maxval(abs(u(1:npfb,1))), maxval(abs(u(1:npfb,2))), maxval(abs(P(1:npfb))): the per-step line of
 * data_directory/TIME (output.F90:86), from the current state */
vb_status vb_field_maxima(vb_ctx* ctx, double* maxima3);

/* ---- run diagnostics (no counterpart in the reference, which computes neither; defined here and
 * identically in the CPU restatement) ---- */
/* Force of the fluid on each circle (the cylinder of the flow-past-a-cylinder runs):
 * F = sum over the circle's boundary particles b of ds (sigma_b . n_b), sigma = -p I +
 * c_s (G + G^T) + c_p tau, n_b the circle's outward unit normal at b, ds = 2 pi |R| / N_c, G the
 * corrected velocity gradient of gradient_calculation.F90:62-71 over the step's neighbour list, and
 * c_s, c_p the viscous and polymeric coefficients of predictor_mod.F90:54-56.  Boundary particles of
 * circle ib are the wall rows at distance |R| (1e-8 relative) from its centre.  From the state after
 * the velocity update: call after vb_pressure_update or vb_div_free.  force is (nc,2) row-major,
 * counts(nc) the boundary particles found; nc must equal the context's circle count. */
vb_status vb_circle_forces(vb_ctx* ctx, int32_t nc, double* force, int32_t* counts);
/* Nusselt number of the heated-from-below runs, Nu = 1 + sqrt(Ra Pr) <v theta> H / dT, with <.> the
 * mean over the fluid rows not in the bin, H the y extent of the boundary nodes and dT the largest
 * minus the smallest wall-patch temperature (VB_ERR_ARG when that is 0). */
vb_status vb_nusselt(vb_ctx* ctx, double* nu);

/* ---- scalars and exports (for the host driver and parity checks; these synchronise) ---- */
/* ints: np, npfb, nbS, nmirror, maxneighbours, LS_iters, itime, n_inbin, ncx, ncy, nct, status
 * dbls: dt, LS_residual, time, umax, x_mint, x_maxt, y_mint, y_maxt */
vb_status vb_get_scalars(vb_ctx* ctx, int32_t* ints12, double* dbls8);
/* irelation, vrelation, dP_mp of the m mirrors (rows npfb+1..np), 1-based parents */
vb_status vb_export_mirrors(vb_ctx* ctx, int32_t m, int32_t* irelation, int32_t* vrelation, double* dP_mp);
/* The cell lists of the step's divide (neighbour_finding_mod.F90:56-127), as the reference holds them,
 * 1-based: cellpart(1..np) the cell of each particle (0 for binned particles), ist(1..nct+1) the first
 * position of each cell in ip (ist(nct+1) - 1 particles binned into cells), ip(1..np) the particles in
 * cell order, stable within a cell (0 past the binned count).  np and nct1 = nct + 1 must equal the
 * step's (vb_get_scalars). */
vb_status vb_export_cells(vb_ctx* ctx, int32_t np, int32_t* cellpart, int32_t nct1, int32_t* ist, int32_t* ip);
/* ij_count(1..n) and ij_link(1..n, 1..w), slot-major with leading dimension n, 1-based, 0-padded */
vb_status vb_export_neighbours(vb_ctx* ctx, int32_t n, int32_t w, int32_t* ij_count, int32_t* ij_link);
/* ij_w_G(n,w,2), ij_w_L(n,w) slot-major; kgcm as four columns K11,K12,K21,K22 (n each) */
vb_status vb_export_weights(vb_ctx* ctx, int32_t n, int32_t w, double* w_G, double* w_L, double* kgcm);
vb_status vb_export_ushift(vb_ctx* ctx, int32_t n, double* ushift);      /* (n,2) column-major */
vb_status vb_export_grad_p(vb_ctx* ctx, int32_t n, double* grad_p);      /* (n,2) column-major */
/* PPE before the solve, ELL form: diag(n), col(n,w)/val(n,w) slot-major (1-based columns; slots
 * beyond ij_count are 0), rhs(n).  In the reference's Numerical-Recipes storage row i holds
 * A(i)=diag and, at ija(i)+k-1, the k-th slot. */
vb_status vb_export_ppe(vb_ctx* ctx, int32_t n, int32_t w, double* diag, int32_t* col, double* val, double* rhs);

/* ---- measurement (no reference counterpart) ---- */
/* Times `reps` launches of the BiCGSTAB SpMV kernel on the context's current PPE matrix with HIP
 * events on the context's stream; returns the mean duration (microseconds) and the algorithmic HBM
 * bytes of one launch (row data + ELL columns/values of the nnz actually used, gathered vector
 * counted once).  Leaves the particle state untouched; call after a step. */
vb_status vb_profile_spmv(vb_ctx* ctx, int32_t reps, double* avg_us, double* alg_bytes);
/* Times `reps` launches of a plain device copy of `bytes` bytes (kernel vb_copy_calibration, 16-byte
 * loads and stores) between two temporary buffers on the context's stream: the HBM rate a streaming
 * kernel reaches, and the known-size kernel that scales hardware byte counters. */
vb_status vb_profile_copy(vb_ctx* ctx, int64_t bytes, int32_t reps, double* avg_us);

/* ---- stand-alone device kernels on caller-owned device memory (stream 0) ---- */
/* nnf_models_mod.F90:8-110 (no call sites in the reference): model 0 newtonian (copies mu_or_sr),
 * 1 power law, 2 Bingham, 3 Cross, 4 Carreau; d_sr and d_mu_eff are device pointers of n doubles */
vb_status vb_nnf(int32_t model, int32_t n, const double* d_sr, double mu, double* d_mu_eff);
/* fac and Wab of sphtools.F90:11-76 on n device values */
vb_status vb_kernel_functions(int32_t n, const double* d_q, double ad_7, double ad_7h, double* d_fac, double* d_wab);
/* log-conformation update of n particles (predictor_mod.F90:165-245): tau, grad_vel as (n,4)
 * row-major device arrays, trace in/out; kick selects the reference's B perturbation */
vb_status vb_logconf(int32_t n, const double* d_tau, const double* d_trace, const double* d_gradvel,
                     double dt, double wi, double eps_ptt, int32_t kick, double* d_tau_out, double* d_trace_out);
/* vb_logconf without any device allocation (no counterpart in the reference: for callers that
 * cannot allocate -- code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as vb_logconf; with kick = 1 the per-particle kick table
 * is built in a fixed device scratch array of the library (one per GPU, held until the process
 * ends) instead of a hipMalloc'd buffer, so n is at most 1,048,576 (VB_ERR_CAPACITY above).
 * Every call on a GPU shares that GPU's scratch: calls that overlap on one GPU (several host
 * threads, or other streams) race and must be serialised by the caller. */
vb_status vb_logconf_noalloc(int32_t n, const double* d_tau, const double* d_trace, const double* d_gradvel,
                             double dt, double wi, double eps_ptt, int32_t kick, double* d_tau_out,
                             double* d_trace_out);
/* Adjoint of vb_logconf (no counterpart in the reference, which has no adjoint; one piece of an
 * adjoint step): given the cotangents of its outputs, d_tau_out_bar (n,4) and d_trace_out_bar (n),
 * writes J^T times them for the inputs at the same (tau, trace, grad_vel, Wi, dt): d_tau_bar (n,4),
 * d_trace_bar (n), d_gradvel_bar (n,4), and per-particle contributions d_wi_bar (n) and d_dt_bar (n)
 * to the shared Wi and dt (the caller sums them).  Exact FP64 derivatives of the branch each value
 * takes (the A21 = 0 branch and the discriminant clamp of predictor_mod.F90:268,280 included).
 * All pointers are device arrays and required. */
vb_status vb_logconf_adjoint(int32_t n, const double* d_tau, const double* d_trace, const double* d_gradvel,
                             double dt, double wi, double eps_ptt, int32_t kick, const double* d_tau_out_bar,
                             const double* d_trace_out_bar, double* d_tau_bar, double* d_trace_bar,
                             double* d_gradvel_bar, double* d_wi_bar, double* d_dt_bar);

#ifdef __cplusplus
}
#endif
#endif
