!! Fortran side of the gfx950 VEBISPH step.
!!
!! module vebisph_gpu   iso_c_binding interfaces to every entry point of include/vebisph_gpu.h
!!                      (libvebisph_hip.so), with the vb_params / vb_boundary records as bind(C)
!!                      derived types.
!! module divfree_gpu   a drop-in for the reference's module divfree (source/base/div_free.F90:1-124):
!!                      the public subroutine div_free(x_mint,y_mint,x_maxt,y_maxt) runs the same
!!                      call sequence (div_free.F90:38-122) on the device, with the particle state
!!                      resident in HBM between calls.  vbf_init / vbf_upload / vbf_download move
!!                      the state between the reference's host arrays (r(npar,2), u(npar,2),
!!                      tau_p(npar,2,2), theta, trace_conf, p, inbin) and the device context.
!! vbf_drive            a C-callable driver (used by the tests): builds the reference-shaped host
!!                      arrays, runs nsteps of div_free through the two modules and hands the state
!!                      back.
module vebisph_gpu
  use iso_c_binding
  implicit none
  public

  integer(c_int), parameter :: VB_OK = 0, VB_ERR_OUT_OF_GRID = 1, VB_ERR_CELL_OVERFLOW = 2, &
       VB_ERR_ALL_BINNED = 3, VB_ERR_CAPACITY = 4, VB_WARN_BICGSTAB_BREAKDOWN = 5, &
       VB_WARN_MAXITER = 6, VB_ERR_HIP = 7, VB_ERR_ARG = 8
  integer(c_int), parameter :: VB_NPLINK = 75, VB_MAX_PATCHES = 16, VB_MAX_CIRCLES = 8
  integer(c_int), parameter :: VB_PREC_FP64 = 0, VB_PREC_MIXED = 1

  !! INDAT run parameters (input.F90:18-37); layout of vb_params
  type, bind(C) :: vb_params
     real(c_double) :: h, dx, dv, Ra, Pr, El, beta
     real(c_double) :: body_force(2)
     real(c_double) :: dt_coef_cfl, dt_coef_visc, ve_nonlinearity, time
     integer(c_int32_t) :: itime, i_open_domain, i_shift_coeff, precision, reference_kick, &
          serial_reductions, fluid_model
  end type vb_params

  !! IBOUND boundary description (input.F90:51-81); arrays row-major, see vebisph_gpu.h
  type, bind(C) :: vb_boundary
     integer(c_int32_t) :: nb_patches
     type(c_ptr) :: b_node, b_type, b_vel, b_thermal, b_periodic_parent
     integer(c_int32_t) :: nb_circles
     type(c_ptr) :: c_centre, c_radius, c_omega, c_vel, c_thermal
  end type vb_boundary

  interface
     function vb_create(prm, bnd, npfb, nbs, ctx) bind(C, name="vb_create") result(s)
       import :: c_int, c_int32_t, c_ptr, vb_params, vb_boundary
       type(vb_params), intent(in) :: prm
       type(vb_boundary), intent(in) :: bnd
       integer(c_int32_t), value :: npfb, nbs
       type(c_ptr), intent(out) :: ctx
       integer(c_int) :: s
     end function vb_create

     function vb_destroy(ctx) bind(C, name="vb_destroy") result(s)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: s
     end function vb_destroy

     function vb_last_error(ctx) bind(C, name="vb_last_error") result(msg)
       import :: c_ptr
       type(c_ptr), value :: ctx
       type(c_ptr) :: msg
     end function vb_last_error

     function vb_upload_state(ctx, n, x, y, u, v, theta, t11, t12, t21, t22, trc, p, inbin) &
          bind(C, name="vb_upload_state") result(s)
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: n
       real(c_double), intent(in) :: x(*), y(*), u(*), v(*), theta(*), t11(*), t12(*), t21(*), &
            t22(*), trc(*), p(*)
       integer(c_int32_t), intent(in) :: inbin(*)
       integer(c_int) :: s
     end function vb_upload_state

     function vb_download_state(ctx, n, x, y, u, v, theta, t11, t12, t21, t22, trc, p, inbin) &
          bind(C, name="vb_download_state") result(s)
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: n
       real(c_double), intent(out) :: x(*), y(*), u(*), v(*), theta(*), t11(*), t12(*), t21(*), &
            t22(*), trc(*), p(*)
       integer(c_int32_t), intent(out) :: inbin(*)
       integer(c_int) :: s
     end function vb_download_state

     !! ---- the phases of one step ----
     function vb_set_time_step(ctx) bind(C, name="vb_set_time_step") result(s)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: s
     end function vb_set_time_step
     function vb_put_in_bin(ctx) bind(C, name="vb_put_in_bin") result(s)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: s
     end function vb_put_in_bin
     function vb_create_mirror_particles(ctx) bind(C, name="vb_create_mirror_particles") result(s)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: s
     end function vb_create_mirror_particles
     function vb_particle_allocation(ctx) bind(C, name="vb_particle_allocation") result(s)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: s
     end function vb_particle_allocation
     function vb_kernel_calculation(ctx) bind(C, name="vb_kernel_calculation") result(s)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: s
     end function vb_kernel_calculation
     function vb_locate_free_surf_part(ctx) bind(C, name="vb_locate_free_surf_part") result(s)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: s
     end function vb_locate_free_surf_part
     function vb_new_fick_shift(ctx) bind(C, name="vb_new_fick_shift") result(s)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: s
     end function vb_new_fick_shift
     function vb_prediction_step(ctx) bind(C, name="vb_prediction_step") result(s)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: s
     end function vb_prediction_step
     function vb_mirror_velocities(ctx) bind(C, name="vb_mirror_velocities") result(s)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: s
     end function vb_mirror_velocities
     !! iters / residual: c_loc of the outputs, or c_null_ptr (no synchronisation)
     function vb_ppe_solve(ctx, tol, itermax, iters, residual) bind(C, name="vb_ppe_solve") result(s)
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), value :: tol
       integer(c_int32_t), value :: itermax
       type(c_ptr), value :: iters, residual
       integer(c_int) :: s
     end function vb_ppe_solve
     !! SPARSE_PBICGSTAB_nonull(A,ija,x,B,n,nmax,itermax,tol,iter,residual), linear_solver.F90:275-285
     function vb_sparse_pbicgstab_nonull(ctx, a, ija, x, b, n, nmax, itermax, tol, iter, residual) &
          bind(C, name="vb_sparse_pbicgstab_nonull") result(s)
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: a(*)
       integer(c_int32_t), intent(in) :: ija(*)
       real(c_double), intent(inout) :: x(*)
       real(c_double), intent(in) :: b(*)
       integer(c_int32_t), value :: n, nmax, itermax
       real(c_double), value :: tol
       integer(c_int32_t), intent(out) :: iter
       real(c_double), intent(out) :: residual
       integer(c_int) :: s
     end function vb_sparse_pbicgstab_nonull
     !! SPARSE_PBICGSTAB(A,ija,x,B,n,nmax,itermax,tol,iter,residual), linear_solver.F90:82-227
     function vb_sparse_pbicgstab(ctx, a, ija, x, b, n, nmax, itermax, tol, iter, residual) &
          bind(C, name="vb_sparse_pbicgstab") result(s)
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: a(*)
       integer(c_int32_t), intent(in) :: ija(*)
       real(c_double), intent(inout) :: x(*)
       real(c_double), intent(in) :: b(*)
       integer(c_int32_t), value :: n, nmax, itermax
       real(c_double), value :: tol
       integer(c_int32_t), intent(out) :: iter
       real(c_double), intent(out) :: residual
       integer(c_int) :: s
     end function vb_sparse_pbicgstab
     function vb_pressure_update(ctx) bind(C, name="vb_pressure_update") result(s)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: s
     end function vb_pressure_update
     function vb_final_advection(ctx) bind(C, name="vb_final_advection") result(s)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: s
     end function vb_final_advection
     function vb_div_free(ctx) bind(C, name="vb_div_free") result(s)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: s
     end function vb_div_free

     !! ---- scalars and exports ----
     function vb_get_scalars(ctx, ints12, dbls8) bind(C, name="vb_get_scalars") result(s)
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(out) :: ints12(12)
       real(c_double), intent(out) :: dbls8(8)
       integer(c_int) :: s
     end function vb_get_scalars
     function vb_export_mirrors(ctx, m, irelation, vrelation, dp_mp) bind(C, name="vb_export_mirrors") result(s)
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: m
       integer(c_int32_t), intent(out) :: irelation(*), vrelation(*)
       real(c_double), intent(out) :: dp_mp(*)
       integer(c_int) :: s
     end function vb_export_mirrors
     function vb_export_cells(ctx, np, cellpart, nct1, ist, ip) bind(C, name="vb_export_cells") result(s)
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: np, nct1
       integer(c_int32_t), intent(out) :: cellpart(*), ist(*), ip(*)
       integer(c_int) :: s
     end function vb_export_cells
     function vb_export_neighbours(ctx, n, w, ij_count, ij_link) bind(C, name="vb_export_neighbours") result(s)
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: n, w
       integer(c_int32_t), intent(out) :: ij_count(*), ij_link(*)
       integer(c_int) :: s
     end function vb_export_neighbours
     function vb_export_weights(ctx, n, w, w_g, w_l, kgcm) bind(C, name="vb_export_weights") result(s)
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: n, w
       real(c_double), intent(out) :: w_g(*), w_l(*), kgcm(*)
       integer(c_int) :: s
     end function vb_export_weights
     function vb_export_ushift(ctx, n, ushift) bind(C, name="vb_export_ushift") result(s)
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: n
       real(c_double), intent(out) :: ushift(*)
       integer(c_int) :: s
     end function vb_export_ushift
     function vb_export_grad_p(ctx, n, grad_p) bind(C, name="vb_export_grad_p") result(s)
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: n
       real(c_double), intent(out) :: grad_p(*)
       integer(c_int) :: s
     end function vb_export_grad_p
     function vb_export_ppe(ctx, n, w, diag, col, val, rhs) bind(C, name="vb_export_ppe") result(s)
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: n, w
       real(c_double), intent(out) :: diag(*), val(*), rhs(*)
       integer(c_int32_t), intent(out) :: col(*)
       integer(c_int) :: s
     end function vb_export_ppe

     !! ---- output quantities of the driver ----
     function vb_secondary_properties(ctx, n, vort, u_rms) bind(C, name="vb_secondary_properties") result(s)
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: n
       real(c_double), intent(out) :: vort(*), u_rms
       integer(c_int) :: s
     end function vb_secondary_properties
     !! the assembled PPE as an operator (y = A x, or A^T x with transpose = 1), for an adjoint solve
     function vb_ppe_operator(ctx, transpose, n, x, y) bind(C, name="vb_ppe_operator") result(s)
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: transpose, n
       real(c_double), intent(in) :: x(*)
       real(c_double), intent(out) :: y(*)
       integer(c_int) :: s
     end function vb_ppe_operator
     !! the SPH pressure gradient as an operator (x: n values -> y: 2n) and its transpose (2n -> n)
     function vb_grad_operator(ctx, transpose, n, x, y) bind(C, name="vb_grad_operator") result(s)
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: transpose, n
       real(c_double), intent(in) :: x(*)
       real(c_double), intent(out) :: y(*)
       integer(c_int) :: s
     end function vb_grad_operator
     !! A^T lambda = g with the reference's BiCGSTAB on the transposed PPE (an adjoint solve)
     function vb_ppe_adjoint_solve(ctx, n, g, lambda, tol, itermax, iters, residual) &
          bind(C, name="vb_ppe_adjoint_solve") result(s)
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: n, itermax
       real(c_double), intent(in) :: g(*)
       real(c_double), intent(out) :: lambda(*)
       real(c_double), value :: tol
       integer(c_int32_t), intent(out) :: iters
       real(c_double), intent(out) :: residual
       integer(c_int) :: s
     end function vb_ppe_adjoint_solve
     function vb_free_surface(ctx, n, n_surf, conc) bind(C, name="vb_free_surface") result(s)
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: n
       integer(c_int32_t), intent(out) :: n_surf(*)
       real(c_double), intent(out) :: conc(*)
       integer(c_int) :: s
     end function vb_free_surface
     function vb_field_maxima(ctx, maxima) bind(C, name="vb_field_maxima") result(s)
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: maxima(3)
       integer(c_int) :: s
     end function vb_field_maxima

     !! ---- run diagnostics (no reference counterpart) ----
     function vb_circle_forces(ctx, nc, force, counts) bind(C, name="vb_circle_forces") result(s)
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: nc
       real(c_double), intent(out) :: force(*)
       integer(c_int32_t), intent(out) :: counts(*)
       integer(c_int) :: s
     end function vb_circle_forces
     function vb_nusselt(ctx, nu) bind(C, name="vb_nusselt") result(s)
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: nu
       integer(c_int) :: s
     end function vb_nusselt

     !! ---- measurement ----
     function vb_profile_spmv(ctx, reps, avg_us, alg_bytes) bind(C, name="vb_profile_spmv") result(s)
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: reps
       real(c_double), intent(out) :: avg_us, alg_bytes
       integer(c_int) :: s
     end function vb_profile_spmv
     function vb_profile_copy(ctx, bytes, reps, avg_us) bind(C, name="vb_profile_copy") result(s)
       import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: bytes
       integer(c_int32_t), value :: reps
       real(c_double), intent(out) :: avg_us
       integer(c_int) :: s
     end function vb_profile_copy

     !! ---- stand-alone kernels on device pointers ----
     function vb_nnf(model, n, d_sr, mu, d_mu_eff) bind(C, name="vb_nnf") result(s)
       import :: c_int, c_int32_t, c_double, c_ptr
       integer(c_int32_t), value :: model, n
       type(c_ptr), value :: d_sr, d_mu_eff
       real(c_double), value :: mu
       integer(c_int) :: s
     end function vb_nnf
     function vb_kernel_functions(n, d_q, ad_7, ad_7h, d_fac, d_wab) bind(C, name="vb_kernel_functions") result(s)
       import :: c_int, c_int32_t, c_double, c_ptr
       integer(c_int32_t), value :: n
       type(c_ptr), value :: d_q, d_fac, d_wab
       real(c_double), value :: ad_7, ad_7h
       integer(c_int) :: s
     end function vb_kernel_functions
     function vb_logconf(n, d_tau, d_trace, d_gradvel, dt, wi, eps_ptt, kick, d_tau_out, d_trace_out) &
          bind(C, name="vb_logconf") result(s)
       import :: c_int, c_int32_t, c_double, c_ptr
       integer(c_int32_t), value :: n, kick
       type(c_ptr), value :: d_tau, d_trace, d_gradvel, d_tau_out, d_trace_out
       real(c_double), value :: dt, wi, eps_ptt
       integer(c_int) :: s
     end function vb_logconf
     function vb_adjoint_record(ctx, nsteps, ckpt_every, tol, itermax, ppe_iters) &
          bind(C, name="vb_adjoint_record") result(s)
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx, ppe_iters
       integer(c_int32_t), value :: nsteps, ckpt_every, itermax
       real(c_double), value :: tol
       integer(c_int) :: s
     end function vb_adjoint_record
     function vb_adjoint_backward(ctx, n, x_bar, y_bar, u_bar, v_bar, theta_bar, tau11_bar, tau12_bar, &
          tau21_bar, tau22_bar, trace_bar, tol, itermax, grads3, x_bar0, y_bar0, u_bar0, v_bar0, theta_bar0, &
          tau11_bar0, tau12_bar0, tau21_bar0, tau22_bar0, trace_bar0, adj_iters) &
          bind(C, name="vb_adjoint_backward") result(s)
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx, x_bar, y_bar, u_bar, v_bar, theta_bar, tau11_bar, tau12_bar, tau21_bar, &
            tau22_bar, trace_bar, x_bar0, y_bar0, u_bar0, v_bar0, theta_bar0, tau11_bar0, tau12_bar0, &
            tau21_bar0, tau22_bar0, trace_bar0, adj_iters
       integer(c_int32_t), value :: n, itermax
       real(c_double), value :: tol
       real(c_double), intent(out) :: grads3(3)
       integer(c_int) :: s
     end function vb_adjoint_backward
     function vb_logconf_noalloc(n, d_tau, d_trace, d_gradvel, dt, wi, eps_ptt, kick, d_tau_out, d_trace_out) &
          bind(C, name="vb_logconf_noalloc") result(s)
       import :: c_int, c_int32_t, c_double, c_ptr
       integer(c_int32_t), value :: n, kick
       type(c_ptr), value :: d_tau, d_trace, d_gradvel, d_tau_out, d_trace_out
       real(c_double), value :: dt, wi, eps_ptt
       integer(c_int) :: s
     end function vb_logconf_noalloc
     function vb_logconf_adjoint(n, d_tau, d_trace, d_gradvel, dt, wi, eps_ptt, kick, d_tau_out_bar, &
          d_trace_out_bar, d_tau_bar, d_trace_bar, d_gradvel_bar, d_wi_bar, d_dt_bar) &
          bind(C, name="vb_logconf_adjoint") result(s)
       import :: c_int, c_int32_t, c_double, c_ptr
       integer(c_int32_t), value :: n, kick
       type(c_ptr), value :: d_tau, d_trace, d_gradvel, d_tau_out_bar, d_trace_out_bar, d_tau_bar, &
            d_trace_bar, d_gradvel_bar, d_wi_bar, d_dt_bar
       real(c_double), value :: dt, wi, eps_ptt
       integer(c_int) :: s
     end function vb_logconf_adjoint
  end interface

contains

  !! vb_last_error as a Fortran string
  function vb_error_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: cs(:)
    type(c_ptr) :: p
    integer :: k
    interface
       function c_strlen(str) bind(C, name="strlen") result(l)
         import :: c_ptr, c_size_t
         type(c_ptr), value :: str
         integer(c_size_t) :: l
       end function c_strlen
    end interface
    p = vb_last_error(ctx)
    if (.not. c_associated(p)) then
       msg = ""
       return
    end if
    call c_f_pointer(p, cs, [int(c_strlen(p))])
    allocate(character(len=size(cs)) :: msg)
    do k = 1, size(cs)
       msg(k:k) = cs(k)
    end do
  end function vb_error_message

end module vebisph_gpu


module divfree_gpu
  use iso_c_binding
  use vebisph_gpu
  implicit none

  private
  public :: div_free, vbf_init, vbf_finalize, vbf_upload, vbf_download, vbf_scalars
  public :: vbf_ctx, vbf_status, vbf_stop_on_error, LS_iters, LS_residual, dt, time, itime

  type(c_ptr), save :: vbf_ctx = c_null_ptr
  integer(c_int), save :: vbf_status = VB_OK
  logical, save :: vbf_stop_on_error = .true.     !! STOP like the reference on an error status
  !! the scalars the reference keeps in common_2d after a step
  integer(c_int32_t), save :: LS_iters = 0, itime = 0
  real(c_double), save :: LS_residual = 0.0d0, dt = 0.0d0, time = 0.0d0

contains

  !! Error statuses stop the run as the reference's STOP does (neighbour_finding_mod.F90:82-103,
  !! put_in_bin.F90:51-54); the BiCGSTAB warnings are kept in vbf_status and the run goes on.
  logical function vbf_ok(s, what)
    integer(c_int), intent(in) :: s
    character(len=*), intent(in) :: what
    vbf_status = s
    vbf_ok = (s == VB_OK .or. s == VB_WARN_BICGSTAB_BREAKDOWN .or. s == VB_WARN_MAXITER)
    if (s == VB_WARN_BICGSTAB_BREAKDOWN) write(6, *) 'BICGSTAB Fails'
    if (.not. vbf_ok) then
       write(6, *) 'vebisph_gpu: ', what, ' failed with status ', s, ': ', vb_error_message(vbf_ctx)
       if (vbf_stop_on_error) stop
    end if
  end function vbf_ok

  subroutine vbf_init(prm, bnd, npfb, nbs)
    type(vb_params), intent(in) :: prm
    type(vb_boundary), intent(in) :: bnd
    integer(c_int32_t), intent(in) :: npfb, nbs
    if (c_associated(vbf_ctx)) call vbf_finalize()
    if (.not. vbf_ok(vb_create(prm, bnd, npfb, nbs, vbf_ctx), 'vb_create')) return
    time = prm%time
    itime = prm%itime
  end subroutine vbf_init

  subroutine vbf_finalize()
    integer(c_int) :: s
    if (c_associated(vbf_ctx)) s = vb_destroy(vbf_ctx)
    vbf_ctx = c_null_ptr
  end subroutine vbf_finalize

  !! Host arrays in the reference's shapes (common_2d.F90:49-61): rows 1..n are sent.
  subroutine vbf_upload(n, r, u, theta, tau_p, trace_conf, p, inbin)
    integer(c_int32_t), intent(in) :: n
    real(c_double), intent(in), contiguous :: r(:,:), u(:,:), theta(:), tau_p(:,:,:), trace_conf(:), p(:)
    logical(c_bool), intent(in) :: inbin(:)
    integer(c_int32_t), allocatable :: ib(:)
    allocate(ib(n))
    ib = merge(1_c_int32_t, 0_c_int32_t, inbin(1:n))
    if (.not. vbf_ok(vb_upload_state(vbf_ctx, n, r(:,1), r(:,2), u(:,1), u(:,2), theta, &
         tau_p(:,1,1), tau_p(:,1,2), tau_p(:,2,1), tau_p(:,2,2), trace_conf, p, ib), 'vb_upload_state')) return
  end subroutine vbf_upload

  subroutine vbf_download(n, r, u, theta, tau_p, trace_conf, p, inbin)
    integer(c_int32_t), intent(in) :: n
    real(c_double), intent(inout), contiguous :: r(:,:), u(:,:), theta(:), tau_p(:,:,:), trace_conf(:), p(:)
    logical(c_bool), intent(inout) :: inbin(:)
    integer(c_int32_t), allocatable :: ib(:)
    allocate(ib(n))
    if (.not. vbf_ok(vb_download_state(vbf_ctx, n, r(:,1), r(:,2), u(:,1), u(:,2), theta, &
         tau_p(:,1,1), tau_p(:,1,2), tau_p(:,2,1), tau_p(:,2,2), trace_conf, p, ib), 'vb_download_state')) return
    inbin(1:n) = ib /= 0
  end subroutine vbf_download

  !! LS_iters, LS_residual, dt, time, itime of the last step (one synchronisation)
  subroutine vbf_scalars()
    integer(c_int32_t) :: iv(12)
    real(c_double) :: dv(8)
    if (.not. vbf_ok(vb_get_scalars(vbf_ctx, iv, dv), 'vb_get_scalars')) return
    LS_iters = iv(6)
    itime = iv(7)
    dt = dv(1)
    LS_residual = dv(2)
    time = dv(3)
  end subroutine vbf_scalars

  !! One time step: the call sequence of div_free.F90:38-122.  The padded box (x_mint..y_maxt of
  !! sph2D_incom.F90:64-67) is computed by the context from the boundary description; the
  !! arguments are kept for the reference's call signature.
  subroutine div_free(x_mint, y_mint, x_maxt, y_maxt)
    real(c_double) :: x_mint, x_maxt, y_mint, y_maxt
    integer(c_int) :: ls
    if (.not. vbf_ok(vb_set_time_step(vbf_ctx), 'set_time_step')) return                !! :39-53
    if (.not. vbf_ok(vb_put_in_bin(vbf_ctx), 'put_in_bin')) return                       !! :57
    if (.not. vbf_ok(vb_create_mirror_particles(vbf_ctx), 'create_mirror_particles')) return
    if (.not. vbf_ok(vb_particle_allocation(vbf_ctx), 'particle_allocation')) return
    if (.not. vbf_ok(vb_kernel_calculation(vbf_ctx), 'kernel_calculation')) return
    if (.not. vbf_ok(vb_locate_free_surf_part(vbf_ctx), 'locate_free_surf_part')) return
    if (.not. vbf_ok(vb_new_fick_shift(vbf_ctx), 'new_fick_shift')) return             !! :65-78
    if (.not. vbf_ok(vb_prediction_step(vbf_ctx), 'prediction_step')) return           !! :82
    if (.not. vbf_ok(vb_mirror_velocities(vbf_ctx), 'mirror_velocities')) return       !! :85
    ls = vb_ppe_solve(vbf_ctx, 1.0d-8, 1000_c_int32_t, c_null_ptr, c_null_ptr)       !! :86
    if (.not. vbf_ok(ls, 'ppe_solve_in_house_simple_dirichlet')) return
    if (.not. vbf_ok(vb_pressure_update(vbf_ctx), 'calc_pressure_gradient')) return    !! :93-102
    if (.not. vbf_ok(vb_final_advection(vbf_ctx), 'final_advection')) return          !! :111-117
    call vbf_scalars()
    if (vbf_status == VB_OK) vbf_status = ls
  end subroutine div_free

end module divfree_gpu


!! C-callable driver: nsteps of div_free from state arrays in per-component columns (length npfb).
!! The state is copied into arrays shaped as the reference's (r(npfb,2), u(npfb,2),
!! tau_p(npfb,2,2), logical inbin) and moved through vbf_upload / vbf_download.  Returns the status
!! of the last step (errors do not STOP here), LS_iters of the last step and the time reached.
function vbf_drive(prm, bnd, npfb, nbs, nsteps, x, y, u, v, theta, t11, t12, t21, t22, trc, p, &
     inbin, ls_iters_out, time_out) bind(C, name="vbf_drive") result(status)
  use iso_c_binding
  use vebisph_gpu
  use divfree_gpu
  implicit none
  type(vb_params), intent(in) :: prm
  type(vb_boundary), intent(in) :: bnd
  integer(c_int32_t), value :: npfb, nbs, nsteps
  real(c_double), intent(inout) :: x(npfb), y(npfb), u(npfb), v(npfb), theta(npfb), t11(npfb), &
       t12(npfb), t21(npfb), t22(npfb), trc(npfb), p(npfb)
  integer(c_int32_t), intent(inout) :: inbin(npfb)
  integer(c_int32_t), intent(out) :: ls_iters_out
  real(c_double), intent(out) :: time_out
  integer(c_int) :: status
  real(c_double), allocatable :: r_h(:,:), u_h(:,:), tau_h(:,:,:)
  logical(c_bool), allocatable :: inbin_h(:)
  real(c_double) :: x_mint, x_maxt, y_mint, y_maxt
  integer :: k

  vbf_stop_on_error = .false.
  allocate(r_h(npfb,2), u_h(npfb,2), tau_h(npfb,2,2), inbin_h(npfb))
  r_h(:,1) = x; r_h(:,2) = y
  u_h(:,1) = u; u_h(:,2) = v
  tau_h(:,1,1) = t11; tau_h(:,1,2) = t12; tau_h(:,2,1) = t21; tau_h(:,2,2) = t22
  inbin_h = inbin /= 0
  x_mint = 0.0d0; x_maxt = 0.0d0; y_mint = 0.0d0; y_maxt = 0.0d0
  ls_iters_out = 0
  time_out = 0.0d0

  call vbf_init(prm, bnd, npfb, nbs)
  status = vbf_status
  if (status /= VB_OK) return
  call vbf_upload(npfb, r_h, u_h, theta, tau_h, trc, p, inbin_h)
  status = vbf_status
  if (status /= VB_OK) then
     call vbf_finalize()
     return
  end if
  do k = 1, nsteps
     call div_free(x_mint, y_mint, x_maxt, y_maxt)
     status = vbf_status
     if (status /= VB_OK .and. status /= VB_WARN_BICGSTAB_BREAKDOWN .and. status /= VB_WARN_MAXITER) exit
  end do
  if (status == VB_OK .or. status == VB_WARN_BICGSTAB_BREAKDOWN .or. status == VB_WARN_MAXITER) then
     call vbf_download(npfb, r_h, u_h, theta, tau_h, trc, p, inbin_h)
     if (vbf_status /= VB_OK) status = vbf_status
     x = r_h(:,1); y = r_h(:,2)
     u = u_h(:,1); v = u_h(:,2)
     t11 = tau_h(:,1,1); t12 = tau_h(:,1,2); t21 = tau_h(:,2,1); t22 = tau_h(:,2,2)
     inbin = merge(1_c_int32_t, 0_c_int32_t, inbin_h)
     ls_iters_out = LS_iters
     time_out = time
  end if
  call vbf_finalize()
end function vbf_drive
