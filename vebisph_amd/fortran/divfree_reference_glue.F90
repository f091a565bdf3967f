!! Module divfree for the reference's own build: compiled in place of source/base/div_free.F90 (whose
!! module has the same name and the same public div_free(x_mint,y_mint,x_maxt,y_maxt),
!! div_free.F90:1-22), it runs each step on the GPU through module divfree_gpu and keeps the
!! reference's module globals (common_2d.F90:20-80) as the driver sph2D_incom.F90 expects them:
!!
!!   first call      copies the run parameters, boundary description and particle state from
!!                   common_2d into a device context (vbf_init, vbf_upload);
!!   every call      one step on the device, then dt, np, nmirror, n_inbin, LS_iters, LS_residual;
!!   output steps    the particle state is copied back into common_2d when the driver will write
!!                   it (time + dt > next_dump, output.F90:17, or the run ends), so the state
!!                   crosses PCIe only for output (with n_surf and conc in an open domain); when
!!                   time > next_dump2 the vorticity and RMS
!!                   velocity of secondary_properties (div_free.F90:152-185) come from the device
!!                   into common_2d's vort and unit 192, as the reference's step computes them.
!! Build: add vebisph_amd/fortran/vebisph_gpu.F90 and this file to the reference's source list in
!! place of div_free.F90 and link with -L<lib> -lvebisph_hip (see INTEGRATION.md of this package).
module divfree
  use iso_c_binding
  use kind_parameters
  use common_parameter
  use common_2d
  use vebisph_gpu, only: vb_params, vb_boundary, vb_get_scalars, vb_secondary_properties, vb_free_surface, &
       VB_PREC_FP64, VB_PREC_MIXED
  use divfree_gpu, only: gpu_div_free => div_free, vbf_init, vbf_upload, vbf_download, vbf_ctx, &
       vbf_status
  implicit none

  private
  public :: div_free, vbf_precision

  integer(c_int32_t), save :: vbf_precision = VB_PREC_FP64   !! set before the first step for mixed
  logical, save :: started = .false.
  !! boundary arrays in the row-major layout of vb_boundary (kept for the context's lifetime)
  real(c_double), allocatable, target, save :: bnode_rm(:), bvel_c(:), bth_c(:), ccen_rm(:), crad_c(:), &
       comg_c(:), cvel_rm(:), cth_c(:)
  integer(c_int32_t), allocatable, target, save :: btype_c(:), bpp_c(:)
#if const_mod==1
  real(c_double), allocatable, save :: tau_nt(:,:,:), trace_nt(:)
#endif

contains

  subroutine div_free(x_mint, y_mint, x_maxt, y_maxt)
    real(rkind) :: x_mint, x_maxt, y_mint, y_maxt
    integer(c_int32_t) :: ivals(12)
    real(c_double) :: dvals(8), u_rms
    integer(c_int) :: s

    !! the flags div_free.F90:31-37 sets
    lagrangian = .true.
    external_forcing = .false.
    fick_shifting = .true.
    akinci = .false.
    output_everytime = .false.
    output_mirrors = .false.

    if (.not. started) call start()
    call gpu_div_free(x_mint, y_mint, x_maxt, y_maxt)
    s = vb_get_scalars(vbf_ctx, ivals, dvals)
    np = ivals(1); nmirror = ivals(4); maxneighbours = ivals(5); LS_iters = ivals(6); n_inbin = ivals(8)
    dt = dvals(1); LS_residual = dvals(2); umax = dvals(4)
    !! secondary_properties, div_free.F90:152-185 (time is still the step's start time here)
    if (time .gt. next_dump2 .or. output_everytime) then
       next_dump2 = next_dump2 + dt_out
       s = vb_secondary_properties(vbf_ctx, int(npfb, c_int32_t), vort, u_rms)
       write(192,*) time, u_rms
       flush(192)
    end if
    if (time + dt > next_dump .or. time + dt > tmax) then
#if const_mod==1
       call vbf_download(int(npfb, c_int32_t), r, u, theta, tau_nt, trace_nt, p, inbin)
#else
       call vbf_download(int(npfb, c_int32_t), r, u, theta, tau_p, trace_conf, p, inbin)
#endif
       !! the open domain's n_surf and conc columns of the PART files (freesurf_mod.F90:29-37, :95-100)
       if (i_open_domain .eq. 1) s = vb_free_surface(vbf_ctx, int(npfb, c_int32_t), n_surf, conc)
    end if
  end subroutine div_free

  subroutine start()
    type(vb_params) :: prm
    type(vb_boundary) :: bnd
    integer :: nb, nc
    prm%h = h; prm%dx = dx; prm%dv = dv
    prm%Ra = Ra; prm%Pr = Pr; prm%El = El; prm%beta = beta
    prm%body_force = body_force
    prm%dt_coef_cfl = dt_coef_cfl; prm%dt_coef_visc = dt_coef_visc
    prm%ve_nonlinearity = ve_nonlinearity
    prm%time = time
    prm%itime = int(itime, c_int32_t) - 1          !! the driver has already counted this step
    prm%i_open_domain = int(i_open_domain, c_int32_t)
    prm%i_shift_coeff = int(i_shift_coeff, c_int32_t)
    prm%precision = vbf_precision
    !! the fluid model is the reference build's compile-time switch (Makefile: -Dconst_mod)
#if const_mod==1
    prm%fluid_model = 1
#else
    prm%fluid_model = 2
#endif
    prm%reference_kick = 1
    prm%serial_reductions = 0

    nb = nb_patches
    allocate(bnode_rm(2*nb), bvel_c(nb), bth_c(nb), btype_c(nb), bpp_c(nb))
    bnode_rm(1:2*nb:2) = b_node(1:nb,1); bnode_rm(2:2*nb:2) = b_node(1:nb,2)
    bvel_c = b_vel(1:nb); bth_c = b_thermal(1:nb)
    btype_c = int(b_type(1:nb), c_int32_t); bpp_c = int(b_periodic_parent(1:nb), c_int32_t)
    bnd%nb_patches = int(nb, c_int32_t)
    bnd%b_node = c_loc(bnode_rm); bnd%b_type = c_loc(btype_c); bnd%b_vel = c_loc(bvel_c)
    bnd%b_thermal = c_loc(bth_c); bnd%b_periodic_parent = c_loc(bpp_c)
    nc = nb_circles
    bnd%nb_circles = int(nc, c_int32_t)
    bnd%c_centre = c_null_ptr; bnd%c_radius = c_null_ptr; bnd%c_omega = c_null_ptr
    bnd%c_vel = c_null_ptr; bnd%c_thermal = c_null_ptr
    if (nc > 0) then
       allocate(ccen_rm(2*nc), crad_c(nc), comg_c(nc), cvel_rm(2*nc), cth_c(nc))
       ccen_rm(1:2*nc:2) = c_centre(1:nc,1); ccen_rm(2:2*nc:2) = c_centre(1:nc,2)
       cvel_rm(1:2*nc:2) = c_vel(1:nc,1); cvel_rm(2:2*nc:2) = c_vel(1:nc,2)
       crad_c = c_radius(1:nc); comg_c = c_omega(1:nc); cth_c = c_thermal(1:nc)
       bnd%c_centre = c_loc(ccen_rm); bnd%c_radius = c_loc(crad_c); bnd%c_omega = c_loc(comg_c)
       bnd%c_vel = c_loc(cvel_rm); bnd%c_thermal = c_loc(cth_c)
    end if

    call vbf_init(prm, bnd, int(npfb, c_int32_t), int(nbS, c_int32_t))
#if const_mod==1
    !! the Newtonian build allocates no polymeric stress (input.F90:109-112): zero stand-ins
    allocate(tau_nt(size(r,1),2,2), trace_nt(size(r,1)))
    tau_nt = 0.0d0; trace_nt = 2.0d0
    call vbf_upload(int(npfb, c_int32_t), r, u, theta, tau_nt, trace_nt, p, inbin)
#else
    call vbf_upload(int(npfb, c_int32_t), r, u, theta, tau_p, trace_conf, p, inbin)
#endif
    started = .true.
  end subroutine start

end module divfree
