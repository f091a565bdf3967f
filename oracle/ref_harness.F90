!! ------------------------------------------------------------------------------------------------
!! Test-only harness around the UNMODIFIED reference sources (jrcking/vebisph, Fortran).
!!
!! This file is written for this repository; it contains no reference code. It is compiled together
!! with the reference's own module sources (read in place, never copied) into oracle/_ref/libvbref.so
!! by oracle/Makefile. It exposes bind(C) entry points that
!!   * set the reference's module globals the way source/base/input.F90:12-180 (getdata) and
!!     source/base/sph2D_incom.F90:63-75 do, without going through the text input files,
!!   * run one reference time step (divfree::div_free, source/base/div_free.F90:22-124), or the same
!!     step phase by phase through the reference's public routines, and
!!   * copy module arrays back out so tests can compare them with the C restatement and the GPU path.
!!
!! It is test infrastructure: only tests, the smoke check and the CPU-baseline leg of bench.py use it.
!! ------------------------------------------------------------------------------------------------
module vbref_harness
  use iso_c_binding
  use kind_parameters
  use common_parameter
  use common_2d
  use cell_griddin
  use divfree
  use neighbour_finding
  use mirror_boundaries
  use freesurf
  use part_shift
  use predictor_mod
  use calculation_gradient
  use linear_solver
  use nnf_models
  implicit none

  real(rkind), save :: hx_mint, hx_maxt, hy_mint, hy_maxt
  real(rkind), allocatable, save :: h_grad_p(:,:)
  integer(ikind), save :: cap = 0

contains

  !! prm: 1 h, 2 dx, 3 dv, 4 Ra, 5 Pr, 6 El, 7 beta, 8 body_force(1), 9 body_force(2),
  !!      10 dt_coef_cfl, 11 dt_coef_visc, 12 ve_nonlinearity, 13 time
  !! iprm: 1 npfb, 2 nbS, 3 i_open_domain, 4 i_shift_coeff, 5 capacity (rows), 6 itime
  subroutine vbref_setup(iprm, prm, nbp, bnode, btype, bvel, btherm, bpp, &
       ncirc, ccen, crad, comg, cvel, ctherm) bind(C, name="vbref_setup")
    integer(c_int), intent(in) :: iprm(6)
    real(c_double), intent(in) :: prm(13)
    integer(c_int), value :: nbp, ncirc
    real(c_double), intent(in) :: bnode(nbp,2), bvel(nbp), btherm(nbp)
    integer(c_int), intent(in) :: btype(nbp), bpp(nbp)
    real(c_double), intent(in) :: ccen(max(ncirc,1),2), crad(max(ncirc,1)), comg(max(ncirc,1))
    real(c_double), intent(in) :: cvel(max(ncirc,1),2), ctherm(max(ncirc,1))
    integer(ikind) :: i, ibm1, ibp1
    real(rkind) :: p_dist, tmp1

    call vbref_free()
    call vbref_set_threads(1)

    h = prm(1); dx = prm(2); dv = prm(3); Ra = prm(4); Pr = prm(5); El = prm(6); beta = prm(7)
    body_force(1) = prm(8); body_force(2) = prm(9)
    dt_coef_cfl = prm(10); dt_coef_visc = prm(11); ve_nonlinearity = prm(12); alpha_evss = 0.0d0
    time = prm(13)
    npfb = iprm(1); nbS = iprm(2); i_open_domain = iprm(3); i_shift_coeff = iprm(4)
    itime = iprm(6)
    np = npfb
    cap = max(iprm(5), npar)
    i_RESTART = .false.

    !! as input.F90:44-48
    nmirror_esti = int(0.5*npfb)
    if(npfb.le.500) nmirror_esti = 10*npfb
    npfb_old = 1
    maxneighbours_old = 1
    allocate(ij_w_G(1,1,1), ij_w_L(1,1), kgcm(1,1,1))

    !! boundaries, as input.F90:51-81 and the generator's edge vectors (datclass.F90:730-744)
    nb_patches = nbp
    allocate(b_node(nb_patches,2), b_edge(nb_patches,2), b_vel(nb_patches), b_thermal(nb_patches))
    allocate(b_type(nb_patches), b_periodic_parent(nb_patches))
    do i = 1, nb_patches
       b_node(i,:) = bnode(i,:)
       b_type(i) = btype(i); b_vel(i) = bvel(i); b_thermal(i) = btherm(i)
       b_periodic_parent(i) = bpp(i)
    end do
    do i = 1, nb_patches
       ibp1 = mod(i,nb_patches) + 1
       b_edge(i,:) = b_node(ibp1,:) - b_node(i,:)
    end do
    nb_circles = ncirc
    if(nb_circles.ne.0) then
       allocate(c_centre(4*nb_circles,2), c_radius(4*nb_circles), c_omega(4*nb_circles))
       allocate(c_vel(4*nb_circles,2), c_thermal(4*nb_circles))
       do i = 1, nb_circles
          c_centre(i,:) = ccen(i,:); c_radius(i) = crad(i); c_omega(i) = comg(i)
          c_vel(i,:) = cvel(i,:); c_thermal(i) = ctherm(i)
       end do
    end if
    !! corner angles, as input.F90:255-273
    allocate(b_corner_angle(nb_patches))
    do i = 1, nb_patches
       ibm1 = mod(i+nb_patches-2,nb_patches)+1
       ibp1 = mod(i,nb_patches)+1
       p_dist = b_edge(ibm1,2)*b_node(ibp1,1) - b_edge(ibm1,1)*b_node(ibp1,2) + &
            b_node(i,1)*b_node(ibm1,2) - b_node(i,2)*b_node(ibm1,1)
       p_dist = -1.0d0*p_dist/sqrt(dot_product(b_edge(ibm1,:),b_edge(ibm1,:)))
       tmp1 = sqrt(dot_product(b_edge(i,:),b_edge(i,:)))
       if(p_dist.le.0.0d0.and.abs(p_dist).gt.abs(tmp1))then
          p_dist = -tmp1
       else if(p_dist.gt.0.0d0.and.abs(p_dist).gt.abs(tmp1))then
          p_dist = tmp1
       end if
       b_corner_angle(i) = asin(p_dist/tmp1)
    end do
    xb_min = minval(b_node(:,1)); xb_max = maxval(b_node(:,1))
    yb_min = minval(b_node(:,2)); yb_max = maxval(b_node(:,2))

    !! particle arrays, as input.F90:105-114
    allocate(r(cap,dims), u(cap,dims))
    allocate(p(cap), conc(cap), a_out(cap), theta(cap))
    allocate(vort(cap))
    allocate(tau_p(cap,dims,dims))
    allocate(trace_conf(cap)); trace_conf = 2.0
    allocate(n_surf(cap), div_r(cap))
    allocate(inbin(cap))
    r = 0.0d0; u = 0.0d0; p = 0.0d0; theta = 0.0d0; tau_p = 0.0d0; a_out = 0.0d0
    conc = 0.0d0; div_r = 0.0d0
    !! solver outputs as a fresh program has them: SPARSE_PBICGSTAB* leave them unset when rho = 0
    !! at the first iteration (linear_solver.F90:138-141)
    LS_iters = 0; LS_residual = 0.0d0
    n_surf(:) = 0_ikind
    n_free = 0
    allocate(p_free(cap))
    inbin(:) = .false.

    !! derived constants, as input.F90:164-172
    h2 = h*h
    sup_size = 3.0d0*h
    sup_size_2 = 9.0d0*h2
    uno_sup_size = 1.0d0/(3.0d0*h)
    ad_7 = 7.0d0/478.0d0/pi/h/h
    ad_7h = ad_7/h
    eta2 = h*h*1.d-10

    !! padded box and cell grid, as sph2D_incom.F90:64-75
    hx_mint = xb_min - sup_size*1.5d0
    hx_maxt = xb_max + sup_size*1.5d0
    hy_mint = yb_min - sup_size*1.5d0
    hy_maxt = yb_max + sup_size*1.5d0
    call cell_generator(hx_mint,hx_maxt,hy_mint,hy_maxt,uno_sup_size,ncx,ncy,nct)
    ncell_m = nct
    nsheet_m = nct
    next_dump = 1.0d300; next_dump2 = 1.0d300
  end subroutine vbref_setup

  !! OpenMP thread count of the reference loops.  1 (set by vbref_setup) makes the reference
  !! deterministic: its reductions then sum in index order and its random_number draws
  !! (predictor_mod.F90:370, inside a parallel loop) go to particles in index order.
  subroutine vbref_set_threads(n) bind(C, name="vbref_set_threads")
    use omp_lib
    integer(c_int), value :: n
    call omp_set_num_threads(n)
  end subroutine vbref_set_threads

  subroutine vbref_free() bind(C, name="vbref_free")
    if(allocated(r)) deallocate(r)
    if(allocated(u)) deallocate(u)
    if(allocated(p)) deallocate(p)
    if(allocated(conc)) deallocate(conc)
    if(allocated(a_out)) deallocate(a_out)
    if(allocated(theta)) deallocate(theta)
    if(allocated(vort)) deallocate(vort)
    if(allocated(tau_p)) deallocate(tau_p)
    if(allocated(trace_conf)) deallocate(trace_conf)
    if(allocated(n_surf)) deallocate(n_surf)
    if(allocated(div_r)) deallocate(div_r)
    if(allocated(inbin)) deallocate(inbin)
    if(allocated(p_free)) deallocate(p_free)
    if(allocated(ij_w_G)) deallocate(ij_w_G)
    if(allocated(ij_w_L)) deallocate(ij_w_L)
    if(allocated(kgcm)) deallocate(kgcm)
    if(allocated(b_node)) deallocate(b_node)
    if(allocated(b_edge)) deallocate(b_edge)
    if(allocated(b_vel)) deallocate(b_vel)
    if(allocated(b_thermal)) deallocate(b_thermal)
    if(allocated(b_type)) deallocate(b_type)
    if(allocated(b_periodic_parent)) deallocate(b_periodic_parent)
    if(allocated(b_corner_angle)) deallocate(b_corner_angle)
    if(allocated(c_centre)) deallocate(c_centre)
    if(allocated(c_radius)) deallocate(c_radius)
    if(allocated(c_omega)) deallocate(c_omega)
    if(allocated(c_vel)) deallocate(c_vel)
    if(allocated(c_thermal)) deallocate(c_thermal)
    if(allocated(irelation)) deallocate(irelation)
    if(allocated(vrelation)) deallocate(vrelation)
    if(allocated(dP_mp)) deallocate(dP_mp)
    if(allocated(ij_count)) deallocate(ij_count)
    if(allocated(ij_link)) deallocate(ij_link)
    if(allocated(ushift)) deallocate(ushift)
    if(allocated(r0)) deallocate(r0)
    if(allocated(u0)) deallocate(u0)
    if(allocated(h_grad_p)) deallocate(h_grad_p)
    if(allocated(surf_norm)) deallocate(surf_norm)
    if(allocated(ppe_smooth)) deallocate(ppe_smooth)
  end subroutine vbref_free

  !! Particle state in and out (1..n rows). tau is (n,4) = tau11, tau12, tau21, tau22.
  subroutine vbref_set_particles(n, x, y, ux, uy, th, tau, trc, pr_in, inb) bind(C, name="vbref_set_particles")
    integer(c_int), value :: n
    real(c_double), intent(in) :: x(n), y(n), ux(n), uy(n), th(n), tau(n,4), trc(n), pr_in(n)
    integer(c_int), intent(in) :: inb(n)
    integer(ikind) :: i
    do i = 1, n
       r(i,1) = x(i); r(i,2) = y(i); u(i,1) = ux(i); u(i,2) = uy(i); theta(i) = th(i)
       tau_p(i,1,1) = tau(i,1); tau_p(i,1,2) = tau(i,2); tau_p(i,2,1) = tau(i,3); tau_p(i,2,2) = tau(i,4)
       trace_conf(i) = trc(i); p(i) = pr_in(i)
       inbin(i) = (inb(i).ne.0)
    end do
  end subroutine vbref_set_particles

  subroutine vbref_get_particles(n, x, y, ux, uy, th, tau, trc, pr_out, inb) bind(C, name="vbref_get_particles")
    integer(c_int), value :: n
    real(c_double), intent(out) :: x(n), y(n), ux(n), uy(n), th(n), tau(n,4), trc(n), pr_out(n)
    integer(c_int), intent(out) :: inb(n)
    integer(ikind) :: i
    do i = 1, n
       x(i) = r(i,1); y(i) = r(i,2); ux(i) = u(i,1); uy(i) = u(i,2); th(i) = theta(i)
       !! the Newtonian build's getdata allocates no polymeric stress (input.F90:109-112)
       if (allocated(tau_p)) then
          tau(i,1) = tau_p(i,1,1); tau(i,2) = tau_p(i,1,2); tau(i,3) = tau_p(i,2,1); tau(i,4) = tau_p(i,2,2)
       else
          tau(i,:) = 0.0d0
       end if
       trc(i) = 2.0d0
       if (allocated(trace_conf)) trc(i) = trace_conf(i)
       pr_out(i) = p(i)
       inb(i) = merge(1, 0, logical(inbin(i)))
    end do
  end subroutine vbref_get_particles

  !! ints: 1 np, 2 npfb, 3 nbS, 4 nmirror, 5 maxneighbours, 6 LS_iters, 7 itime, 8 n_inbin, 9 ncx, 10 ncy, 11 nct
  !! dbls: 1 dt, 2 LS_residual, 3 time, 4 umax, 5 x_mint, 6 x_maxt, 7 y_mint, 8 y_maxt
  subroutine vbref_get_scalars(ints, dbls) bind(C, name="vbref_get_scalars")
    integer(c_int), intent(out) :: ints(11)
    real(c_double), intent(out) :: dbls(8)
    ints = (/ np, npfb, nbS, nmirror, maxneighbours, LS_iters, itime, n_inbin, ncx, ncy, nct /)
    dbls = (/ dt, LS_residual, time, umax, hx_mint, hx_maxt, hy_mint, hy_maxt /)
  end subroutine vbref_get_scalars

  !! One complete reference time step (sph2D_incom.F90:91-98).
  subroutine vbref_step() bind(C, name="vbref_step")
    itime = itime + 1
    call div_free(hx_mint, hy_mint, hx_maxt, hy_maxt)
    time = time + dt
  end subroutine vbref_step

  !! The same step split into the phases of div_free.F90:38-122, so intermediate arrays can be read.
  !! Phase k runs the k-th block; calling 1..12 in order equals vbref_step (checked by the tests).
  subroutine vbref_phase(k) bind(C, name="vbref_phase")
    integer(c_int), value :: k
    integer(ikind) :: i
    real(rkind) :: u_abs, dt_visc, dt_CFL
    select case(k)
    case(1)   !! div_free.F90:39-53 (set_time_step restated from div_free.F90:126-146, a private routine)
       itime = itime + 1
       lagrangian = .true.; external_forcing = .false.; fick_shifting = .true.; akinci = .false.
       output_everytime = .false.; output_mirrors = .false.
       allocate(r0(2*npfb,dims), u0(2*npfb,dims))
       u0(1:npfb,:) = u(1:npfb,:)
       r0(1:npfb,:) = r(1:npfb,:)
       umax = 1.0d-10
       do i = nbS+1, npfb
          u_abs = sqrt(dot_product(u(i,:),u(i,:)))
          if(u_abs.gt.umax) umax = u_abs
       end do
       dt_CFL = dt_coef_cfl*h/umax
       dt_visc = dt_coef_visc*h*h*min(sqrt(Ra/Pr),sqrt(Ra*Pr))
       dt = min(dt_CFL, dt_visc)
       do i = nbS+1, npfb
          r(i,:) = r0(i,:) + u0(i,:)*dt
       end do
    case(2)
       call put_in_bin(hx_mint, hx_maxt, hy_mint, hy_maxt)
    case(3)
       call create_mirror_particles
    case(4)
       call particle_allocation(hx_mint, hx_maxt, hy_mint, hy_maxt, 0)
    case(5)
       call kernel_calculation
    case(6)
       call locate_free_surf_part
    case(7)   !! div_free.F90:65-78
       allocate(ushift(npar,dims))
       allocate(dr_fs(npfb,dims))
       call new_fick_shift
       ushift(1:npfb,:) = dr_fs(:,:)/dt
       deallocate(dr_fs)
       ushift(1:nbS,:) = -u(1:nbS,:)
    case(8)
       call prediction_step
    case(9)
       call mirror_velocities
    case(10)
       call ppe_solve_in_house_simple_dirichlet
    case(11)  !! div_free.F90:93-102
       if(allocated(h_grad_p)) deallocate(h_grad_p)
       allocate(h_grad_p(npfb,dims))
       call calc_pressure_gradient(h_grad_p)
       do i = nbS+1, npfb
          u(i,:) = u(i,:) + dt*(body_force(:) - h_grad_p(i,:))
          if(ij_count(i).le.3) u(i,:) = u0(i,:) + dt*body_force(:)
       end do
    case(12)  !! div_free.F90:111-122 (secondary_properties only writes output files)
       do i = nbS+1, npfb
          r(i,:) = r0(i,:) + 0.5d0*(u(i,:)+u0(i,:))*dt + ushift(i,:)*dt
       end do
       deallocate(irelation, dP_mp, vrelation, ij_count, ij_link, ushift)
       if(allocated(surf_norm)) deallocate(surf_norm)
       deallocate(r0, u0)
       time = time + dt
    end select
  end subroutine vbref_phase

  !! Mirror maps (valid after phase 3): rows npfb+1..np.
  subroutine vbref_get_mirrors(m, irel, vrel, dp) bind(C, name="vbref_get_mirrors")
    integer(c_int), value :: m
    integer(c_int), intent(out) :: irel(m), vrel(m)
    real(c_double), intent(out) :: dp(m)
    integer(ikind) :: j
    do j = 1, m
       irel(j) = irelation(npfb+j); vrel(j) = vrelation(npfb+j); dp(j) = dP_mp(npfb+j)
    end do
  end subroutine vbref_get_mirrors

  !! Neighbour lists (valid after phase 4): count(n), link(n,w) slot-major, 1-based particle indices.
  subroutine vbref_get_neighbours(n, w, cnt, link) bind(C, name="vbref_get_neighbours")
    integer(c_int), value :: n, w
    integer(c_int), intent(out) :: cnt(n), link(n,w)
    integer(ikind) :: i, k
    link = 0
    do i = 1, n
       cnt(i) = ij_count(i)
       do k = 1, min(ij_count(i), w)
          link(i,k) = ij_link(i,k)
       end do
    end do
  end subroutine vbref_get_neighbours

  !! Kernel weights (valid after phase 5): wg(n,w,2), wl(n,w), kg(n,4) = K11,K12,K21,K22.
  subroutine vbref_get_weights(n, w, wg, wl, kg) bind(C, name="vbref_get_weights")
    integer(c_int), value :: n, w
    real(c_double), intent(out) :: wg(n,w,2), wl(n,w), kg(n,4)
    integer(ikind) :: i, k
    wg = 0.0d0; wl = 0.0d0
    do i = 1, n
       do k = 1, min(ij_count(i), w)
          wg(i,k,1) = ij_w_G(i,k,1); wg(i,k,2) = ij_w_G(i,k,2); wl(i,k) = ij_w_L(i,k)
       end do
       kg(i,1) = kgcm(i,1,1); kg(i,2) = kgcm(i,1,2); kg(i,3) = kgcm(i,2,1); kg(i,4) = kgcm(i,2,2)
    end do
  end subroutine vbref_get_weights

  !! Shifting velocity (valid after phase 7).
  subroutine vbref_get_ushift(n, us) bind(C, name="vbref_get_ushift")
    integer(c_int), value :: n
    real(c_double), intent(out) :: us(n,2)
    us(:,:) = ushift(1:n,:)
  end subroutine vbref_get_ushift

  !! Pressure gradient (valid after phase 11).
  subroutine vbref_get_grad_p(n, gp) bind(C, name="vbref_get_grad_p")
    integer(c_int), value :: n
    real(c_double), intent(out) :: gp(n,2)
    gp(:,:) = h_grad_p(1:n,:)
  end subroutine vbref_get_grad_p

  !! The reference BiCGSTAB (linear_solver.F90:275-425) on a caller-supplied row-indexed matrix.
  subroutine vbref_bicgstab_nonull(n, nnz, a, ija_in, x, b, itermax, tol, iters, resid) &
       bind(C, name="vbref_bicgstab_nonull")
    integer(c_int), value :: n, nnz, itermax
    real(c_double), value :: tol
    real(c_double), intent(in) :: a(nnz), b(n)
    integer(c_int), intent(in) :: ija_in(nnz)
    real(c_double), intent(inout) :: x(n)
    integer(c_int), intent(out) :: iters
    real(c_double), intent(out) :: resid
    real(rkind), allocatable :: aa(:), bb(:), xx(:)
    integer(ikind), allocatable :: jj(:)
    integer(ikind) :: it
    real(rkind) :: rs
    allocate(aa(nnz), bb(n), xx(n), jj(nnz))
    aa = a; bb = b; xx = x; jj = ija_in
    rs = 0.0d0
    call SPARSE_PBICGSTAB_nonull(aa, jj, xx, bb, n, nnz, itermax, tol, it, rs)
    x = xx; iters = it; resid = rs
    deallocate(aa, bb, xx, jj)
  end subroutine vbref_bicgstab_nonull

  !! The open-domain BiCGSTAB (linear_solver.F90:82-227) on a caller-supplied row-indexed matrix.
  subroutine vbref_bicgstab_open(n, nnz, a, ija_in, x, b, itermax, tol, iters, resid) &
       bind(C, name="vbref_bicgstab_open")
    integer(c_int), value :: n, nnz, itermax
    real(c_double), value :: tol
    real(c_double), intent(in) :: a(nnz), b(n)
    integer(c_int), intent(in) :: ija_in(nnz)
    real(c_double), intent(inout) :: x(n)
    integer(c_int), intent(out) :: iters
    real(c_double), intent(out) :: resid
    real(rkind), allocatable :: aa(:), bb(:), xx(:)
    integer :: jj(nnz), nn, nm, itm, it
    real(rkind) :: rs, tl
    allocate(aa(nnz), bb(n), xx(n))
    aa = a; bb = b; xx = x; jj = ija_in
    nn = n; nm = nnz; itm = itermax; tl = tol
    rs = 0.0d0; it = 0
    call SPARSE_PBICGSTAB(aa, jj, xx, bb, nn, nm, itm, tl, it, rs)
    x = xx; iters = it; resid = rs
    deallocate(aa, bb, xx)
  end subroutine vbref_bicgstab_open

  !! Free-surface quantities of rows 1..n after locate_free_surf_part (freesurf_mod.F90:17-68):
  !! n_surf, conc, div_r, and in an open domain the smoothed surf_norm (n,2) and ppe_smooth while
  !! they are allocated (zeros otherwise).
  subroutine vbref_get_surface(n, ns, co, dr, sn, ps) bind(C, name="vbref_get_surface")
    integer(c_int), value :: n
    integer(c_int), intent(out) :: ns(n)
    real(c_double), intent(out) :: co(n), dr(n), sn(n,2), ps(n)
    ns(:) = n_surf(1:n); co(:) = conc(1:n); dr(:) = div_r(1:n)
    sn = 0.0d0; ps = 0.0d0
    if(allocated(surf_norm)) sn(:,:) = surf_norm(1:n,:)
    if(allocated(ppe_smooth)) ps(:) = ppe_smooth(1:n)
  end subroutine vbref_get_surface

  !! ---- the reference's own file reader and writer (input.F90, output.F90) ----

  !! getdata (input.F90:12-180) on INDAT, IBOUND and IPART (TEMP_REC.dat when INDAT says restart)
  !! in the current directory, then the padded box and cell grid of sph2D_incom.F90:64-75.  Arrays
  !! the reference allocates without setting on the closed-domain path (conc, a_out, vort) and
  !! never-set scalars (av_conc, LS_iters, LS_residual, n_inbin) are zeroed, so written files are
  !! reproducible.
  subroutine vbref_getdata() bind(C, name="vbref_getdata")
    use input, only: getdata
    call vbref_free()
    call vbref_set_threads(1)
    call getdata()
    conc = 0.0d0; a_out = 0.0d0; vort = 0.0d0
    av_conc = 0.0d0; LS_iters = 0; LS_residual = 0.0d0; n_inbin = 0
    cap = npar
    hx_mint = xb_min - sup_size*1.5d0
    hx_maxt = xb_max + sup_size*1.5d0
    hy_mint = yb_min - sup_size*1.5d0
    hy_maxt = yb_max + sup_size*1.5d0
    call cell_generator(hx_mint,hx_maxt,hy_mint,hy_maxt,uno_sup_size,ncx,ncy,nct)
    ncell_m = nct
    nsheet_m = nct
    next_dump = time; next_dump2 = time
    n_dump = 0
  end subroutine vbref_getdata

  !! INDAT values and counts as getdata left them.
  !! ints: 1 i_RESTART, 2 i_open_domain, 3 i_shift_coeff, 4 np, 5 npfb, 6 nbS, 7 itime, 8 nb_patches,
  !!       9 nb_circles, 10 n_dump
  !! dbls: 1 tmax, 2 dt_out, 3 h, 4 dx, 5 dt, 6-7 body_force, 8 Ra, 9 Pr, 10 El, 11 beta,
  !!       12 dt_coef_cfl, 13 dt_coef_visc, 14 ve_nonlinearity, 15 alpha_evss, 16 dv, 17 time
  subroutine vbref_get_run_params(ints, dbls) bind(C, name="vbref_get_run_params")
    integer(c_int), intent(out) :: ints(10)
    real(c_double), intent(out) :: dbls(17)
    ints = (/ merge(1, 0, i_RESTART), i_open_domain, i_shift_coeff, np, npfb, nbS, itime, nb_patches, &
              nb_circles, n_dump /)
    dbls = (/ tmax, dt_out, h, dx, dt, body_force(1), body_force(2), Ra, Pr, El, beta, dt_coef_cfl, &
              dt_coef_visc, ve_nonlinearity, alpha_evss, dv, time /)
  end subroutine vbref_get_run_params

  !! Boundary arrays as read (patch arrays (nbp,2) column-major; periodic parents only meaningful
  !! for type-2 patches; circles (ncirc[,2]))
  subroutine vbref_get_boundary(nbp, ncirc, bnode, bedge, btype, bvel, btherm, bpp, ccen, crad, comg, cvel, &
       ctherm) bind(C, name="vbref_get_boundary")
    integer(c_int), value :: nbp, ncirc
    real(c_double), intent(out) :: bnode(nbp,2), bedge(nbp,2), bvel(nbp), btherm(nbp)
    integer(c_int), intent(out) :: btype(nbp), bpp(nbp)
    real(c_double), intent(out) :: ccen(max(ncirc,1),2), crad(max(ncirc,1)), comg(max(ncirc,1))
    real(c_double), intent(out) :: cvel(max(ncirc,1),2), ctherm(max(ncirc,1))
    integer(ikind) :: i
    do i = 1, nbp
       bnode(i,:) = b_node(i,:); bedge(i,:) = b_edge(i,:); btype(i) = b_type(i)
       bvel(i) = b_vel(i); btherm(i) = b_thermal(i)
       bpp(i) = 0
       if(b_type(i).eq.2) bpp(i) = b_periodic_parent(i)
    end do
    do i = 1, ncirc
       ccen(i,:) = c_centre(i,:); crad(i) = c_radius(i); comg(i) = c_omega(i)
       cvel(i,:) = c_vel(i,:); ctherm(i) = c_thermal(i)
    end do
  end subroutine vbref_get_boundary

  !! next_dump and next_dump2 of the driver: output writes files when time > next_dump
  !! (output.F90:18), div_free computes the vorticity when time > next_dump2 (div_free.F90:154)
  subroutine vbref_set_dumps(nd, nd2, dtout) bind(C, name="vbref_set_dumps")
    real(c_double), value :: nd, nd2, dtout
    next_dump = nd; next_dump2 = nd2; dt_out = dtout
  end subroutine vbref_set_dumps

  subroutine vbref_get_vort(n, vo, a_o) bind(C, name="vbref_get_vort")
    integer(c_int), value :: n
    real(c_double), intent(out) :: vo(n), a_o(n)
    vo = vort(1:n)
    a_o = a_out(1:n)
  end subroutine vbref_get_vort

  !! The units of sph2D_incom.F90:52-61 (data_directory/ must exist in the current directory)
  subroutine vbref_open_units(restart) bind(C, name="vbref_open_units")
    integer(c_int), value :: restart
    if(restart.ne.0) then
       open(unit=92,file='./data_directory/TIME',status='unknown',access='append')
       open(unit=93,file='./data_directory/LS',status='unknown',access='append')
       open(unit=94,file='./data_directory/TIME_OUT',status='unknown',access='append')
    else
       open(unit=92,file='./data_directory/TIME')
       open(unit=93,file='./data_directory/LS')
       open(unit=94,file='./data_directory/TIME_OUT')
    end if
  end subroutine vbref_open_units

  subroutine vbref_close_units() bind(C, name="vbref_close_units")
    close(92); close(93); close(94); close(192)
  end subroutine vbref_close_units

  !! the reference's output (output.F90)
  subroutine vbref_output() bind(C, name="vbref_output")
    external :: output
    call output
  end subroutine vbref_output

  !! list-directed records of this compiler, for the writer's format tests: one value per record
  subroutine vbref_ld_write(path, plen, n, vals) bind(C, name="vbref_ld_write")
    integer(c_int), value :: plen, n
    character(kind=c_char), intent(in) :: path(plen)
    real(c_double), intent(in) :: vals(n)
    character(len=plen) :: fname
    integer(ikind) :: i
    do i = 1, plen
       fname(i:i) = path(i)
    end do
    open(unit=77, file=fname, status='replace')
    do i = 1, n
       write(77,*) vals(i)
    end do
    close(77)
  end subroutine vbref_ld_write

  !! The reference's non-Newtonian viscosity models (nnf_models_mod.F90:8-110, no call sites in the
  !! reference) on n shear rates: model 0 nnf_newtonian (mu_eff = its per-particle mu array, given
  !! here in sr), 1 nnf_power_law, 2 nnf_bingham, 3 nnf_cross, 4 nnf_carreau with viscosity mu.
  subroutine vbref_nnf(model, n, sr, mu, out) bind(C, name="vbref_nnf")
    integer(c_int), value :: model, n
    real(c_double), value :: mu
    real(c_double), intent(in) :: sr(n)
    real(c_double), intent(out) :: out(n)
    real(rkind), allocatable :: s(:), o(:)
    real(rkind) :: m
    integer(ikind) :: nn
    allocate(s(n), o(n))
    s = sr; nn = n; m = mu; o = 0.0d0
    select case (model)
    case (0)
       call nnf_newtonian(s, nn, o)
    case (1)
       call nnf_power_law(s, m, nn, o)
    case (2)
       call nnf_bingham(s, m, nn, o)
    case (3)
       call nnf_cross(s, m, nn, o)
    case (4)
       call nnf_carreau(s, m, nn, o)
    end select
    out = o
    deallocate(s, o)
  end subroutine vbref_nnf

end module vbref_harness
