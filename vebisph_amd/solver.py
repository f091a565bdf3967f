"""Host-side mirror of the reference's step interface over the gfx950 C ABI.

`Solver` keeps the reference's routine names (div_free.F90:22-124 and the subroutines it calls), so a
driver written against the reference reads the same:

    s = Solver(case)
    s.div_free()                         # one step (sph2D_incom.F90:97)
    # or phase by phase:
    s.set_time_step(); s.put_in_bin(); s.create_mirror_particles(); s.particle_allocation()
    s.kernel_calculation(); s.locate_free_surf_part(); s.new_fick_shift(); s.prediction_step()
    s.mirror_velocities(); s.ppe_solve_in_house_simple_dirichlet(); s.calc_pressure_gradient()
    s.final_advection()

Errors that STOP the reference raise VbError; the BiCGSTAB breakdown / iteration-limit messages of
the reference become warnings recorded in `last_warning`.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import NPLINK, VbBoundary, VbError, VbParams

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _i(a):
    return a.ctypes.data_as(_ip) if a is not None else None


def make_params(case, precision=_lib.PREC_FP64, reference_kick=True, time=0.0, itime=0,
                serial_reductions=False) -> VbParams:
    """vb_params of a Case (the INDAT values, input.F90:18-37)."""
    p = VbParams()
    p.h, p.dx, p.dv = case.h, case.dx, case.dv
    p.Ra, p.Pr, p.El, p.beta = case.Ra, case.Pr, case.El, case.beta
    p.body_force[0], p.body_force[1] = case.body_force
    p.dt_coef_cfl, p.dt_coef_visc = case.dt_coef_cfl, case.dt_coef_visc
    p.ve_nonlinearity = case.ve_nonlinearity
    p.time, p.itime = time, itime
    p.i_open_domain, p.i_shift_coeff = case.i_open_domain, case.i_shift_coeff
    p.precision = int(precision)
    p.reference_kick = 1 if reference_kick else 0
    p.serial_reductions = 1 if serial_reductions else 0
    p.fluid_model = getattr(case, "const_mod", 2)
    return p


def make_boundary(case, keep: list) -> VbBoundary:
    """vb_boundary of a Case (IBOUND, input.F90:51-81); the arrays it points to are appended to
    `keep`, which must outlive every call that reads the record."""
    def f64(a):
        keep.append(np.ascontiguousarray(a, dtype=np.float64))
        return _d(keep[-1])

    def i32(a):
        keep.append(np.ascontiguousarray(a, dtype=np.int32))
        return _i(keep[-1])
    b = VbBoundary()
    b.nb_patches = case.nb_patches
    b.b_node, b.b_type = f64(case.b_node), i32(case.b_type)
    b.b_vel, b.b_thermal = f64(case.b_vel), f64(case.b_thermal)
    b.b_periodic_parent = i32(case.b_periodic_parent)
    b.nb_circles = case.nb_circles
    if case.nb_circles:
        b.c_centre, b.c_radius = f64(case.c_centre), f64(case.c_radius)
        b.c_omega, b.c_vel = f64(case.c_omega), f64(case.c_vel)
        b.c_thermal = f64(case.c_thermal)
    return b


def initial_state(case):
    """The state a case starts from: fluid at rest, conformation I (tau = 0, trace 2), p = 0."""
    n = case.npfb
    return dict(x=case.x, y=case.y, u=case.u, v=case.v, theta=case.theta, tau=np.zeros((n, 4)),
                trace=np.full(n, 2.0), p=np.zeros(n), inbin=np.zeros(n, np.int32))


STATE_COLUMNS = ("x", "y", "u", "v", "theta", "tau11", "tau12", "tau21", "tau22", "trace", "p")


def _state_columns(s):
    tau = np.asarray(s["tau"], dtype=np.float64)
    cols = {"x": s["x"], "y": s["y"], "u": s["u"], "v": s["v"], "theta": s["theta"],
            "tau11": tau[:, 0], "tau12": tau[:, 1], "tau21": tau[:, 2], "tau22": tau[:, 3],
            "trace": s["trace"], "p": s["p"]}
    return [np.ascontiguousarray(cols[k], dtype=np.float64) for k in STATE_COLUMNS]


def write_case_file(path, case, precision=_lib.PREC_FP64, state=None):
    """The case file read by the C++ driver vb_drive (layout in vebisph_amd/cpp/vb_drive.cpp)."""
    p = make_params(case, precision)
    s = initial_state(case) if state is None else state
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64).tobytes()
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32).tobytes()
    with open(path, "wb") as f:
        f.write(b"VBC1")
        f.write(i32([case.npfb, case.nbS, case.nb_patches, case.nb_circles]))
        f.write(bytes(p))
        f.write(f64(case.b_node) + i32(case.b_type) + f64(case.b_vel) + f64(case.b_thermal)
                + i32(case.b_periodic_parent))
        if case.nb_circles:
            f.write(f64(case.c_centre) + f64(case.c_radius) + f64(case.c_omega) + f64(case.c_vel)
                    + f64(case.c_thermal))
        for c in _state_columns(s):
            f.write(c.tobytes())
        f.write(i32(s["inbin"]))


def read_state_file(path, n):
    """The state file written by vb_drive, as a dict like Solver.get_state()."""
    raw = open(path, "rb").read()
    cols = np.frombuffer(raw[:88 * n], dtype=np.float64).reshape(11, n)
    d = dict(zip(STATE_COLUMNS, cols))
    out = {k: d[k].copy() for k in ("x", "y", "u", "v", "theta", "trace", "p")}
    out["tau"] = np.stack([d["tau11"], d["tau12"], d["tau21"], d["tau22"]], axis=1)
    out["inbin"] = np.frombuffer(raw[88 * n:88 * n + 4 * n], dtype=np.int32).copy()
    return out


class Solver:
    """One gfx950 context holding a Case's particle state in HBM."""

    def __init__(self, case, precision=_lib.PREC_FP64, reference_kick=True, time=0.0, itime=0,
                 state=None, serial_reductions=False):
        L = _lib.lib()
        self.case = case
        p = make_params(case, precision, reference_kick, time, itime, serial_reductions)
        self._keep = []
        b = make_boundary(case, self._keep)
        h = C.c_void_p()
        st = L.vb_create(C.byref(p), C.byref(b), case.npfb, case.nbS, C.byref(h))
        if st != 0:
            raise VbError(st, "vb_create failed")
        self.h = h
        self.last_warning = 0
        self.set_state(initial_state(case) if state is None else state)

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().vb_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, st):
        if st in _lib.WARNINGS:
            self.last_warning = st
            return st
        if st != 0:
            msg = _lib.lib().vb_last_error(self.h)
            raise VbError(st, msg.decode() if msg else "")
        return st

    # ---- state ----
    def set_state(self, s):
        n = self.case.npfb
        arrs = _state_columns(s)
        inb = np.ascontiguousarray(s["inbin"], dtype=np.int32)
        self._chk(_lib.lib().vb_upload_state(self.h, n, *[_d(a) for a in arrs], _i(inb)))

    def get_state(self, n=None):
        n = self.case.npfb if n is None else n
        out = {k: np.zeros(n) for k in ("x", "y", "u", "v", "theta", "trace", "p")}
        taus = [np.zeros(n) for _ in range(4)]
        inb = np.zeros(n, np.int32)
        self._chk(_lib.lib().vb_download_state(
            self.h, n, _d(out["x"]), _d(out["y"]), _d(out["u"]), _d(out["v"]), _d(out["theta"]),
            *[_d(t) for t in taus], _d(out["trace"]), _d(out["p"]), _i(inb)))
        out["tau"] = np.stack(taus, axis=1)
        out["inbin"] = inb
        return out

    def scalars(self):
        ints = np.zeros(12, np.int32); dbls = np.zeros(8)
        self._chk(_lib.lib().vb_get_scalars(self.h, _i(ints), _d(dbls)))
        keys_i = ("np", "npfb", "nbS", "nmirror", "maxneighbours", "LS_iters", "itime", "n_inbin",
                  "ncx", "ncy", "nct", "status")
        keys_d = ("dt", "LS_residual", "time", "umax", "x_mint", "x_maxt", "y_mint", "y_maxt")
        d = {k: int(v) for k, v in zip(keys_i, ints)}
        d.update({k: float(v) for k, v in zip(keys_d, dbls)})
        return d

    # ---- the step (reference names) ----
    def set_time_step(self):
        return self._chk(_lib.lib().vb_set_time_step(self.h))

    def put_in_bin(self):
        return self._chk(_lib.lib().vb_put_in_bin(self.h))

    def create_mirror_particles(self):
        return self._chk(_lib.lib().vb_create_mirror_particles(self.h))

    def particle_allocation(self):
        return self._chk(_lib.lib().vb_particle_allocation(self.h))

    def kernel_calculation(self):
        return self._chk(_lib.lib().vb_kernel_calculation(self.h))

    def locate_free_surf_part(self):
        return self._chk(_lib.lib().vb_locate_free_surf_part(self.h))

    def new_fick_shift(self):
        return self._chk(_lib.lib().vb_new_fick_shift(self.h))

    def prediction_step(self):
        return self._chk(_lib.lib().vb_prediction_step(self.h))

    def mirror_velocities(self):
        return self._chk(_lib.lib().vb_mirror_velocities(self.h))

    def ppe_solve_in_house_simple_dirichlet(self, tol=1e-8, itermax=1000):
        it = C.c_int32(0); res = C.c_double(0.0)
        self._chk(_lib.lib().vb_ppe_solve(self.h, tol, itermax, C.byref(it), C.byref(res)))
        return it.value, res.value

    def sparse_pbicgstab_nonull(self, a, ija, b, x0=None, itermax=1000, tol=1e-8, open_domain=False):
        """SPARSE_PBICGSTAB_nonull (linear_solver.F90:275) on a caller's NR row-indexed system, or with
        open_domain SPARSE_PBICGSTAB (:82, Jacobi, no mean removal)."""
        n = b.shape[0]
        x = np.ascontiguousarray(np.zeros(n) if x0 is None else x0, dtype=np.float64).copy()
        a = np.ascontiguousarray(a, dtype=np.float64); ija = np.ascontiguousarray(ija, dtype=np.int32)
        bb = np.ascontiguousarray(b, dtype=np.float64)
        it = C.c_int32(0); res = C.c_double(0.0)
        fn = _lib.lib().vb_sparse_pbicgstab if open_domain else _lib.lib().vb_sparse_pbicgstab_nonull
        st = self._chk(fn(self.h, _d(a), _i(ija), _d(x), _d(bb), n, a.shape[0], itermax, tol, C.byref(it),
                          C.byref(res)))
        return x, it.value, res.value, st

    def calc_pressure_gradient(self):
        return self._chk(_lib.lib().vb_pressure_update(self.h))

    def final_advection(self):
        return self._chk(_lib.lib().vb_final_advection(self.h))

    def div_free(self):
        return self._chk(_lib.lib().vb_div_free(self.h))

    step = div_free

    def phase(self, k):
        """Phase k of the same 12-way split the reference harness uses."""
        fns = [self.set_time_step, self.put_in_bin, self.create_mirror_particles,
               self.particle_allocation, self.kernel_calculation, self.locate_free_surf_part,
               self.new_fick_shift, self.prediction_step, self.mirror_velocities,
               self.ppe_solve_in_house_simple_dirichlet, self.calc_pressure_gradient,
               self.final_advection]
        return fns[k - 1]()

    # ---- exports (parity checks) ----
    def mirrors(self):
        m = self.scalars()["nmirror"]
        irel = np.zeros(m, np.int32); vrel = np.zeros(m, np.int32); dp = np.zeros(m)
        self._chk(_lib.lib().vb_export_mirrors(self.h, m, _i(irel), _i(vrel), _d(dp)))
        return irel, vrel, dp

    def cells(self):
        """The step's cell lists, 1-based as the reference's divide leaves them: cellpart (np), ist
        (nct + 1), ip (np; 0 past the binned count)."""
        s = self.scalars()
        np_, nct = s["np"], s["nct"]
        cp = np.zeros(np_, np.int32); ist = np.zeros(nct + 1, np.int32); ip = np.zeros(np_, np.int32)
        self._chk(_lib.lib().vb_export_cells(self.h, np_, _i(cp), nct + 1, _i(ist), _i(ip)))
        return cp, ist, ip

    def neighbours(self, w=NPLINK):
        n = self.case.npfb
        cnt = np.zeros(n, np.int32); link = np.zeros(n * w, np.int32)
        self._chk(_lib.lib().vb_export_neighbours(self.h, n, w, _i(cnt), _i(link)))
        return cnt, link.reshape(w, n).T

    def weights(self, w):
        n = self.case.npfb
        wg = np.zeros(2 * w * n); wl = np.zeros(w * n); kg = np.zeros(4 * n)
        self._chk(_lib.lib().vb_export_weights(self.h, n, w, _d(wg), _d(wl), _d(kg)))
        return (wg.reshape(2, w, n).transpose(2, 1, 0), wl.reshape(w, n).T, kg.reshape(4, n).T)

    def ushift(self):
        n = self.case.npfb
        us = np.zeros(2 * n)
        self._chk(_lib.lib().vb_export_ushift(self.h, n, _d(us)))
        return us.reshape(2, n).T

    def grad_p(self):
        n = self.case.npfb
        g = np.zeros(2 * n)
        self._chk(_lib.lib().vb_export_grad_p(self.h, n, _d(g)))
        return g.reshape(2, n).T

    def secondary_properties(self):
        """secondary_properties (div_free.F90:148-192): the vorticity of rows 1..npfb and the RMS
        velocity, from the last step's velocities and weights."""
        n = self.case.npfb
        vort = np.zeros(n); u_rms = C.c_double(0.0)
        self._chk(_lib.lib().vb_secondary_properties(self.h, n, _d(vort), C.byref(u_rms)))
        return vort, u_rms.value

    def ppe_operator(self, x, transpose=False, solver_rows=False):
        """y = A x or A^T x with A the PPE the last solve assembled (the adjoint's operator); with
        solver_rows through the solvers' own SpMV rows (the adjoint solve's A^T rows, the forward FP64
        solve's compact rows)."""
        n = self.case.npfb
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(n)
        mode = (2 if transpose else 3) if solver_rows else (1 if transpose else 0)
        self._chk(_lib.lib().vb_ppe_operator(self.h, mode, n, _d(x), _d(y)))
        return y

    def grad_operator(self, x, transpose=False):
        """The SPH pressure-gradient operator of the step's lists (mirrors at their parent's value):
        G x as an (npfb, 2) array, or G^T g for g of shape (npfb, 2)."""
        n = self.case.npfb
        if transpose:
            xin = np.ascontiguousarray(np.asarray(x, dtype=np.float64).T).ravel()
            y = np.zeros(n)
        else:
            xin = np.ascontiguousarray(x, dtype=np.float64)
            y = np.zeros(2 * n)
        self._chk(_lib.lib().vb_grad_operator(self.h, 1 if transpose else 0, n, _d(xin), _d(y)))
        return y if transpose else y.reshape(2, n).T

    def ppe_adjoint_solve(self, g, tol=1e-8, itermax=1000):
        """lambda with A^T lambda = g (the reference's BiCGSTAB on the transposed PPE): (lambda, iters,
        residual, status)."""
        n = self.case.npfb
        g = np.ascontiguousarray(g, dtype=np.float64)
        lam = np.zeros(n)
        it = C.c_int32(0); res = C.c_double(0.0)
        st = self._chk(_lib.lib().vb_ppe_adjoint_solve(self.h, n, _d(g), _d(lam), C.c_double(tol), itermax,
                                                      C.byref(it), C.byref(res)))
        return lam, it.value, res.value, st

    # ---- the adjoint of a run of steps (no counterpart in the reference) ----
    ADJ_FIELDS = ("x", "y", "u", "v", "theta", "tau11", "tau12", "tau21", "tau22", "trace")

    def adjoint_record(self, nsteps, ckpt_every, tol=1e-8, itermax=1000):
        """nsteps steps from the current state, recorded for adjoint_backward (the state every
        ckpt_every steps and every step's pressure kept on the device).  Returns the PPE iteration
        counts and the status (VB_WARN_* when a solve stopped at itermax)."""
        it = np.zeros(nsteps, np.int32)
        st = self._chk(_lib.lib().vb_adjoint_record(self.h, nsteps, ckpt_every, C.c_double(tol), itermax,
                                                    _i(it)))
        self._adj_nsteps = nsteps
        return it, st

    def adjoint_backward(self, bars, tol=1e-8, itermax=1000, initial=False):
        """The reverse sweep for dL/d(final state) = bars (a dict over ADJ_FIELDS, missing = 0):
        dict(grads = [dL/dWi, dL/dbeta, dL/dRa], initial = dL/d(start state) when asked for,
        adj_iters = adjoint PPE iterations per step, status)."""
        n = self.case.npfb
        ins = [np.ascontiguousarray(bars[k], dtype=np.float64) if bars.get(k) is not None else None
               for k in self.ADJ_FIELDS]
        outs = [np.zeros(n) for _ in self.ADJ_FIELDS] if initial else [None] * len(self.ADJ_FIELDS)
        g = np.zeros(3)
        it = np.zeros(self._adj_nsteps, np.int32)
        st = self._chk(_lib.lib().vb_adjoint_backward(
            self.h, n, *[_d(a) for a in ins], C.c_double(tol), itermax, _d(g), *[_d(a) for a in outs], _i(it)))
        return {"grads": g, "initial": dict(zip(self.ADJ_FIELDS, outs)) if initial else None,
                "adj_iters": it, "status": st}

    def free_surface(self, n=None):
        """n_surf and conc of rows 1..n (the PART file columns, output.F90:47-49) of the last step."""
        n = self.case.npfb if n is None else n
        ns = np.zeros(n, np.int32); co = np.zeros(n)
        self._chk(_lib.lib().vb_free_surface(self.h, n, _i(ns), _d(co)))
        return ns, co

    def field_maxima(self):
        """max |u|, |v|, |p| over rows 1..npfb (the per-step line of data_directory/TIME)."""
        m = np.zeros(3)
        self._chk(_lib.lib().vb_field_maxima(self.h, _d(m)))
        return m

    def circle_forces(self):
        """Force of the fluid on each circle, (nc, 2), and the boundary particles found on each."""
        nc = self.case.nb_circles
        f = np.zeros(2 * max(nc, 1)); cnt = np.zeros(max(nc, 1), np.int32)
        self._chk(_lib.lib().vb_circle_forces(self.h, nc, _d(f), _i(cnt)))
        return f[:2 * nc].reshape(nc, 2), cnt[:nc]

    def nusselt(self):
        nu = C.c_double(0.0)
        self._chk(_lib.lib().vb_nusselt(self.h, C.byref(nu)))
        return nu.value

    def profile_spmv(self, reps=20):
        """Mean duration (us) and algorithmic bytes of one BiCGSTAB SpMV launch."""
        us = C.c_double(0.0); by = C.c_double(0.0)
        self._chk(_lib.lib().vb_profile_spmv(self.h, reps, C.byref(us), C.byref(by)))
        return us.value, by.value

    def profile_copy(self, nbytes=1 << 30, reps=10):
        """Mean duration (us) of a plain device copy of `nbytes` bytes (HBM streaming rate)."""
        us = C.c_double(0.0)
        self._chk(_lib.lib().vb_profile_copy(self.h, nbytes, reps, C.byref(us)))
        return us.value

    def ppe_ell(self, w):
        n = self.case.npfb
        diag = np.zeros(n); col = np.zeros(w * n, np.int32); val = np.zeros(w * n); rhs = np.zeros(n)
        self._chk(_lib.lib().vb_export_ppe(self.h, n, w, _d(diag), _i(col), _d(val), _d(rhs)))
        return diag, col.reshape(w, n).T, val.reshape(w, n).T, rhs


def ell_to_nr(diag, col, val, cnt):
    """Reference Numerical-Recipes row-indexed arrays (1-based ija) from the ELL export."""
    n = diag.shape[0]
    nnz = n + 2 + int(cnt.sum())
    a = np.zeros(nnz); ija = np.zeros(nnz, np.int32)
    a[:n] = diag
    pos = n + 2
    for i in range(n):
        ija[i] = pos
        m = int(cnt[i])
        a[pos - 1:pos - 1 + m] = val[i, :m]
        ija[pos - 1:pos - 1 + m] = col[i, :m]
        pos += m
    ija[n] = pos
    return a, ija
