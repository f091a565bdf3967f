"""ctypes wrapper for oracle/liboracle.so, the C restatement of the reference step (vb_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker for tests/, the smoke check and bench.py's CPU-baseline leg.
Same interface as refwrap.RefSolver so tests can run either against the same Case.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
NPLINK = 75

_d = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lib = None


def available() -> bool:
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run make -C oracle oracle)")
        L = C.CDLL(LIB_PATH)
        L.vbo_create.restype = C.c_void_p
        L.vbo_create.argtypes = [_i, _d, C.c_int, _d, _i, _d, _d, _i, C.c_int, _d, _d, _d, _d, _d]
        L.vbo_destroy.argtypes = [C.c_void_p]
        L.vbo_set_particles.argtypes = [C.c_void_p, C.c_int] + [_d] * 8 + [_i]
        L.vbo_get_particles.argtypes = [C.c_void_p, C.c_int] + [_d] * 8 + [_i]
        L.vbo_get_scalars.argtypes = [C.c_void_p, _i, _d]
        L.vbo_phase.argtypes = [C.c_void_p, C.c_int]
        L.vbo_step.argtypes = [C.c_void_p]
        L.vbo_get_mirrors.argtypes = [C.c_void_p, C.c_int, _i, _i, _d]
        L.vbo_get_neighbours.argtypes = [C.c_void_p, C.c_int, C.c_int, _i, _i]
        L.vbo_get_cells.argtypes = [C.c_void_p, C.c_int, _i, C.c_int, _i, _i]
        L.vbo_get_weights.argtypes = [C.c_void_p, C.c_int, C.c_int, _d, _d, _d]
        L.vbo_get_ushift.argtypes = [C.c_void_p, C.c_int, _d]
        L.vbo_get_grad_vel.argtypes = [C.c_void_p, C.c_int, _d]
        L.vbo_get_ppe_nnz.argtypes = [C.c_void_p]
        L.vbo_get_ppe.argtypes = [C.c_void_p, C.c_int, _d, _i, _d]
        L.vbo_get_grad_p.argtypes = [C.c_void_p, C.c_int, _d]
        L.vbo_secondary_properties.argtypes = [C.c_void_p, C.c_int, _d, C.POINTER(C.c_double)]
        L.vbo_circle_forces.argtypes = [C.c_void_p, C.c_int, _d, _i]
        L.vbo_nusselt.argtypes = [C.c_void_p]
        L.vbo_nusselt.restype = C.c_double
        L.vbo_fac_wab.argtypes = [C.c_int, _d, C.c_double, C.c_double, _d, _d]
        L.vbo_calc_eigens.argtypes = [C.c_int, _d, _d, _d]
        L.vbo_logconf.argtypes = [C.c_int, _d, _d, _d, C.c_double, C.c_double, C.c_double, C.c_int, _d, _d]
        L.vbo_bicgstab_nonull.argtypes = [C.c_int, C.c_int, _d, _i, _d, _d, C.c_int, C.c_double,
                                          C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.vbo_bicgstab_open.argtypes = L.vbo_bicgstab_nonull.argtypes
        L.vbo_get_surface.argtypes = [C.c_void_p, C.c_int, _i, _d, _d, _d, _d]
        L.vbo_spmv_nr.argtypes = [C.c_int, _d, _i, _d, _d]
        L.vbo_nnf.argtypes = [C.c_int, C.c_int, _d, C.c_double, _d]
        _lib = L
    return _lib


def _c(a, dt=np.float64):
    return np.ascontiguousarray(np.asarray(a, dtype=dt))


def initial_state(case):
    n = case.npfb
    return dict(x=case.x, y=case.y, u=case.u, v=case.v, theta=case.theta, tau=np.zeros((n, 4)),
                trace=np.full(n, 2.0), p=np.zeros(n), inbin=np.zeros(n, np.int32))


class OracleSolver:
    def __init__(self, case, time=0.0, itime=0, state=None):
        L = lib()
        self.case = case
        nc = case.nb_circles
        iprm = _c([case.npfb, case.nbS, case.i_open_domain, case.i_shift_coeff, 0, itime,
                   getattr(case, "const_mod", 2)], np.int32)
        prm = _c(list(case.params_vector()) + [time])
        if nc:
            cc, cr, co, cv, ct = (_c(case.c_centre), _c(case.c_radius), _c(case.c_omega),
                                  _c(case.c_vel), _c(case.c_thermal))
        else:
            cc, cr, co, cv, ct = (np.zeros(2), np.zeros(1), np.zeros(1), np.zeros(2), np.zeros(1))
        self.h = L.vbo_create(iprm, prm, case.nb_patches, _c(case.b_node), _c(case.b_type, np.int32),
                              _c(case.b_vel), _c(case.b_thermal), _c(case.b_periodic_parent, np.int32),
                              nc, cc, cr, co, cv, ct)
        self.set_state(initial_state(case) if state is None else state)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.vbo_destroy(self.h)
            self.h = None

    def set_state(self, s):
        n = self.case.npfb
        lib().vbo_set_particles(self.h, n, _c(s["x"]), _c(s["y"]), _c(s["u"]), _c(s["v"]),
                                _c(s["theta"]), _c(s["tau"]), _c(s["trace"]), _c(s["p"]),
                                _c(s["inbin"], np.int32))

    def get_state(self, n=None):
        n = self.case.npfb if n is None else n
        out = {k: np.zeros(n) for k in ("x", "y", "u", "v", "theta", "trace", "p")}
        tau = np.zeros((n, 4)); inb = np.zeros(n, np.int32)
        lib().vbo_get_particles(self.h, n, out["x"], out["y"], out["u"], out["v"], out["theta"], tau,
                                out["trace"], out["p"], inb)
        out["tau"] = tau; out["inbin"] = inb
        return out

    def scalars(self):
        ints = np.zeros(12, np.int32); dbls = np.zeros(8)
        lib().vbo_get_scalars(self.h, ints, dbls)
        keys_i = ("np", "npfb", "nbS", "nmirror", "maxneighbours", "LS_iters", "itime", "n_inbin",
                  "ncx", "ncy", "nct", "status")
        keys_d = ("dt", "LS_residual", "time", "umax", "x_mint", "x_maxt", "y_mint", "y_maxt")
        d = {k: int(v) for k, v in zip(keys_i, ints)}
        d.update({k: float(v) for k, v in zip(keys_d, dbls)})
        return d

    def step(self):
        return lib().vbo_step(self.h)

    def phase(self, k):
        return lib().vbo_phase(self.h, k)

    def mirrors(self):
        m = self.scalars()["nmirror"]
        irel = np.zeros(m, np.int32); vrel = np.zeros(m, np.int32); dp = np.zeros(m)
        lib().vbo_get_mirrors(self.h, m, irel, vrel, dp)
        return irel, vrel, dp

    def neighbours(self, w=NPLINK):
        n = self.case.npfb
        cnt = np.zeros(n, np.int32); link = np.zeros(n * w, np.int32)
        lib().vbo_get_neighbours(self.h, n, w, cnt, link)
        return cnt, link.reshape(w, n).T

    def cells(self):
        s = self.scalars()
        np_, nct = s["np"], s["nct"]
        cp = np.zeros(np_, np.int32); ist = np.zeros(nct + 1, np.int32); ip = np.zeros(np_, np.int32)
        lib().vbo_get_cells(self.h, np_, cp, nct + 1, ist, ip)
        return cp, ist, ip

    def weights(self, w):
        n = self.case.npfb
        wg = np.zeros(2 * w * n); wl = np.zeros(w * n); kg = np.zeros(4 * n)
        lib().vbo_get_weights(self.h, n, w, wg, wl, kg)
        return (wg.reshape(2, w, n).transpose(2, 1, 0), wl.reshape(w, n).T, kg.reshape(4, n).T)

    def ushift(self):
        n = self.case.npfb
        us = np.zeros(2 * n)
        lib().vbo_get_ushift(self.h, n, us)
        return us.reshape(2, n).T

    def grad_vel(self):
        n = self.case.npfb
        g = np.zeros(4 * n)
        lib().vbo_get_grad_vel(self.h, n, g)
        return g.reshape(n, 4)

    def ppe(self):
        nnz = lib().vbo_get_ppe_nnz(self.h)
        a = np.zeros(nnz); ija = np.zeros(nnz, np.int32); b = np.zeros(self.case.npfb)
        lib().vbo_get_ppe(self.h, nnz, a, ija, b)
        return a, ija, b

    def grad_p(self):
        n = self.case.npfb
        g = np.zeros(2 * n)
        lib().vbo_get_grad_p(self.h, n, g)
        return g.reshape(2, n).T

    def secondary_properties(self):
        """Vorticity of rows 1..npfb and RMS velocity (after phase 11, before phase 12)."""
        n = self.case.npfb
        vort = np.zeros(n); u_rms = C.c_double(0.0)
        lib().vbo_secondary_properties(self.h, n, vort, C.byref(u_rms))
        return vort, u_rms.value

    def circle_forces(self):
        """Force of the fluid on each circle (after phase 11), (nc, 2), and particles per circle."""
        nc = self.case.nb_circles
        f = np.zeros(2 * max(nc, 1)); cnt = np.zeros(max(nc, 1), np.int32)
        lib().vbo_circle_forces(self.h, nc, f, cnt)
        return f[:2 * nc].reshape(nc, 2), cnt[:nc]

    def nusselt(self):
        return lib().vbo_nusselt(self.h)

    def surface(self, n=None):
        """n_surf, conc, div_r, smoothed surface normal (n, 2), ppe_smooth of rows 1..n (from phase 6)."""
        n = self.case.npfb if n is None else n
        ns = np.zeros(n, np.int32); co = np.zeros(n); dr = np.zeros(n); sn = np.zeros(2 * n); ps = np.zeros(n)
        lib().vbo_get_surface(self.h, n, ns, co, dr, sn, ps)
        return {"nsurf": ns, "conc": co, "divr": dr, "sn": sn.reshape(2, n).T, "psmooth": ps}

    def set_rows(self, s, n):
        """Rows 1..n of the state (n up to np: the mirror rows of the current step too)."""
        lib().vbo_set_particles(self.h, n, _c(s["x"][:n]), _c(s["y"][:n]), _c(s["u"][:n]), _c(s["v"][:n]),
                                _c(s["theta"][:n]), _c(s["tau"][:n]), _c(s["trace"][:n]), _c(s["p"][:n]),
                                _c(s["inbin"][:n], np.int32))


def fac_wab(q, ad_7, ad_7h):
    q = _c(q); f = np.zeros_like(q); w = np.zeros_like(q)
    lib().vbo_fac_wab(q.shape[0], q, ad_7, ad_7h, f, w)
    return f, w


def calc_eigens(a_n4):
    a = _c(a_n4); n = a.shape[0]
    lam = np.zeros((n, 2)); r = np.zeros((n, 4))
    lib().vbo_calc_eigens(n, a, lam, r)
    return lam, r


def logconf(tau_n4, trc, gradvel_n4, dt, wi, eps_ptt=0.0, kick=False):
    tau = _c(tau_n4); n = tau.shape[0]
    to = np.zeros((n, 4)); tr = np.zeros(n)
    lib().vbo_logconf(n, tau, _c(trc), _c(gradvel_n4), dt, wi, eps_ptt, int(kick), to, tr)
    return to, tr


def bicgstab_nonull(a, ija, b, x0=None, itermax=1000, tol=1e-8):
    n = b.shape[0]
    x = _c(np.zeros(n) if x0 is None else x0).copy()
    it = C.c_int(0); res = C.c_double(0.0)
    st = lib().vbo_bicgstab_nonull(n, a.shape[0], _c(a), _c(ija, np.int32), x, _c(b), itermax, tol,
                                   C.byref(it), C.byref(res))
    return x, it.value, res.value, st


def bicgstab_open(a, ija, b, x0=None, itermax=1000, tol=1e-8):
    n = b.shape[0]
    x = _c(np.zeros(n) if x0 is None else x0).copy()
    it = C.c_int(0); res = C.c_double(0.0)
    st = lib().vbo_bicgstab_open(n, a.shape[0], _c(a), _c(ija, np.int32), x, _c(b), itermax, tol,
                                 C.byref(it), C.byref(res))
    return x, it.value, res.value, st


def spmv_nr(a, ija, x):
    n = x.shape[0]
    y = np.zeros(n)
    lib().vbo_spmv_nr(n, _c(a), _c(ija, np.int32), _c(x), y)
    return y


def nnf(model, sr, mu):
    sr = _c(sr); out = np.zeros_like(sr)
    lib().vbo_nnf(model, sr.shape[0], sr, mu, out)
    return out
