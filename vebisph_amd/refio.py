"""The reference's run files from Python: ctypes over libvebisph_refio.so (include/vebisph_refio.h).

A run directory holds INDAT (run parameters, input.F90:18-37), IBOUND (boundary patches and circles,
input.F90:51-81) and IPART (one x y u v theta dv record per particle, input.F90:139-144); the
reference's program writes data_directory/PART<n>, TIME, LS, TIME_OUT and TEMP_REC.dat
(output.F90:1-89).  read_case_dir/write_case_dir map a directory to and from a cases.Case;
read_part and read_temp_rec read the output files back.  vb_drive (``vb_drive run DIR``) runs such
a directory on the GPU.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os

import numpy as np

from . import cases

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libvebisph_refio.so")
MAX_PATCHES, MAX_CIRCLES = 16, 8

_d = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


class Indat(C.Structure):
    _fields_ = [("i_restart", C.c_int32), ("i_open_domain", C.c_int32), ("tmax", C.c_double),
                ("i_shift_coeff", C.c_int32), ("np", C.c_int32), ("npfb", C.c_int32), ("nbS", C.c_int32),
                ("dt_out", C.c_double), ("h", C.c_double), ("dx", C.c_double), ("dt", C.c_double),
                ("body_force", C.c_double * 2), ("Ra", C.c_double), ("Pr", C.c_double), ("El", C.c_double),
                ("beta", C.c_double), ("dt_coef_cfl", C.c_double), ("dt_coef_visc", C.c_double),
                ("ve_nonlinearity", C.c_double), ("alpha_evss", C.c_double)]


class Ibound(C.Structure):
    _fields_ = [("nb_patches", C.c_int32), ("b_node", (C.c_double * 2) * MAX_PATCHES),
                ("b_edge", (C.c_double * 2) * MAX_PATCHES), ("b_type", C.c_int32 * MAX_PATCHES),
                ("b_vel", C.c_double * MAX_PATCHES), ("b_thermal", C.c_double * MAX_PATCHES),
                ("b_periodic_parent", C.c_int32 * MAX_PATCHES), ("u_inflow", C.c_double),
                ("nb_circles", C.c_int32), ("c_centre", (C.c_double * 2) * MAX_CIRCLES),
                ("c_radius", C.c_double * MAX_CIRCLES), ("c_omega", C.c_double * MAX_CIRCLES),
                ("c_vel", (C.c_double * 2) * MAX_CIRCLES), ("c_thermal", C.c_double * MAX_CIRCLES)]


class Restart(C.Structure):
    _fields_ = [("itime", C.c_int32), ("time", C.c_double), ("n_dump", C.c_int32), ("np", C.c_int32),
                ("npfb", C.c_int32), ("nbS", C.c_int32), ("np_inflow", C.c_int32), ("xb_min", C.c_double),
                ("xb_max", C.c_double), ("yb_min", C.c_double), ("yb_max", C.c_double), ("dx", C.c_double),
                ("av_conc", C.c_double)]


_lib = None


class _OptD:
    """an FP64 array argument that may be None (a NULL pointer)"""
    @classmethod
    def from_param(cls, a):
        return None if a is None else _d.from_param(a)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(LIB_PATH)
        L.vbio_last_error.restype = C.c_char_p
        L.vbio_read_indat.argtypes = [C.c_char_p, C.POINTER(Indat)]
        L.vbio_write_indat.argtypes = [C.c_char_p, C.POINTER(Indat)]
        L.vbio_read_ibound.argtypes = [C.c_char_p, C.POINTER(Ibound)]
        L.vbio_write_ibound.argtypes = [C.c_char_p, C.POINTER(Ibound)]
        L.vbio_check_ibound.argtypes = [C.POINTER(Ibound)]
        L.vbio_read_ipart.argtypes = [C.c_char_p, C.c_int32] + [_d] * 5 + [C.POINTER(C.c_double)]
        L.vbio_write_ipart.argtypes = [C.c_char_p, C.c_int32] + [_d] * 5 + [C.c_double]
        L.vbio_write_part.argtypes = [C.c_char_p, C.c_int32] + [_d] * 8 + [_i] + [_d] + [_OptD] * 3 + [_i]
        L.vbio_read_part.argtypes = [C.c_char_p, C.c_int32, C.POINTER(C.c_int32), _d]
        L.vbio_read_part_ncols.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.POINTER(C.c_int32), _d]
        L.vbio_write_temp_rec.argtypes = ([C.c_char_p, C.POINTER(Restart), C.c_int32] + [_d] * 6
                                          + [C.c_double, _i] + [_OptD] * 4 + [_i])
        L.vbio_read_temp_rec.argtypes = ([C.c_char_p, C.POINTER(Restart), C.c_int32] + [C.c_void_p] * 12)
        L.vbio_append_record.argtypes = [C.c_char_p, C.c_char_p, _d]
        L.vbio_format_real.argtypes = [C.c_double, C.c_char_p, C.c_int32]
        _lib = L
    return _lib


class RefIOError(RuntimeError):
    pass


def _chk(status):
    if status != 0:
        raise RefIOError(lib().vbio_last_error().decode())


def _b(path):
    return os.fsencode(path)


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


# ---- INDAT / IBOUND / IPART -----------------------------------------------------------------------
def read_indat(path) -> Indat:
    o = Indat()
    _chk(lib().vbio_read_indat(_b(path), C.byref(o)))
    return o


def read_ibound(path) -> Ibound:
    o = Ibound()
    _chk(lib().vbio_read_ibound(_b(path), C.byref(o)))
    return o


def check_ibound(b: Ibound) -> None:
    _chk(lib().vbio_check_ibound(C.byref(b)))


def read_ipart(path, npfb):
    x, y, u, v, th = (np.zeros(npfb) for _ in range(5))
    dv = C.c_double(0.0)
    _chk(lib().vbio_read_ipart(_b(path), npfb, x, y, u, v, th, C.byref(dv)))
    return x, y, u, v, th, dv.value


@dataclasses.dataclass
class RunDir:
    """A run directory: the case, and the INDAT values a Case does not hold."""
    case: cases.Case
    tmax: float
    dt_out: float
    dt: float = 0.0
    restart: bool = False
    alpha_evss: float = 0.0


def read_case_dir(path) -> RunDir:
    """getdata (input.F90:12-180) on a run directory, without restart: INDAT, IBOUND, IPART."""
    ind = read_indat(os.path.join(path, "INDAT"))
    ib = read_ibound(os.path.join(path, "IBOUND"))
    check_ibound(ib)
    x, y, u, v, th, dv = read_ipart(os.path.join(path, "IPART"), ind.npfb)
    nb, nc = ib.nb_patches, ib.nb_circles
    arr = lambda a, n, w=None: np.array([[a[k][j] for j in range(w)] for k in range(n)]) if w else np.array(a[:n])
    case = cases.Case(
        name=os.path.basename(os.path.abspath(path)), h=ind.h, dx=ind.dx, dv=dv, Ra=ind.Ra, Pr=ind.Pr,
        El=ind.El, beta=ind.beta, body_force=(ind.body_force[0], ind.body_force[1]),
        dt_coef_cfl=ind.dt_coef_cfl, dt_coef_visc=ind.dt_coef_visc, ve_nonlinearity=ind.ve_nonlinearity,
        i_open_domain=ind.i_open_domain, i_shift_coeff=ind.i_shift_coeff,
        b_node=arr(ib.b_node, nb, 2), b_type=np.array(ib.b_type[:nb], np.int32), b_vel=arr(ib.b_vel, nb),
        b_thermal=arr(ib.b_thermal, nb), b_periodic_parent=np.array(ib.b_periodic_parent[:nb], np.int32),
        c_centre=arr(ib.c_centre, nc, 2) if nc else None, c_radius=arr(ib.c_radius, nc) if nc else None,
        c_omega=arr(ib.c_omega, nc) if nc else None, c_vel=arr(ib.c_vel, nc, 2) if nc else None,
        c_thermal=arr(ib.c_thermal, nc) if nc else None,
        nbS=ind.nbS, x=x, y=y, u=u, v=v, theta=th)
    return RunDir(case=case, tmax=ind.tmax, dt_out=ind.dt_out, dt=ind.dt, restart=bool(ind.i_restart),
                  alpha_evss=ind.alpha_evss)


def write_case_dir(path, case: cases.Case, tmax: float, dt_out: float, dt: float = 0.0,
                   restart: bool = False) -> None:
    """INDAT, IBOUND and IPART of a case in the reference's formats (as its generator writes them:
    b_edge(ib) = b_node(ib+1) - b_node(ib))."""
    os.makedirs(path, exist_ok=True)
    ind = Indat(i_restart=int(restart), i_open_domain=case.i_open_domain, tmax=tmax,
                i_shift_coeff=case.i_shift_coeff, np=case.npfb, npfb=case.npfb, nbS=case.nbS, dt_out=dt_out,
                h=case.h, dx=case.dx, dt=dt, Ra=case.Ra, Pr=case.Pr, El=case.El, beta=case.beta,
                dt_coef_cfl=case.dt_coef_cfl, dt_coef_visc=case.dt_coef_visc,
                ve_nonlinearity=case.ve_nonlinearity, alpha_evss=0.0)
    ind.body_force[0], ind.body_force[1] = case.body_force
    _chk(lib().vbio_write_indat(_b(os.path.join(path, "INDAT")), C.byref(ind)))
    ib = Ibound()
    nb = case.nb_patches
    ib.nb_patches = nb
    edge = case.b_edge
    for k in range(nb):
        ib.b_node[k][0], ib.b_node[k][1] = case.b_node[k]
        ib.b_edge[k][0], ib.b_edge[k][1] = edge[k]
        ib.b_type[k] = int(case.b_type[k]); ib.b_vel[k] = case.b_vel[k]; ib.b_thermal[k] = case.b_thermal[k]
        ib.b_periodic_parent[k] = int(case.b_periodic_parent[k]) if case.b_type[k] == 2 else 0
    ib.nb_circles = case.nb_circles
    for k in range(case.nb_circles):
        ib.c_centre[k][0], ib.c_centre[k][1] = case.c_centre[k]
        ib.c_radius[k] = case.c_radius[k]; ib.c_omega[k] = case.c_omega[k]
        ib.c_vel[k][0], ib.c_vel[k][1] = case.c_vel[k]
        ib.c_thermal[k] = case.c_thermal[k]
    _chk(lib().vbio_write_ibound(_b(os.path.join(path, "IBOUND")), C.byref(ib)))
    _chk(lib().vbio_write_ipart(_b(os.path.join(path, "IPART")), case.npfb, _c(case.x), _c(case.y),
                                _c(case.u), _c(case.v), _c(case.theta), case.dv))


# ---- outputs ---------------------------------------------------------------------------------------
PART_COLUMNS = ("x", "y", "u", "v", "theta", "p", "conc", "a_out", "n_surf", "vort", "tau11", "tau12", "tau22")


def read_part(path, cap, newtonian=False):
    """The 13 columns of a PART file (output.F90:44-51) as a dict of arrays (the Newtonian build's
    file: the first 10)."""
    nc = 10 if newtonian else 13
    cols = np.zeros(nc * cap)
    n = C.c_int32(0)
    _chk(lib().vbio_read_part_ncols(_b(path), nc, cap, C.byref(n), cols))
    cols = cols.reshape(nc, cap)[:, :n.value]
    return dict(zip(PART_COLUMNS, cols))


def write_part(path, state, vort, a_out=None, conc=None, n_surf=None, newtonian=False):
    """A PART file of a state dict (Solver.get_state layout) with the reference's columns (the
    Newtonian build's: without the stress columns)."""
    n = state["x"].shape[0]
    z = np.zeros(n)
    tau = state["tau"]
    ts = (None, None, None) if newtonian else (_c(tau[:, 0]), _c(tau[:, 1]), _c(tau[:, 3]))
    _chk(lib().vbio_write_part(_b(path), n, _c(state["x"]), _c(state["y"]), _c(state["u"]), _c(state["v"]),
                               _c(state["theta"]), _c(state["p"]), _c(z if conc is None else conc),
                               _c(state["trace"] if a_out is None else a_out),
                               _c(np.zeros(n) if n_surf is None else n_surf, np.int32), _c(vort),
                               *ts, _c(state["inbin"], np.int32)))


def read_temp_rec(path, newtonian=False):
    """TEMP_REC.dat: the header and the particle columns (eight values per particle line; the
    Newtonian build's file has no stress lines, and the tau columns come back as 0)."""
    hd = Restart()
    _chk(lib().vbio_read_temp_rec(_b(path), C.byref(hd), 0, *([None] * 12)))
    n = hd.npfb
    cols = {k: np.zeros(n) for k in ("x", "y", "u", "v", "theta", "p", "tau11", "tau12", "tau21", "tau22")}
    dv = C.c_double(0.0)
    ns = np.zeros(n, np.int32)
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    _chk(lib().vbio_read_temp_rec(_b(path), C.byref(hd), n, ptr(cols["x"]), ptr(cols["y"]), ptr(cols["u"]),
                                  ptr(cols["v"]), ptr(cols["theta"]), ptr(cols["p"]), C.cast(C.byref(dv), C.c_void_p),
                                  ptr(ns), *([None] * 4 if newtonian else
                                             [ptr(cols["tau11"]), ptr(cols["tau12"]), ptr(cols["tau21"]),
                                              ptr(cols["tau22"])])))
    cols["n_surf"] = ns
    cols["dv"] = dv.value
    return hd, cols


def read_table(path, nitems):
    """A per-step text file (TIME: 6 items, LS: 2, TIME_OUT: 4, fort.192: 2) as an (nrec, nitems)
    array; a record may continue over several lines."""
    with open(path) as f:
        toks = [float(t.replace("D", "E")) for t in f.read().split()]
    if len(toks) % nitems:
        raise RefIOError(f"{path}: {len(toks)} values is not a whole number of {nitems}-item records")
    return np.array(toks).reshape(-1, nitems)


def format_real(x: float) -> str:
    buf = C.create_string_buffer(64)
    n = lib().vbio_format_real(float(x), buf, 64)
    if n < 0:
        raise RefIOError("format buffer too small")
    return buf.value.decode()
