"""ctypes binding of the Fortran driver in libvebisph_f.so (vebisph_amd/fortran/vebisph_gpu.F90).

`drive()` runs nsteps of the Fortran `div_free` (module divfree_gpu, the drop-in for the reference's
module divfree) on a case: the Fortran side builds the reference-shaped host arrays r(npfb,2),
u(npfb,2), tau_p(npfb,2,2) and logical inbin, uploads them through the iso_c_binding interface,
steps, and downloads the state.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from .solver import initial_state, make_boundary, make_params

FLIB_PATH = os.path.join(_lib.HERE, "libvebisph_f.so")
FEXPORTS = ["vbf_drive"]

_flib = None


def flib():
    global _flib
    if _flib is None:
        if not os.path.exists(FLIB_PATH):
            raise RuntimeError(f"{FLIB_PATH} is not built; run python -m vebisph_amd.build")
        _lib.lib()                                   # libvebisph_hip.so first (its dependency)
        L = C.CDLL(FLIB_PATH)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int32)
        L.vbf_drive.argtypes = ([C.POINTER(_lib.VbParams), C.POINTER(_lib.VbBoundary), C.c_int32,
                                 C.c_int32, C.c_int32] + [dp] * 11 + [ip, ip, dp])
        L.vbf_drive.restype = C.c_int
        _flib = L
    return _flib


def drive(case, nsteps, state=None, precision=_lib.PREC_FP64, time=0.0, itime=0):
    """Returns (status, LS_iters of the last step, time reached, state dict as Solver.get_state)."""
    L = flib()
    keep = []
    p = make_params(case, precision, True, time, itime, False)
    b = make_boundary(case, keep)
    s = initial_state(case) if state is None else state
    n = case.npfb
    col = lambda a: np.array(a, dtype=np.float64, copy=True)
    out = {k: col(s[k]) for k in ("x", "y", "u", "v", "theta", "trace", "p")}
    tau = np.asarray(s["tau"], dtype=np.float64)
    taus = [col(tau[:, q]) for q in range(4)]
    inb = np.array(s["inbin"], dtype=np.int32, copy=True)
    it = C.c_int32(0)
    t = C.c_double(0.0)
    d = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    st = L.vbf_drive(C.byref(p), C.byref(b), n, case.nbS, nsteps, d(out["x"]), d(out["y"]),
                     d(out["u"]), d(out["v"]), d(out["theta"]), *[d(q) for q in taus],
                     d(out["trace"]), d(out["p"]), inb.ctypes.data_as(C.POINTER(C.c_int32)),
                     C.byref(it), C.byref(t))
    out["tau"] = np.stack(taus, axis=1)
    out["inbin"] = inb
    return st, it.value, t.value, out
