"""ctypes binding of libvebisph_hip.so (the C ABI in include/vebisph_gpu.h).

Loading fails loudly when the library has not been built: there is no CPU fallback in the product.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvebisph_hip.so")

VB_OK = 0
STATUS = {0: "VB_OK", 1: "VB_ERR_OUT_OF_GRID", 2: "VB_ERR_CELL_OVERFLOW", 3: "VB_ERR_ALL_BINNED",
          4: "VB_ERR_CAPACITY", 5: "VB_WARN_BICGSTAB_BREAKDOWN", 6: "VB_WARN_MAXITER",
          7: "VB_ERR_HIP", 8: "VB_ERR_ARG"}
WARNINGS = (5, 6)
NPLINK = 75
PREC_FP64, PREC_MIXED = 0, 1

EXPORTS = [
    "vb_create", "vb_destroy", "vb_last_error", "vb_upload_state", "vb_download_state",
    "vb_set_time_step", "vb_put_in_bin", "vb_create_mirror_particles", "vb_particle_allocation",
    "vb_kernel_calculation", "vb_locate_free_surf_part", "vb_new_fick_shift", "vb_prediction_step",
    "vb_mirror_velocities", "vb_ppe_solve", "vb_sparse_pbicgstab_nonull", "vb_sparse_pbicgstab",
    "vb_pressure_update",
    "vb_final_advection", "vb_div_free",
    "vb_get_scalars", "vb_export_mirrors", "vb_export_cells", "vb_export_neighbours", "vb_export_weights",
    "vb_export_ushift", "vb_export_grad_p", "vb_export_ppe", "vb_secondary_properties",
    "vb_ppe_operator", "vb_grad_operator", "vb_ppe_adjoint_solve", "vb_adjoint_record",
    "vb_adjoint_backward", "vb_free_surface", "vb_field_maxima", "vb_circle_forces", "vb_nusselt", "vb_profile_spmv", "vb_profile_copy",
    "vb_nnf",
    "vb_kernel_functions", "vb_logconf", "vb_logconf_noalloc", "vb_logconf_adjoint",
]


class VbParams(C.Structure):
    _fields_ = [("h", C.c_double), ("dx", C.c_double), ("dv", C.c_double), ("Ra", C.c_double),
                ("Pr", C.c_double), ("El", C.c_double), ("beta", C.c_double),
                ("body_force", C.c_double * 2), ("dt_coef_cfl", C.c_double),
                ("dt_coef_visc", C.c_double), ("ve_nonlinearity", C.c_double), ("time", C.c_double),
                ("itime", C.c_int32), ("i_open_domain", C.c_int32), ("i_shift_coeff", C.c_int32),
                ("precision", C.c_int32), ("reference_kick", C.c_int32),
                ("serial_reductions", C.c_int32), ("fluid_model", C.c_int32)]


class VbBoundary(C.Structure):
    _fields_ = [("nb_patches", C.c_int32), ("b_node", C.POINTER(C.c_double)),
                ("b_type", C.POINTER(C.c_int32)), ("b_vel", C.POINTER(C.c_double)),
                ("b_thermal", C.POINTER(C.c_double)), ("b_periodic_parent", C.POINTER(C.c_int32)),
                ("nb_circles", C.c_int32), ("c_centre", C.POINTER(C.c_double)),
                ("c_radius", C.POINTER(C.c_double)), ("c_omega", C.POINTER(C.c_double)),
                ("c_vel", C.POINTER(C.c_double)), ("c_thermal", C.POINTER(C.c_double))]


class VbError(RuntimeError):
    def __init__(self, status, msg=""):
        self.status = status
        super().__init__(f"{STATUS.get(status, status)}: {msg}")


_lib = None


def built() -> bool:
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        if not built():
            raise RuntimeError(f"{LIB_PATH} is not built; run python -m vebisph_amd.build "
                               "(the GPU path has no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int32)
        L.vb_create.argtypes = [C.POINTER(VbParams), C.POINTER(VbBoundary), C.c_int32, C.c_int32,
                                C.POINTER(vp)]
        L.vb_destroy.argtypes = [vp]
        L.vb_last_error.argtypes = [vp]
        L.vb_last_error.restype = C.c_char_p
        L.vb_upload_state.argtypes = [vp, C.c_int32] + [dp] * 11 + [ip]
        L.vb_download_state.argtypes = [vp, C.c_int32] + [dp] * 11 + [ip]
        for name in ("vb_set_time_step", "vb_put_in_bin", "vb_create_mirror_particles",
                     "vb_particle_allocation", "vb_kernel_calculation", "vb_locate_free_surf_part",
                     "vb_new_fick_shift", "vb_prediction_step", "vb_mirror_velocities",
                     "vb_pressure_update", "vb_final_advection", "vb_div_free"):
            getattr(L, name).argtypes = [vp]
        L.vb_ppe_solve.argtypes = [vp, C.c_double, C.c_int32, ip, dp]
        L.vb_sparse_pbicgstab_nonull.argtypes = [vp, dp, ip, dp, dp, C.c_int32, C.c_int32, C.c_int32,
                                                 C.c_double, ip, dp]
        L.vb_sparse_pbicgstab.argtypes = L.vb_sparse_pbicgstab_nonull.argtypes
        L.vb_get_scalars.argtypes = [vp, ip, dp]
        L.vb_export_mirrors.argtypes = [vp, C.c_int32, ip, ip, dp]
        L.vb_export_cells.argtypes = [vp, C.c_int32, ip, C.c_int32, ip, ip]
        L.vb_export_neighbours.argtypes = [vp, C.c_int32, C.c_int32, ip, ip]
        L.vb_export_weights.argtypes = [vp, C.c_int32, C.c_int32, dp, dp, dp]
        L.vb_export_ushift.argtypes = [vp, C.c_int32, dp]
        L.vb_export_grad_p.argtypes = [vp, C.c_int32, dp]
        L.vb_export_ppe.argtypes = [vp, C.c_int32, C.c_int32, dp, ip, dp, dp]
        L.vb_adjoint_record.argtypes = [vp, C.c_int32, C.c_int32, C.c_double, C.c_int32, ip]
        L.vb_adjoint_backward.argtypes = ([vp, C.c_int32] + [dp] * 10 + [C.c_double, C.c_int32, dp]
                                          + [dp] * 10 + [ip])
        L.vb_profile_spmv.argtypes = [vp, C.c_int32, dp, dp]
        L.vb_profile_copy.argtypes = [vp, C.c_int64, C.c_int32, dp]
        L.vb_nnf.argtypes = [C.c_int32, C.c_int32, vp, C.c_double, vp]
        L.vb_kernel_functions.argtypes = [C.c_int32, vp, C.c_double, C.c_double, vp, vp]
        L.vb_logconf.argtypes = [C.c_int32, vp, vp, vp, C.c_double, C.c_double, C.c_double,
                                 C.c_int32, vp, vp]
        L.vb_logconf_noalloc.argtypes = L.vb_logconf.argtypes
        L.vb_logconf_adjoint.argtypes = [C.c_int32, vp, vp, vp, C.c_double, C.c_double, C.c_double,
                                         C.c_int32, vp, vp, vp, vp, vp, vp, vp]
        for name in EXPORTS:
            if name != "vb_last_error":
                getattr(L, name).restype = C.c_int32
        _lib = L
    return _lib
