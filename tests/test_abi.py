"""The C ABI (include/vebisph_gpu.h) and its Fortran interface, checked without a GPU: the libraries
load, export every declared entry point, and the record layouts agree between C, ctypes and
Fortran.  No compute call is made."""
import ctypes as C
import os
import re
import shutil
import subprocess

import pytest

from vebisph_amd import _lib, build, fortran_drive

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vebisph_gpu.h")
F90 = os.path.join(ROOT, "vebisph_amd", "fortran", "vebisph_gpu.F90")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:vb_status|const char\*)\s+(vb_\w+)\s*\(", text, flags=re.M)))


def fortran_bindings():
    return sorted(set(re.findall(r'bind\(C,\s*name="(vb_\w+)"\)', open(F90).read())))


def test_header_declares_the_step():
    names = header_functions()
    for n in ("vb_create", "vb_destroy", "vb_div_free", "vb_ppe_solve", "vb_sparse_pbicgstab_nonull",
              "vb_particle_allocation", "vb_kernel_calculation", "vb_prediction_step", "vb_nnf"):
        assert n in names
    assert sorted(_lib.EXPORTS) == names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    missing = [n for n in header_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_fortran_interface_covers_the_header():
    assert fortran_bindings() == header_functions()


def test_fortran_library_exports_driver():
    if not os.path.exists(fortran_drive.FLIB_PATH):
        if not os.path.exists(build.FLANG):
            pytest.skip("flang absent and libvebisph_f.so not built")
        build.build_fortran()
    L = fortran_drive.flib()
    for n in fortran_drive.FEXPORTS:
        assert hasattr(L, n)


def _c_layout():
    """sizeof / offsetof of the public records, from a C program compiled against the header."""
    if shutil.which("gcc") is None:
        pytest.skip("gcc absent")
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "vebisph_gpu.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(vb_params), offsetof(vb_params, body_force),
         offsetof(vb_params, itime), offsetof(vb_params, fluid_model),
         sizeof(vb_boundary), offsetof(vb_boundary, c_centre));
  return 0;
}
"""
    d = os.path.join(ROOT, "vebisph_amd", "_build")
    os.makedirs(d, exist_ok=True)
    cfile, exe = os.path.join(d, "abi_layout.c"), os.path.join(d, "abi_layout")
    open(cfile, "w").write(src)
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), cfile, "-o", exe], check=True)
    return [int(v) for v in subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()]


def test_record_layout_c_vs_ctypes():
    P, B = _lib.VbParams, _lib.VbBoundary
    assert _c_layout() == [C.sizeof(P), P.body_force.offset, P.itime.offset, P.fluid_model.offset,
                           C.sizeof(B), B.c_centre.offset]


def test_record_layout_c_vs_fortran():
    if not os.path.exists(build.FLANG):
        pytest.skip("flang absent")
    if not os.path.exists(os.path.join(build.BUILD, "vebisph_gpu.mod")):
        build.build_fortran()
    src = """
program layout
  use iso_c_binding
  use vebisph_gpu
  type(vb_params), target :: p
  type(vb_boundary), target :: b
  print '(6i8)', c_sizeof(p), loc_diff(c_loc(p%body_force), c_loc(p)), loc_diff(c_loc(p%itime), c_loc(p)), &
       loc_diff(c_loc(p%fluid_model), c_loc(p)), c_sizeof(b), loc_diff(c_loc(b%c_centre), c_loc(b))
contains
  integer function loc_diff(a, o)
    type(c_ptr), intent(in) :: a, o
    integer(c_intptr_t) :: ia, io
    ia = transfer(a, ia); io = transfer(o, io)
    loc_diff = int(ia - io)
  end function
end program
"""
    d = build.BUILD
    f, exe = os.path.join(d, "abi_layout.f90"), os.path.join(d, "abi_layout_f")
    open(f, "w").write(src)
    # the record types only: compile the program against the module, link nothing from the GPU side
    obj = os.path.join(d, "vebisph_gpu_f.o")
    subprocess.run([build.FLANG, "-I", d, f, obj, "-o", exe, f"-L{build.PKG}", "-lvebisph_hip",
                    f"-Wl,-rpath,{build.PKG}"], check=True)
    got = [int(v) for v in subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()]
    assert got == _c_layout()


@pytest.mark.parametrize("const_mod", [2, 1])
def test_reference_glue_compiles_against_reference_modules(const_mod):
    """divfree_reference_glue.F90 replaces the reference's module divfree inside the reference's own
    build: it must compile against the reference's module files (oracle/_ref/obj, and oracle/_ref/newt/obj
    for the Newtonian build, built from the reference sources by oracle/Makefile) with the build's
    -Dconst_mod."""
    ref_obj = os.path.join(ROOT, "oracle", "_ref", *(("newt", "obj") if const_mod == 1 else ("obj",)))
    if not os.path.exists(os.path.join(ref_obj, "common_2d.mod")) or not os.path.exists(build.FLANG):
        pytest.skip("reference module files not built here")
    if not os.path.exists(os.path.join(build.BUILD, "divfree_gpu.mod")):
        build.build_fortran()
    out = os.path.join(build.BUILD, f"glue_check_{const_mod}")
    os.makedirs(out, exist_ok=True)
    subprocess.run([build.FLANG, "-cpp", f"-Dconst_mod={const_mod}", "-c", "-fPIC", "-module-dir", out, "-I",
                    build.BUILD, "-I", ref_obj,
                    os.path.join(ROOT, "vebisph_amd", "fortran", "divfree_reference_glue.F90"),
                    "-o", os.path.join(out, "glue.o")], check=True)


def test_cpp_driver_rejects_bad_input(tmp_path):
    """vb_drive's own argument and file checks (no device call is reached)."""
    exe = os.path.join(build.PKG, "vb_drive")
    if not os.path.exists(exe):
        build.build_cpp_driver()
    assert subprocess.run([exe], capture_output=True).returncode == 2
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"XXXX")
    assert subprocess.run([exe, str(bad), "1", str(tmp_path / "o")], capture_output=True).returncode == 2
    from vebisph_amd import cases, solver
    good = tmp_path / "c.bin"
    solver.write_case_file(str(good), cases.make("poiseuille_2500"))
    data = good.read_bytes()
    (tmp_path / "trunc.bin").write_bytes(data[:-100])
    r = subprocess.run([exe, str(tmp_path / "trunc.bin"), "1", str(tmp_path / "o")], capture_output=True, text=True)
    assert r.returncode == 2 and "truncated" in r.stderr


# ---- the run-file library (include/vebisph_refio.h) ----
REFIO_HEADER = os.path.join(ROOT, "include", "vebisph_refio.h")


def refio_functions():
    text = re.sub(r"/\*.*?\*/", "", open(REFIO_HEADER).read(), flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int32_t|const char\*)\s+(vbio_\w+)\s*\(", text, flags=re.M)))


def test_refio_library_exports_every_declared_symbol():
    from vebisph_amd import refio
    names = refio_functions()
    assert "vbio_read_indat" in names and "vbio_write_part" in names and "vbio_read_temp_rec" in names
    L = refio.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_refio_record_layout_c_vs_ctypes():
    from vebisph_amd import refio
    if shutil.which("gcc") is None:
        pytest.skip("gcc absent")
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "vebisph_refio.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(vbio_indat), offsetof(vbio_indat, body_force),
         offsetof(vbio_indat, alpha_evss), sizeof(vbio_ibound), offsetof(vbio_ibound, u_inflow),
         offsetof(vbio_ibound, c_thermal), sizeof(vbio_restart), offsetof(vbio_restart, xb_min),
         offsetof(vbio_restart, av_conc));
  return 0;
}
"""
    d = os.path.join(ROOT, "vebisph_amd", "_build")
    os.makedirs(d, exist_ok=True)
    cfile, exe = os.path.join(d, "refio_layout.c"), os.path.join(d, "refio_layout")
    open(cfile, "w").write(src)
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), cfile, "-o", exe], check=True)
    got = [int(v) for v in subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()]
    I, B, R = refio.Indat, refio.Ibound, refio.Restart
    assert got == [C.sizeof(I), I.body_force.offset, I.alpha_evss.offset, C.sizeof(B), B.u_inflow.offset,
                   B.c_thermal.offset, C.sizeof(R), R.xb_min.offset, R.av_conc.offset]


def test_create_refuses_npfb_beyond_32bit_slot_indices():
    """Neighbour slots are addressed k*npfb + i with 32-bit ints: vb_create refuses npfb above
    (2^31 - 1) / VB_NPLINK with VB_ERR_CAPACITY before it makes any device call (so this runs
    without a GPU), and leaves the output handle NULL."""
    from vebisph_amd import cases, solver
    L = _lib.lib()
    case = cases.make("poiseuille_2500")
    keep = []
    p = solver.make_params(case)
    b = solver.make_boundary(case, keep)
    h = C.c_void_p(1234)
    limit = (2 ** 31 - 1) // _lib.NPLINK
    assert limit == 28_633_115
    for npfb in (limit + 1, 2 ** 31 - 1):
        assert L.vb_create(C.byref(p), C.byref(b), npfb, 0, C.byref(h)) == 4       # VB_ERR_CAPACITY
        assert h.value is None
        h = C.c_void_p(1234)
    assert L.vb_create(C.byref(p), C.byref(b), 0, 0, C.byref(h)) == 8               # VB_ERR_ARG
