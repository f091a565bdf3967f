"""The reference generator's cases (vebisph_amd.generator, equal to datclass.F90's output) on the
gfx950 path: read from their run directories, two steps against the oracle (itself bit-exact against
the reference on these cases), and `vb_drive run` on a generated directory.

Tolerances as in tests/test_gpu_parity.py: neighbour lists exact; positions after each step to the
FP64 field bar 1e-10 (the BiCGSTAB's parallel dot products round differently), pressure 1e-3.
Cases 1 and 2 are Newtonian (El = 0): the viscoelastic step's polymeric coefficient is 0/0 there, in
the reference as here."""
import os

import numpy as np
import pytest

from vebisph_amd import generator, refio, solver
from oracle import oraclewrap as ow

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("k", [3, 4, 5, 6, 7, 8])
def test_generator_case_against_oracle(gpu_lib, k, tmp_path):
    generator.write_run_dir(k, str(tmp_path))
    case = refio.read_case_dir(str(tmp_path)).case
    print(f"generator case {k}: npfb={case.npfb} nbS={case.nbS} open={case.i_open_domain}", flush=True)
    G = solver.Solver(case); O = ow.OracleSolver(case)
    for step in range(2):
        for ph in range(1, 13):
            G.phase(ph); O.phase(ph)
            if ph == 4:
                cg, lg = G.neighbours(); co, lo = O.neighbours()
                assert np.array_equal(cg, co) and np.array_equal(lg, lo)
            if ph == 10:
                s = G.scalars()
                assert s["status"] in (0, 6) and s["LS_residual"] <= 1e-8
                assert _rel(G.get_state()["p"], O.get_state()["p"]) <= 1e-3
        a, b = G.get_state(), O.get_state()
        print(f"  step {step + 1}: x {_rel(a['x'], b['x']):.2e} y {_rel(a['y'], b['y']):.2e} "
              f"u {_rel(a['u'], b['u']):.2e} p {_rel(a['p'], b['p']):.2e}", flush=True)
        assert _rel(a["x"], b["x"]) <= 1e-10 and _rel(a["y"], b["y"]) <= 1e-10
        assert _rel(a["u"], b["u"]) <= 1e-5 and _rel(a["v"], b["v"]) <= 1e-5


def test_vb_drive_runs_generated_directory(gpu_lib, tmp_path):
    """generator -> vb_drive run: the flow past a cylinder (case 5) for three steps, a PART file per
    step, and the state the ctypes path reaches in the same three steps."""
    import subprocess
    generator.write_run_dir(5, str(tmp_path))
    ind = (tmp_path / "INDAT").read_text().replace("      0.0099999998 :dt_out", "      0.0000000001 :dt_out")
    (tmp_path / "INDAT").write_text(ind)
    exe = os.path.join(os.path.dirname(solver.__file__), "vb_drive")
    r = subprocess.run([exe, "run", str(tmp_path), "fp64", "3"], capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0
    case = refio.read_case_dir(str(tmp_path)).case
    G = solver.Solver(case)
    for _ in range(3):
        G.div_free()
    st = G.get_state()
    part = refio.read_part(str(tmp_path / "data_directory" / "PART2"), case.npfb)
    keep = st["inbin"] == 0
    for key in ("x", "y", "u", "v", "p"):
        assert np.array_equal(part[key], st[key][keep]), key
