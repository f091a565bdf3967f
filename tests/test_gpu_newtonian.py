"""The Newtonian build (vb_params.fluid_model = 1; the reference compiled with -Dconst_mod=1: no
polymeric stress, RHS predictor_mod.F90:58-77) on the gfx950 path, against the oracle's Newtonian
branch (itself bit-exact against that reference build, tests/test_oracle_vs_reference.py).

Cases: the generator's Newtonian cases -- the dam break (1, an open domain) and Taylor-Green vortices
(2, all periodic; its initial field has an exactly zero discrete divergence, so the first solve
breaks down at rho = 0 in the reference and here) -- and the flow past a cylinder (5) run Newtonian.
Tolerances as in tests/test_gpu_parity.py: lists exact, positions after each step to the FP64 field
bar 1e-10, pressure 1e-3, ten steps within 10x the oracle's own one-ulp spread."""
import os
import subprocess

import numpy as np
import pytest

from tests.conftest import mixed_floor

from vebisph_amd import generator, refio, solver
from oracle import oraclewrap as ow

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300)


def _case(k, path):
    generator.write_run_dir(k, str(path))
    c = refio.read_case_dir(str(path)).case
    c.const_mod = 1
    return c


@pytest.mark.parametrize("k", [1, 2, 5])
def test_newtonian_steps_against_oracle(gpu_lib, k, tmp_path):
    case = _case(k, tmp_path)
    print(f"generator case {k} Newtonian: npfb={case.npfb}", flush=True)
    G = solver.Solver(case); O = ow.OracleSolver(case)
    for step in range(2):
        for ph in range(1, 13):
            G.phase(ph); O.phase(ph)
            if ph == 4:
                cg, lg = G.neighbours(); co, lo = O.neighbours()
                assert np.array_equal(cg, co) and np.array_equal(lg, lo)
            if ph == 9 and step == 0:
                # the Newtonian predictor from identical inputs (u*, theta*, after the mirror
                # velocities swap them in): to rounding
                a, b = G.get_state(), O.get_state()
                for key in ("u", "v", "theta"):
                    assert np.max(np.abs(a[key] - b[key])) <= 1e-13 * max(1.0, np.abs(b[key]).max()), key
                assert not np.any(a["tau"]) and np.all(a["trace"] == 2.0)
            if ph == 10:
                s = G.scalars()
                assert s["status"] in (0, 5, 6)
                assert _rel(G.get_state()["p"], O.get_state()["p"]) <= 1e-3 or not np.any(O.get_state()["p"])
        a, b = G.get_state(), O.get_state()
        print(f"  step {step + 1}: x {_rel(a['x'], b['x']):.2e} u {_rel(a['u'], b['u']):.2e} "
              f"p {_rel(a['p'], b['p']) if np.any(b['p']) else 0.0:.2e}", flush=True)
        assert _rel(a["x"], b["x"]) <= 1e-10 and _rel(a["y"], b["y"]) <= 1e-10
        assert _rel(a["u"], b["u"]) <= 1e-5 and _rel(a["v"], b["v"]) <= 1e-5
    if case.nb_circles:
        # the force on a circle after the first step's pressure update, on the same state (as
        # tests/test_gpu_parity.py does for the viscoelastic build), no polymeric term; case 5 starts
        # at rest with no body force (p = 0 after one step), so the body-forced cylinder is used
        from vebisph_amd import cases
        cyl = cases.cylinder(nx=160, ny=80, name="cylinder_13k_newtonian")
        cyl.const_mod = 1
        G = solver.Solver(cyl); O = ow.OracleSolver(cyl)
        for ph in range(1, 12):
            G.phase(ph); O.phase(ph)
        n = G.scalars()["np"]
        O.set_rows(G.get_state(n), n)
        fg, _ = G.circle_forces(); fo, _ = O.circle_forces()
        print(f"  force on the cylinder (no polymeric stress): GPU {fg[0]}, oracle {fo[0]}", flush=True)
        assert np.all(np.isfinite(fg)) and np.abs(fg).max() > 0 and _rel(fg, fo) <= 1e-12


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_newtonian_dam_break_ten_steps(gpu_lib, precision, tmp_path):
    from vebisph_amd import _lib
    case = _case(1, tmp_path)
    pert = dict(x=np.nextafter(case.x, np.inf), y=np.nextafter(case.y, np.inf), u=case.u, v=case.v,
                theta=case.theta, tau=np.zeros((case.npfb, 4)), trace=np.full(case.npfb, 2.0),
                p=np.zeros(case.npfb), inbin=np.zeros(case.npfb, np.int32))
    O = ow.OracleSolver(case); O2 = ow.OracleSolver(case, state=pert)
    G = solver.Solver(case, precision=_lib.PREC_MIXED if precision == "mixed" else _lib.PREC_FP64)
    for _ in range(10):
        O.step(); O2.step(); G.div_free()
        assert G.scalars()["status"] in (0, 6)
    a, b, b2 = G.get_state(), O.get_state(), O2.get_state()
    for key in ("x", "y", "u", "v", "theta", "p"):
        spread, err = _rel(b2[key], b[key]), _rel(a[key], b[key])
        print(f"{precision} {key}: gpu-vs-oracle {err:.2e}, oracle one-ulp spread {spread:.2e}", flush=True)
        floor = mixed_floor(key) if precision == "mixed" else 1e-12
        assert err <= max(10.0 * spread, floor), key


def test_vb_drive_newtonian_run(gpu_lib, tmp_path):
    """vb_drive run DIR fp64 3 newtonian on the generated dam break: ten-column PART files and a
    TEMP_REC.dat without stress lines, holding the state the ctypes path reaches."""
    case = _case(1, tmp_path)
    ind = (tmp_path / "INDAT").read_text().replace("      0.0099999998 :dt_out", "      0.0000000001 :dt_out")
    (tmp_path / "INDAT").write_text(ind)
    exe = os.path.join(os.path.dirname(solver.__file__), "vb_drive")
    r = subprocess.run([exe, "run", str(tmp_path), "fp64", "3", "newtonian"], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0
    G = solver.Solver(case)
    for _ in range(3):
        G.div_free()
    st = G.get_state()
    keep = st["inbin"] == 0
    ns, conc = G.free_surface()
    part = refio.read_part(str(tmp_path / "data_directory" / "PART2"), case.npfb, newtonian=True)
    for key in ("x", "y", "u", "v", "p"):
        assert np.array_equal(part[key], st[key][keep]), key
    assert np.array_equal(part["n_surf"], ns[keep]) and np.array_equal(part["conc"], conc[keep])
    assert not np.any(part["a_out"])
    hd, cols = refio.read_temp_rec(str(tmp_path / "TEMP_REC.dat"), newtonian=True)
    assert hd.itime == 3 and np.array_equal(cols["x"], st["x"][keep])
