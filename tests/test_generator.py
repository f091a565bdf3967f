"""vebisph_amd.generator (the eight cases of the reference's generator, source/gen/datclass.F90) against
the reference's own generator: built here from its sources (oracle/_ref/gen2D, oracle/Makefile) when
they are present, and through the fixtures it wrote (tests/golden/generator_cases.json) everywhere.

Bit for bit: INDAT's text (its fixed formats round h, dx, dt_out and the groups to ten decimals, the
values a run uses), every parameter, patch, circle and particle value as the readers parse them."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from vebisph_amd import generator, refio

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden  # noqa: E402

FIXTURES = json.load(open(os.path.join(HERE, "golden", "generator_cases.json")))


@pytest.mark.parametrize("k", range(1, 9))
def test_generator_reproduces_reference_fixtures(k, tmp_path):
    generator.write_run_dir(k, str(tmp_path))
    assert make_golden.run_dir_record(str(tmp_path)) == FIXTURES[str(k)]


@pytest.mark.skipif(not os.path.exists(make_golden.GEN2D), reason="reference generator not built (no sources)")
@pytest.mark.parametrize("k", range(1, 9))
def test_generator_equals_reference_generator(k, tmp_path):
    ref, ours = tmp_path / "ref", tmp_path / "ours"
    ref.mkdir()
    subprocess.run([make_golden.GEN2D], input=f"{k}\n", text=True, cwd=ref, capture_output=True, check=True)
    generator.write_run_dir(k, str(ours))
    assert (ref / "INDAT").read_text() == (ours / "INDAT").read_text()
    a, b = refio.read_case_dir(str(ref)), refio.read_case_dir(str(ours))
    assert (a.tmax, a.dt_out, a.dt, a.alpha_evss) == (b.tmax, b.dt_out, b.dt, b.alpha_evss)
    for key in ("h", "dx", "dv", "Ra", "Pr", "El", "beta", "body_force", "dt_coef_cfl", "dt_coef_visc",
                "ve_nonlinearity", "i_open_domain", "i_shift_coeff", "nbS", "npfb"):
        assert getattr(a.case, key) == getattr(b.case, key), key
    for key in ("x", "y", "u", "v", "theta", "b_node", "b_type", "b_vel", "b_thermal", "b_periodic_parent",
                "c_centre", "c_radius", "c_omega", "c_vel", "c_thermal"):
        va, vb = getattr(a.case, key), getattr(b.case, key)
        assert (va is None) == (vb is None), key
        if va is not None:
            assert np.array_equal(va, vb), key


def test_generator_cli(tmp_path):
    """python -m vebisph_amd.generator CASE DIR writes a run directory the readers accept."""
    r = subprocess.run([sys.executable, "-m", "vebisph_amd.generator", "3", str(tmp_path)], capture_output=True,
                       text=True, cwd=os.path.dirname(HERE))
    assert r.returncode == 0, r.stderr
    rd = refio.read_case_dir(str(tmp_path))
    assert rd.case.npfb == 932 and rd.case.nbS == 62
    assert subprocess.run([sys.executable, "-m", "vebisph_amd.generator", "9", str(tmp_path)],
                          capture_output=True, cwd=os.path.dirname(HERE)).returncode == 2


@pytest.mark.parametrize("name, factor", [("cylinder_2M", 2), ("cylinder_4M", 4)])
def test_cylinder_at_size_is_the_same_flow(name, factor):
    """The larger cylinder workloads (throughput at size) are the 1M flow at sqrt(factor) times the
    resolution: the same domain, circle, body force and fluid parameters, the spacing and smoothing
    length scaled down, and about factor times the particles."""
    from vebisph_amd import cases
    a, b = cases.make("cylinder_1M"), cases.make(name)
    for key in ("Wi", "beta", "Re", "Ra", "Pr", "El", "nb_circles", "i_open_domain", "i_shift_coeff", "const_mod"):
        assert getattr(a, key) == getattr(b, key), key
    np.testing.assert_array_equal(np.asarray(a.body_force), np.asarray(b.body_force))
    np.testing.assert_array_equal(np.asarray(a.c_centre), np.asarray(b.c_centre))
    np.testing.assert_array_equal(np.asarray(a.c_radius), np.asarray(b.c_radius))
    r = np.sqrt(factor)
    assert abs(a.dx / b.dx - r) < 1e-3 * r and abs(a.h / b.h - r) < 1e-3 * r
    assert abs(b.npfb / a.npfb - factor) < 0.01 * factor
    assert np.min(b.x) >= np.min(a.x) - a.dx and np.max(b.x) <= np.max(a.x) + a.dx
