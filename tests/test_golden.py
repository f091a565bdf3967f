"""The C restatement (oracle) against the golden fixtures written by the reference itself
(tests/golden/make_golden.py: oracle/_ref built from the reference's Fortran sources, one thread).
Every phase's outputs must hash identically (bit for bit), every step scalar must be equal, and the
stored final states must match exactly.  Runs without the reference sources."""
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
sys.path.insert(0, GOLDEN)

import make_golden  # noqa: E402

INDEX = json.load(open(os.path.join(GOLDEN, "reference_phases.json")))


@pytest.mark.parametrize("name", sorted(INDEX))
def test_oracle_reproduces_reference_fixtures(name):
    from oracle import oraclewrap
    mk, steps = make_golden.CASES[name]
    case = mk()
    want = INDEX[name]
    assert case.npfb == want["npfb"] and steps == want["steps"]
    O = oraclewrap.OracleSolver(case)
    for s in range(steps):
        for k in range(1, 13):
            O.phase(k)
            got = make_golden.phase_record(O, k)
            assert got == want["phases"][s][k - 1], f"step {s + 1} phase {k}"
    final = O.get_state()
    assert {k: make_golden.digest(v) for k, v in final.items()} == want["final"]
    path = os.path.join(GOLDEN, f"{name}_final.npz")
    if os.path.exists(path):
        ref = np.load(path)
        for k in ref.files:
            assert np.array_equal(np.asarray(final[k]), ref[k]), k


def test_oracle_nnf_reproduces_reference_fixture():
    """The oracle's nnf_models restatement against the reference's own five subroutines
    (nnf_models_mod.F90:8-110, run by make_golden.py over a shear-rate sweep that crosses the
    Bingham threshold and the Cross and Carreau transitions): bit for bit."""
    from oracle import oraclewrap
    ref = np.load(os.path.join(GOLDEN, "nnf_reference.npz"))
    for q, mu in enumerate(ref["mu"]):
        sr = ref[f"sr{q}"]
        assert sr.size == make_golden.nnf_sweep(mu).size
        for model in range(5):
            got = oraclewrap.nnf(model, sr, float(mu))
            assert np.array_equal(got, ref[f"mu{q}_model{model}"]), (mu, model)


def test_nnf_fixture_discriminates_model_constants():
    """A one-constant slip in any model (the exponent 1.5, the Bingham tau_y or alpha, the Cross
    kk or mm, the Carreau kk, mm or aa) moves the fixture's values: the sweep pins every constant."""
    ref = np.load(os.path.join(GOLDEN, "nnf_reference.npz"))
    mu = float(ref["mu"][0]); sr = ref["sr0"]
    slips = {
        1: lambda s: mu * (s + 1.0) ** (1.4 - 1.0),
        2: lambda s: np.where(s <= 1.1e-2 / (mu * 1e2), mu * 1e2, mu + 1.1e-2 / np.maximum(s, 1e-300)),
        3: lambda s: 0.05 * mu + (mu - 0.05 * mu) / (1.0 + (1.1e-3 * s) ** 3.0),
        4: lambda s: 1e-3 + (mu - 1e-3) * (1.0 + (0.14 * s) ** 1.6) ** ((0.1 - 1.0) / 1.6),
    }
    for model, f in slips.items():
        assert not np.allclose(f(sr), ref[f"mu0_model{model}"], rtol=1e-6, atol=0), model
