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
