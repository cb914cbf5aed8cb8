"""The oracle against the reference's own known answers.

test/<case>/CMakeLists.txt:33 holds the L1 density error each reference test
must print (compared as a 13-significant-digit string, test_driver.py:63).
These are the only golden values the reference ships; they pin the whole
chain: ICs at centroids, geometry, flux, BCs, dt rule, RK3, error definition.
"""
import json
import os

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = json.load(open(os.path.join(HERE, "golden", "reference_known_answers.json")))["cases"]


def _run(orc, case):
    ptype = orc.PROBLEM_KEYS[case["problem"]]
    pb = orc.problem(ptype, case["dimensions"])
    tend = case["end"]
    if tend is None:  # problem.cpp:187-202 defaults
        tend = 1.0 if ptype in (orc.VORTEX_XY, orc.VORTEX_ZX, orc.VORTEX_YZ) else 4.0
    return orc.run_reference_case(ptype, case["dimensions"], case["nCells"], tend, case["CFL"])


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_known_answer(orc, case):
    l1, l2, steps, s, U = _run(orc, case)
    expected = float(case["expected"])
    rel = abs(l1 - expected) / expected
    assert rel < 1e-11, (case["name"], "%.12e" % l1, case["expected"], rel)
    # the string the reference's test driver compares (main.cpp:572)
    assert "%.12e" % l1 == case["expected"]
    assert steps > 10
