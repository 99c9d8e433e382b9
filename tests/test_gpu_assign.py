"""
Satisfying assignments on the GPU: the reference's self-reduction over one contraction
(``assign.find_assignment``, 2n input selections on one planned contraction of
``DeviceRel`` boxes) against the reference's golden traces, and the batched device route
(``assign.find_assignments``: the cube search in lexicographic order with witnesses)
against the oracle's restatement, the golden trace, brute force and the first route.
"""

import json
from pathlib import Path

import numpy as np
import pytest

from oracle import brute
from oracle import search as OS
from tensorsat_amd.sat import CNFInstance
from tensorsat_amd.workloads import batched_3sat

pytestmark = pytest.mark.gpu

GOLD = json.loads((Path(__file__).parent / "golden" / "reference_goldens.json").read_text())


def _satisfies(cnf, a):
    return all(any((l > 0) == (a[abs(l) - 1] == "1") for l in cl) for cl in cnf.clauses)


@pytest.mark.parametrize("g", GOLD["assignment_search"])
def test_contraction_loop_reproduces_reference_traces(g):
    from tensorsat_amd.assign import find_assignment

    cnf = CNFInstance.random(g["k"], g["n"], g["m"], rng=g["rng"])
    assert find_assignment(cnf) == g["assignment"]


def test_device_witnesses_match_oracle_and_golden():
    from tensorsat_amd.assign import find_assignments

    cnfs = batched_3sat(1024)
    g = [x for x in GOLD["assignment_search"] if x["k"] == 3][0]
    cnfs.append(CNFInstance.random(3, g["n"], g["m"], rng=g["rng"]))
    got = find_assignments(cnfs)
    _, want = OS.search(cnfs, width=6, order="lex", witness=True, threads=8)
    assert got == want
    assert got[-1] == g["assignment"]
    for c, a in zip(cnfs, got):
        if a is not None:
            assert len(a) == c.num_vars and _satisfies(c, a)
    assert sum(a is not None for a in got) > 500


def test_device_witnesses_equal_the_contraction_loop_and_brute_force():
    from tensorsat_amd.assign import find_assignment, find_assignments

    cnfs = [CNFInstance.random(3, n, int(round(r * n)), rng=500 + n) for n in (8, 10, 12, 14) for r in (3.0, 4.0, 5.0)]
    got = find_assignments(cnfs)
    for c, a in zip(cnfs, got):
        want = None
        if brute.is_sat(c.num_vars, c.clauses):
            want = find_assignment(c)
        assert a == want, c.clauses


def test_lex_order_statistics_match_oracle():
    from tensorsat_amd.search import SearchSolver

    cnfs = batched_3sat(512)
    s = SearchSolver(width=6, split_target=None, order="lex", witness=True, stats=True)
    v = s.run(cnfs)
    ref, ws = OS.search(cnfs, width=6, order="lex", witness=True, threads=8)
    np.testing.assert_array_equal(v, ref["verdict"])
    got = s.instance_stats()
    for f in ("nodes", "leaves", "conflicts", "digest"):
        np.testing.assert_array_equal(got[f], ref[f], err_msg=f)
    assert s.witnesses() == ws
