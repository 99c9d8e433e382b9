"""
Host logic of the contraction classes (SimpleContraction, PlannedContraction, the
autoray-style backend and the CNF diagram API), run with a CPU test box built on the
oracle, against the reference's goldens and brute force.
"""

import json
from pathlib import Path

import numpy as np
import pytest

from oracle import brute
from tensorsat_amd import _autoray as AR
from tensorsat_amd import gates as G
from tensorsat_amd.contractions import PlannedContraction, SimpleContraction
from tensorsat_amd.contractions.cotengra import CotengraContraction
from tensorsat_amd.network import Diagram, Wiring
from tensorsat_amd.sat import CNFInstance

from oracle_box import OracleRel

GOLD = json.loads((Path(__file__).parent / "golden" / "reference_goldens.json").read_text())


def _full_adder():
    fa = GOLD["full_adder_wiring"]
    w = Wiring((2,) * 8, fa["slot_wires_list"], fa["out_wires"])
    return Diagram(w, tuple(getattr(G, b) for b in fa["boxes"]))


@pytest.mark.parametrize("kind", ["simple", "greedy", "elim", "auto"])
def test_full_adder(kind):
    d = _full_adder()
    if kind == "simple":
        c = SimpleContraction.using_planner(OracleRel, d.wiring)
    else:
        c = PlannedContraction(OracleRel, d.wiring, optimize=kind)
    box = c.contract(d)
    assert sorted(box.to_set()) == [tuple(x) for x in GOLD["full_adder_relation"]["set"]]


def test_inputs_relation_and_assignment_search():
    cnf = CNFInstance.random(3, 3, 6, rng=3)
    d = cnf.inputs("01?") >> cnf.diagram()
    # inputs only: cell 102
    inp = cnf.inputs("01?")
    di = Diagram(inp.wiring, (G.bit_0, G.bit_1, G.bit_unk))
    box = SimpleContraction.using_planner(OracleRel, di.wiring).contract(di)
    assert sorted(box.to_set()) == [tuple(x) for x in GOLD["inputs_relation"]["set"]]
    # assignment search with one contraction reused (cell 113)
    contraction = PlannedContraction(OracleRel, (cnf.inputs(None) >> cnf.diagram()).flatten().wiring)
    n = cnf.num_vars
    assignment = ""
    for i in range(n):
        found = None
        for b in "01":
            cand = assignment[:i] + b + "?" * (n - i - 1)
            if contraction.contract((cnf.inputs(cand) >> cnf.diagram()).flatten()):
                found = b
                break
        assert found is not None
        assignment += found
    assert assignment == GOLD["assignment_search"][0]["assignment"]


def test_all_solutions_open_relation():
    g = GOLD["all_solutions"][0]
    cnf = CNFInstance.random(g["k"], g["n"], g["m"], rng=g["rng"])
    d = cnf.diagram()
    box = CotengraContraction(OracleRel, d.wiring, optimize="auto").contract(d)
    assert sorted(box.to_set()) == sorted(tuple(s) for s in g["solutions"])


@pytest.mark.parametrize("seed", range(8))
def test_sat_decisions_match_brute_force(seed):
    n = 8 + seed
    cnf = CNFInstance.random(3, n, int(round(4.26 * n)), rng=seed)
    d = (cnf.inputs(None) >> cnf.diagram()).flatten()
    v = bool(PlannedContraction(OracleRel, d.wiring, optimize="greedy").contract(d))
    assert v == brute.is_sat(n, cnf.clauses)


def test_sliced_planned_contraction():
    cnf = CNFInstance.random(3, 14, 50, rng=2)
    d = (cnf.inputs(None) >> cnf.diagram(mode="clause")).flatten()
    c = PlannedContraction(OracleRel, d.wiring, optimize="elim", slicing_target=5)
    assert len(c.tree.sliced) >= 1
    assert bool(c.contract(d)) == brute.is_sat(14, cnf.clauses)


def test_autoray_seam():
    a = OracleRel(G.or_)
    b = OracleRel(G.and_)
    # "ab c, d c e -> a b d e" with first-occurrence numbering
    r = AR.einsum("abc,dce->abde", a, b)
    ref = OracleRel.contract2(a, [0, 1, 2], b, [3, 2, 5], [0, 1, 3, 5])
    assert r == ref
    assert AR.shape(a) == (2, 2, 2)
    assert AR.transpose(a, [2, 0, 1]).tensor.shape == (2, 2, 2)
    with pytest.raises(NotImplementedError):
        AR.einsum("ab->a", a)


def test_contraction_validation_errors():
    d = _full_adder()
    c = SimpleContraction.using_planner(OracleRel, d.wiring)
    other = Diagram(Wiring((2,) * 3, [(0, 1, 2)], [0]), (G.or_,))
    with pytest.raises(ValueError):
        c.contract(other)
    with pytest.raises(ValueError):
        SimpleContraction(OracleRel, d.wiring, [(0, 1)])
    with pytest.raises(ValueError):
        SimpleContraction(OracleRel, d.wiring, [(0, 9), (0, 1), (0, 1), (0, 1)])


def test_simple_contraction_args_keep_common_wires():
    # simple.py:150-175: result wires lhs_only + common + rhs_only; discarded wires are
    # removed from the last step where they appear
    w = Wiring((2,) * 4, [(0, 1), (1, 2), (2, 3)], [0, 3])
    c = SimpleContraction(OracleRel, w, [(0, 1), (0, 1)])
    (_, lw1, _, rw1, res1), (_, lw2, _, rw2, res2) = c.contract2_args
    assert res1 == (0, 1, 2)
    # step 2: lhs (2,3), rhs (0,1,2): lhs_only + common + rhs_only = (3,2,0,1); wires 1, 2
    # are dropped here, their last appearance
    assert res2 == (3, 0)
    assert c.contract(Diagram(w, (G.bit_unk.reshape(2, 1).repeat(2, 1), np.eye(2, dtype=np.uint8),
                                  np.ones((2, 2), np.uint8)))).tensor.shape == (2, 2)


def test_literal_rewire_ports_match_rewire_array():
    """DeviceRel's MOD256 port rule (kept single ports in increasing order, repeats on
    the diagonal) reproduces the oracle's literal rewire_array on random unsorted,
    discarded and repeated ports: rewire_array(t, ports) == intended(t, literal ports)."""
    from types import SimpleNamespace

    from oracle import finrel as F
    from tensorsat_amd.rel import DeviceRel

    rng = np.random.default_rng(17)
    checked = 0
    for case in range(300):
        r = int(rng.integers(1, 6))
        t = (rng.random((2,) * r) < 0.5).astype(np.uint8)
        ports = [int(p) for p in rng.integers(0, r, size=int(rng.integers(0, r + 3)))]
        if len(set(ports)) == len(ports) == r:
            continue  # pure permutation: a transpose in both
        lit = DeviceRel._literal_ports(SimpleNamespace(_shape=(2,) * r), ports)
        want = F.rewire(t, ports, "MOD256")
        got = ((F.rewire_exact(t, lit) % 256) != 0).astype(np.uint8)
        np.testing.assert_array_equal(got, want, err_msg=f"case {case} ports {ports}")
        checked += lit != ports
    assert checked > 20  # the rule moved ports in many cases
    # unequal extents fail as numpy's broadcast does in the reference
    with pytest.raises(ValueError):
        DeviceRel._literal_ports(SimpleNamespace(_shape=(2, 3, 2)), [1, 0])
