"""
GPU parity of the drop-in box class (DeviceRel) and the contraction classes running on
it, against the oracle and the reference's own goldens and verdict tests.
"""

import itertools
import json
from pathlib import Path

import numpy as np
import pytest
import xxhash

from oracle import brute, finrel as F
from tensorsat_amd import _autoray as AR
from tensorsat_amd import gates as G
from tensorsat_amd.contractions import PlannedContraction, SimpleContraction
from tensorsat_amd.network import Diagram, Wiring
from tensorsat_amd.sat import CNFInstance

pytestmark = pytest.mark.gpu

HERE = Path(__file__).parent
GOLD = json.loads((HERE / "golden" / "reference_goldens.json").read_text())


@pytest.fixture(scope="module")
def DR():
    from tensorsat_amd.rel import DeviceRel

    return DeviceRel


def _rand_case(rng, max_wires=10, max_rank=8, dims=(2,)):
    nw = int(rng.integers(2, max_wires + 1))
    ext = {w: int(rng.choice(dims)) for w in range(nw)}
    lw = [int(x) for x in rng.choice(nw, size=int(rng.integers(1, min(nw, max_rank) + 1)), replace=False)]
    rw = [int(x) for x in rng.choice(nw, size=int(rng.integers(1, min(nw, max_rank) + 1)), replace=False)]
    allw = sorted(set(lw) | set(rw))
    ow = [w for w in allw if rng.random() < 0.45]
    rng.shuffle(ow)
    dens = float(rng.choice([0.2, 0.7, 1.0]))
    lhs = (rng.random(tuple(ext[w] for w in lw)) < dens).astype(np.uint8)
    rhs = (rng.random(tuple(ext[w] for w in rw)) < dens).astype(np.uint8)
    return lhs, lw, rhs, rw, ow


@pytest.mark.parametrize("mode", ["GT0", "MOD256"])
def test_contract2_random_vs_oracle(DR, mode):
    rng = np.random.default_rng(11)
    with DR.mode_scope(mode):
        for case in range(120):
            lhs, lw, rhs, rw, ow = _rand_case(rng)
            got = DR._contract2(DR(lhs), lw, DR(rhs), rw, ow).tensor
            ref = F.contract2(lhs, lw, rhs, rw, ow, mode)
            np.testing.assert_array_equal(got, ref, err_msg=f"case {case}")


def test_contract2_mod256_divergence(DR):
    a = np.ones((2,) * 8, np.uint8)
    with DR.mode_scope("MOD256"):
        assert not bool(DR.contract2(DR(a), range(8), DR(a), range(8), []))
    assert bool(DR.contract2(DR(a), range(8), DR(a), range(8), []))


def test_contract2_wider_finsets_vs_oracle(DR):
    rng = np.random.default_rng(3)
    for case in range(60):
        lhs, lw, rhs, rw, ow = _rand_case(rng, max_wires=6, max_rank=5, dims=(1, 2, 3, 5))
        got = DR._contract2(DR(lhs), lw, DR(rhs), rw, ow).tensor
        np.testing.assert_array_equal(got, F.contract2(lhs, lw, rhs, rw, ow, "GT0"), err_msg=f"case {case}")


def test_contract2_repeated_wires(DR):
    rng = np.random.default_rng(8)
    t = (rng.random((2, 2, 2)) < 0.6).astype(np.uint8)
    u = (rng.random((2, 2)) < 0.6).astype(np.uint8)
    # repeated wire inside lhs (diagonal) and repeated out wire (diagonal embedding)
    got = DR.contract2(DR(t), [0, 1, 0], DR(u), [1, 2], [2, 0, 2]).tensor
    ref = F.contract2(t, [0, 1, 0], u, [1, 2], [2, 0, 2], "GT0")
    np.testing.assert_array_equal(got, ref)


def test_rewire_vs_oracle(DR):
    rng = np.random.default_rng(5)
    for case in range(80):
        r = int(rng.integers(1, 7))
        t = (rng.random((2,) * r) < 0.5).astype(np.uint8)
        k = int(rng.integers(0, r + 3))
        ports = [int(p) for p in rng.integers(0, r, size=k)]
        np.testing.assert_array_equal(DR(t).rewire(ports).tensor, F.rewire(t, ports, "GT0"), err_msg=str(ports))
        # MOD256 follows the reference's rewire_array literally, unsorted and repeated
        # ports included (its slices are added without a transpose)
        with DR.mode_scope("MOD256"):
            np.testing.assert_array_equal(DR(t).rewire(ports).tensor, F.rewire(t, ports, "MOD256"),
                                          err_msg=f"MOD256 {ports}")


def test_rewire_mod256_unsorted_ports_wide(DR):
    """Literal MOD256 rewire on tensors wide enough for the device's ∃-reduction kernels
    (discarding >= 6 axes) with kept ports in decreasing order."""
    rng = np.random.default_rng(23)
    for case in range(12):
        r = int(rng.integers(8, 13))
        t = (rng.random((2,) * r) < 0.02).astype(np.uint8)
        kept = [int(p) for p in rng.choice(r, size=int(rng.integers(1, 4)), replace=False)]
        ports = sorted(kept, reverse=True) + ([kept[0]] if case % 3 == 0 else [])
        with DR.mode_scope("MOD256"):
            got = DR(t).rewire(ports).tensor
        np.testing.assert_array_equal(got, F.rewire(t, ports, "MOD256"), err_msg=f"case {case} {ports}")


def test_contract2_repeated_out_wires_mod256(DR):
    """fin_rel.py:281-287: repeated out wires go through rewire_array relative to the
    sorted einsum output, in MOD256 reproduced literally."""
    rng = np.random.default_rng(29)
    for case in range(40):
        lhs, lw, rhs, rw, ow = _rand_case(rng, max_wires=7, max_rank=5)
        if not ow:
            continue
        ow = list(ow) + [ow[int(rng.integers(0, len(ow)))]]
        rng.shuffle(ow)
        with DR.mode_scope("MOD256"):
            got = DR._contract2(DR(lhs), lw, DR(rhs), rw, ow).tensor
        np.testing.assert_array_equal(got, F.contract2(lhs, lw, rhs, rw, ow, "MOD256"), err_msg=f"case {case} {ow}")


def test_notebook_box_goldens(DR):
    and_, xor_, or_ = DR(G.and_), DR(G.xor_), DR(G.or_)
    assert sorted(and_.rewire([0, 2, 1]).to_set()) == [tuple(p) for p in GOLD["gates"]["and_rewire_021_set"]]
    assert xor_.rewire([0, 2, 1]) == xor_
    cnot = DR.from_callable((2, 2), (2, 2), lambda a, b: (a, a ^ b))
    assert cnot.rewire([0, 1, 3]) == xor_
    assert xor_.rewire([0, 1, 0, 2]) == cnot
    bit = DR(G.bit_unk).shape[0]
    assert DR.contract2(DR.spider(bit, 3), [0, 2, 4], xor_, [4, 1, 3], [0, 1, 2, 3]) == cnot
    f = DR.contract2(or_, [1, 2, 3], and_, [0, 3, 4])
    assert f == DR.from_callable((2, 2, 2), 2, lambda a, b, c: a & (b | c))
    f1 = DR.contract2(or_, [0, 2, 3], or_, [1, 2, 4], out_wires=[0, 1, 2, 3, 4])
    f2 = DR.contract2(f1, [0, 1, 2, 3, 4], and_, [3, 4, 5])
    assert f2 == DR.from_callable((2, 2, 2), 2, lambda a, b, c: (a | c) & (b | c))
    assert DR.prod([DR(G.bit_0), DR(G.bit_1), DR(G.bit_unk)]).to_set() == frozenset(
        tuple(p) for p in GOLD["inputs_relation"]["set"])
    assert hash(and_) == xxhash.xxh64(G.and_.tobytes()).intdigest()


def test_spider_union_stack_getitem_scalar(DR):
    for d, n in [(2, 1), (2, 3), (3, 2), (5, 3), (1, 4)]:
        np.testing.assert_array_equal(DR.spider(d, n).tensor, F.spider(d, n))
    a, b = DR(G.bit_0), DR(G.bit_1)
    assert (a + b) == DR(G.bit_unk)
    st = DR.stack([a, b, DR(G.bit_unk)], 1)
    np.testing.assert_array_equal(st.tensor, F.stack([G.bit_0, G.bit_1, G.bit_unk], 1))
    x = DR(G.or_)
    np.testing.assert_array_equal(x[1, :, :].tensor, G.or_[1])
    np.testing.assert_array_equal(x[(0, slice(None), 1)].tensor, G.or_[0, :, 1])
    with pytest.raises(NotImplementedError):
        x[(slice(0, 1),)]
    assert bool(DR.scalar(True)) and not bool(DR.scalar(False))
    with pytest.raises(ValueError):
        bool(x)
    with pytest.raises(ValueError):
        DR(np.array([2], np.uint8))
    with pytest.raises(ValueError):
        DR.contract2(x, [0, 1], x, [0, 1, 2])


def test_autoray_on_device(DR):
    a, b = DR(G.or_), DR(G.and_)
    r = AR.einsum("abc,dce->abde", a, b)
    np.testing.assert_array_equal(r.tensor, F.contract2(G.or_, [0, 1, 2], G.and_, [3, 2, 5], [0, 1, 3, 5]))
    s = AR.stack([DR(G.bit_0), DR(G.bit_1)], 0)
    np.testing.assert_array_equal(s.tensor, np.eye(2, dtype=np.uint8))


def test_full_adder_and_open_relation(DR):
    fa = GOLD["full_adder_wiring"]
    w = Wiring((2,) * 8, fa["slot_wires_list"], fa["out_wires"])
    d = Diagram(w, tuple(DR(getattr(G, b)) for b in fa["boxes"]))
    for c in (SimpleContraction.using_planner(DR, w), PlannedContraction(DR, w, optimize="auto")):
        assert sorted(c.contract(d).to_set()) == [tuple(x) for x in GOLD["full_adder_relation"]["set"]]
    g = GOLD["all_solutions"][0]
    cnf = CNFInstance.random(g["k"], g["n"], g["m"], rng=g["rng"])
    cd = cnf.diagram()
    box = PlannedContraction(DR, cd.wiring).contract(cd)
    assert sorted(box.to_set()) == sorted(tuple(s) for s in g["solutions"])


def _is_satisfiable(DR, cnf):
    """examples/test_sat_solver.py:5-9 with the device box class."""
    d = (cnf.inputs(None) >> cnf.diagram()).flatten()
    return bool(PlannedContraction(DR, d.wiring, optimize="auto").contract(d))


def test_reference_verdict_suite(DR):
    verdicts = {k: v for k, v in GOLD["verdicts"].items() if not k.startswith("_")}
    for name, expected in verdicts.items():
        cnf = CNFInstance.from_dimacs((HERE / "golden" / "kissat" / "cnf" / f"{name}.cnf").read_text())
        print(f"verdict {name} n={cnf.num_vars}", flush=True)
        assert _is_satisfiable(DR, cnf) == expected, name


def test_unit6_commented_out_in_reference_is_unsat(DR):
    # examples/test_sat_solver.py:182-185 lists unit6 as UNSAT but skips it (it hangs there)
    cnf = CNFInstance.from_dimacs((HERE / "golden" / "kissat" / "cnf" / "unit6.cnf").read_text())
    assert _is_satisfiable(DR, cnf) is False
    assert not brute.is_sat(cnf.num_vars, cnf.clauses)


@pytest.mark.parametrize("g", GOLD["assignment_search"])
def test_assignment_search_traces(DR, g):
    """notebooks 01 cell 113 / 02 cell 36: one contraction reused for 2n candidates."""
    cnf = CNFInstance.random(g["k"], g["n"], g["m"], rng=g["rng"])
    cd = cnf.diagram()
    contraction = PlannedContraction(DR, (cnf.inputs(None) >> cd).flatten().wiring, optimize="auto")
    n = cnf.num_vars
    assignment = ""
    for i in range(n):
        found = None
        for b in "01":
            cand = assignment[:i] + b + "?" * (n - i - 1)
            if contraction.contract((cnf.inputs(cand) >> cd).flatten()):
                found = b
                break
        assert found is not None
        assignment += found
    assert assignment == g["assignment"]
