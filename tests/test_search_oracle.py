"""
Pins the oracle's restatement of the device cube search (``oracle/search.c``): its
verdicts against brute force and the independent DPLL, and every traced leaf relation
against the ``FinRel`` restatement (``oracle.finrel``: the residual clause tensors
multiplied with every wire kept, then reduced over all ports).
"""

import numpy as np
import pytest

from oracle import brute, dpll
from oracle import finrel as F
from oracle import search as OS
from tensorsat_amd.sat import CNFInstance
from tensorsat_amd.workloads import batched_3sat


def _random_mixed(rng, n):
    m = int(rng.integers(0, 5 * n))
    cls = []
    for _ in range(m):
        k = int(rng.choice([1, 2, 3, 3, 3]))
        cls.append(tuple(int(rng.integers(1, n + 1)) * (1 if rng.random() < 0.5 else -1) for _ in range(k)))
    if rng.random() < 0.05:
        cls.append(())
    return CNFInstance(n, cls)


def test_search_verdicts_match_brute_force():
    rng = np.random.default_rng(5)
    cnfs = [_random_mixed(rng, int(rng.integers(1, 17))) for _ in range(400)]
    cnfs += [CNFInstance.random(3, n, int(round(r * n)), rng=40 + n) for n in range(8, 21) for r in (3.5, 4.26, 5.0)]
    for width in (0, 3, 6):
        for ex in (False, True):
            st = OS.search(cnfs, width=width, exhaustive=ex)
            want = np.array([brute.is_sat(c.num_vars, c.clauses) for c in cnfs], dtype=np.int32)
            np.testing.assert_array_equal(st["verdict"], want, err_msg=f"width {width} exhaustive {ex}")


def test_search_verdicts_match_dpll_on_batched_workload():
    cnfs = batched_3sat(512)
    st = OS.search(cnfs, width=6, threads=4)
    want = np.array([dpll.solve(c.num_vars, c.clauses) for c in cnfs], dtype=np.int32)
    np.testing.assert_array_equal(st["verdict"], want)
    # host threads split instances, never a search: statistics are identical
    one = OS.search(cnfs[:64], width=6, threads=1)
    for f in ("verdict", "nodes", "leaves", "conflicts", "digest"):
        np.testing.assert_array_equal(st[f][:64], one[f])


def _leaf_relation_by_finrel(cnf, rec, mode):
    """The leaf's residual network rebuilt from its assignment: each clause without a
    true literal becomes its relation over its free variables (one 0 at the falsifying
    values), multiplied by FinRel._contract2 with every wire kept (out wires = the
    residual variables, ascending)."""
    asg = [(int(rec["assigned"][v >> 6]) >> (v & 63)) & 1 for v in range(cnf.num_vars)]
    val = [(int(rec["value"][v >> 6]) >> (v & 63)) & 1 for v in range(cnf.num_vars)]
    rel, wires = np.ones((), np.uint8), []
    for cl in cnf.clauses:
        lits = list(dict.fromkeys(cl))
        if any(-l in lits for l in lits):
            continue
        if any(asg[abs(l) - 1] and val[abs(l) - 1] == (l > 0) for l in lits):
            continue  # satisfied: the all-ones tensor after slicing
        free = [l for l in lits if not asg[abs(l) - 1]]
        t = np.ones((2,) * len(free), np.uint8)
        t[tuple(0 if l > 0 else 1 for l in free)] = 0
        fw = [abs(l) - 1 for l in free]
        out = sorted(set(wires) | set(fw))
        rel = F.contract2(rel, wires, t, fw, out, mode)
        wires = out
    return rel, wires


@pytest.mark.parametrize("mode", ["GT0", "MOD256"])
def test_search_leaf_relations_match_finrel(mode):
    cnfs = [c for c in batched_3sat(96) if c.num_vars <= 40][:24]
    st, trace = OS.search(cnfs, width=6, exhaustive=True, trace=4000)
    assert len(trace) >= 200
    for rec in trace:
        cnf = cnfs[int(rec["instance"])]
        rel, wires = _leaf_relation_by_finrel(cnf, rec, mode)
        r = int(rec["rank"])
        assert wires == sorted(int(v) for v in rec["vars"][:r])
        # entry e: bit j of e is the value of wires[j] (numpy axis order is wires[0] first)
        bits = 0
        for e in range(1 << r):
            idx = tuple((e >> j) & 1 for j in range(r))
            bits |= int(rel[idx] if r else rel) << e
        assert bits == int(rec["bits"])
        assert int(F.rewire(rel, [], mode)) == int(bits != 0)
    # the digest is the sum of the leaves' mixes when every leaf was traced
    full = np.flatnonzero(np.bincount(trace["instance"], minlength=len(cnfs)) == st["leaves"])
    assert full.size > 5
    for i in full:
        mine = trace[trace["instance"] == i]
        assert sum(OS.leaf_mix(int(b), int(r)) for b, r in zip(mine["bits"], mine["rank"])) % 2**64 == int(
            st["digest"][i])


def test_search_rejects_bad_input():
    with pytest.raises(ValueError):
        OS.search([CNFInstance(4, [(1, 2, 3, 4)])])  # more than 3 distinct literals
    with pytest.raises(ValueError):
        OS.search([CNFInstance(3, [(1, 2)])], width=7)


def test_lex_witness_is_the_smallest_model():
    """order="lex": the witness is the lexicographically smallest satisfying assignment
    (variables in index order, 0 before 1), by enumeration; and it satisfies."""
    rng = np.random.default_rng(13)
    cnfs = [_random_mixed(rng, int(rng.integers(1, 15))) for _ in range(300)]
    cnfs += [CNFInstance.random(3, n, int(round(r * n)), rng=70 + n) for n in range(8, 19) for r in (3.0, 4.0)]
    for width in (0, 3, 6):
        st, ws = OS.search(cnfs, width=width, order="lex", witness=True)
        for c, v, w in zip(cnfs, st["verdict"], ws):
            want = None
            for a in range(1 << c.num_vars):  # a's bit n-1-j is variable j: lexicographic order
                bits = [(a >> (c.num_vars - 1 - j)) & 1 for j in range(c.num_vars)]
                if all(any((l > 0) == bool(bits[abs(l) - 1]) for l in cl) for cl in c.clauses):
                    want = "".join(map(str, bits))
                    break
            assert (v == 1) == (want is not None)
            assert w == want, (c.clauses, w, want)


def test_lex_witness_reference_golden_traces():
    """The reference's assignment searches (notebook 01 cell 113, 02 cell 36)."""
    import json
    from pathlib import Path

    gold = json.loads((Path(__file__).parent / "golden" / "reference_goldens.json").read_text())
    for g in gold["assignment_search"]:
        cnf = CNFInstance.random(g["k"], g["n"], g["m"], rng=g["rng"])
        if max(len(c) for c in cnf.clauses) > 3:
            continue  # the device search packs at most 3 literals per clause
        _, ws = OS.search([cnf], width=6, order="lex", witness=True)
        assert ws[0] == g["assignment"]
