"""
Pins the oracle (``oracle/``) before it is trusted as the checker.

Every oracle function is checked against at least one of: a golden value held by the
reference (notebook outputs / asserts), a closed form, or brute force, chosen so that a
dropped term, wrong sign, swapped index or transposed operand fails.
"""

import itertools
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import brute, finrel as F
from oracle import executor as OX
from tensorsat_amd import gates as G

GOLD = json.loads((Path(__file__).parent / "golden" / "reference_goldens.json").read_text())


def _rel(points, rank):
    t = np.zeros((2,) * rank, dtype=np.uint8)
    for p in points:
        t[tuple(p)] = 1
    return t


def _func(n_in, n_out, f):
    t = np.zeros((2,) * (n_in + n_out), dtype=np.uint8)
    for idx in itertools.product(range(2), repeat=n_in):
        out = f(*idx)
        out = out if isinstance(out, tuple) else (out,)
        t[idx + out] = 1
    return t


CNOT = _func(2, 2, lambda a, b: (a, a ^ b))


def test_gate_goldens():
    np.testing.assert_array_equal(G.and_, np.array(GOLD["gates"]["and_"], dtype=np.uint8))
    # xor/or/not by their truth tables
    for a, b in itertools.product(range(2), repeat=2):
        assert G.or_[a, b, a | b] == 1 and G.or_[a, b, 1 - (a | b)] == 0
        assert G.xor_[a, b, a ^ b] == 1 and G.xor_[a, b, 1 - (a ^ b)] == 0
    assert G.not_.tolist() == [[0, 1], [1, 0]]
    assert G.bit_0.tolist() == [1, 0] and G.bit_1.tolist() == [0, 1] and G.bit_unk.tolist() == [1, 1]


@pytest.mark.parametrize("mode", F.MODES)
def test_rewire_goldens(mode):
    # 01-introduction.ipynb cells 42, 44, 46, 48
    r = F.rewire(G.and_, [0, 2, 1], mode)
    assert sorted(map(tuple, np.argwhere(r).tolist())) == [tuple(p) for p in GOLD["gates"]["and_rewire_021_set"]]
    np.testing.assert_array_equal(F.rewire(G.xor_, [0, 2, 1], mode), G.xor_)
    np.testing.assert_array_equal(F.rewire(CNOT, [0, 1, 3], mode), G.xor_)
    np.testing.assert_array_equal(F.rewire(G.xor_, [0, 1, 0, 2], mode), CNOT)


@pytest.mark.parametrize("mode", F.MODES)
def test_contract2_goldens(mode):
    # cell 51
    np.testing.assert_array_equal(F.contract2(F.spider(2, 3), [0, 2, 4], G.xor_, [4, 1, 3], [0, 1, 2, 3], mode), CNOT)
    # cell 53 (default out wires = sorted symmetric difference)
    np.testing.assert_array_equal(F.contract2(G.or_, [1, 2, 3], G.and_, [0, 3, 4], None, mode),
                                  _func(3, 1, lambda a, b, c: a & (b | c)))
    # cell 55
    f1 = F.contract2(G.or_, [0, 2, 3], G.or_, [1, 2, 4], [0, 1, 2, 3, 4], mode)
    np.testing.assert_array_equal(f1, _func(3, 2, lambda a, b, c: (a | c, b | c)))
    f2 = F.contract2(f1, [0, 1, 2, 3, 4], G.and_, [3, 4, 5], None, mode)
    np.testing.assert_array_equal(f2, _func(3, 1, lambda a, b, c: (a | c) & (b | c)))


def test_mod256_divergence_closed_form():
    # all-ones rank 8 x all-ones rank 8 over 8 contracted wires: count 256
    a = np.ones((2,) * 8, dtype=np.uint8)
    assert F.contract2(a, range(8), a, range(8), [], "MOD256") == 0
    assert F.contract2(a, range(8), a, range(8), [], "GT0") == 1
    # 255 terms (one zero) -> nonzero in both; 512 -> 0 in MOD256
    b = a.copy()
    b[(0,) * 8] = 0
    assert F.contract2(a, range(8), b, range(8), [], "MOD256") == 1
    c = np.ones((2,) * 9, dtype=np.uint8)
    assert F.contract2(c, range(9), c, range(9), [], "MOD256") == 0


def test_numpy_wrap_semantics_behind_mod256():
    # the two numpy behaviours the MOD256 restatement relies on (fin_rel.py:279, einsum on uint8)
    assert int(np.einsum("i->", np.ones(300, np.uint8))) == 44
    assert np.sign(np.array([255, 256, 257, 512], np.uint16), dtype=np.uint8).tolist() == [1, 0, 1, 0]


def test_widening_thresholds():
    x = np.zeros(3, np.uint8)
    assert F.safe_contraction_dtype(255, x)[0].dtype == np.uint8
    assert F.safe_contraction_dtype(256, x)[0].dtype == np.uint16
    assert F.safe_contraction_dtype(65536, x)[0].dtype == np.uint32
    assert F.safe_contraction_dtype(2**32, x)[0].dtype == np.uint64


@pytest.mark.parametrize("seed", range(40))
def test_contract2_literal_equals_closed_form(seed):
    """Literal MOD256 restatement (widening + wrapped einsum + sign) == (count mod 256) != 0,
    and GT0 == brute-force OR-of-AND, on random 0/1 tensors with batch, contracted and
    lhs-only summed wires."""
    rng = np.random.default_rng(seed)
    nw = int(rng.integers(3, 11))
    lw = list(rng.choice(nw, size=int(rng.integers(1, min(nw, 9) + 1)), replace=False))
    rw = list(rng.choice(nw, size=int(rng.integers(1, min(nw, 9) + 1)), replace=False))
    allw = sorted(set(lw) | set(rw))
    ow = [w for w in allw if rng.random() < 0.4]
    rng.shuffle(ow)
    dens = rng.choice([0.3, 0.9, 1.0])
    lhs = (rng.random((2,) * len(lw)) < dens).astype(np.uint8)
    rhs = (rng.random((2,) * len(rw)) < dens).astype(np.uint8)
    for mode in F.MODES:
        lit = F.contract2(lhs, lw, rhs, rw, ow, mode)
        cf = F.contract2_closed_form(lhs, lw, rhs, rw, ow, mode)
        np.testing.assert_array_equal(lit, cf)
    # brute force GT0 by explicit loops over all wire values
    res = np.zeros((2,) * len(ow), dtype=np.uint8)
    for vals in itertools.product(range(2), repeat=len(allw)):
        env = dict(zip(allw, vals))
        if lhs[tuple(env[w] for w in lw)] and rhs[tuple(env[w] for w in rw)]:
            res[tuple(env[w] for w in ow)] = 1
    np.testing.assert_array_equal(F.contract2(lhs, lw, rhs, rw, ow, "GT0"), res)


def test_rewire_array_reference_quirk_documented():
    """rewire_array adds the remaining axes without transposing: with discarded ports and
    non-increasing kept ports it differs from the intended map (a reference defect)."""
    t = np.zeros((2, 2, 2), np.uint8)
    t[0, 0, 1] = 1  # port0=0, port1=0, port2=1
    quirk = F.rewire(t, [2, 0], "MOD256")  # discard port 1, swap 0 and 2
    exact = F.rewire(t, [2, 0], "GT0")
    assert exact[1, 0] == 1 and exact.sum() == 1
    assert quirk[0, 1] == 1 and quirk.sum() == 1  # transposed


def test_count_contract2_exact_big():
    # 70 contracted bits of ones would overflow uint64; object arithmetic keeps it exact
    a = np.ones((2,) * 4, np.uint8)
    c = F.count_contract2(a, range(4), a, range(4), [])
    assert int(c) == 16


def test_spider_union_stack_getitem_as_scalar():
    s = F.spider(2, 3)
    assert s.sum() == 2 and s[0, 0, 0] == 1 and s[1, 1, 1] == 1
    s3 = F.spider(3, 2)
    assert s3.tolist() == np.eye(3, dtype=np.uint8).tolist()
    np.testing.assert_array_equal(F.union(G.bit_0, G.bit_1), G.bit_unk)
    st = F.stack([G.bit_0, G.bit_1], 0)
    assert st.tolist() == [[1, 0], [0, 1]]
    np.testing.assert_array_equal(F.getitem(CNOT, (1, slice(None), slice(None), slice(None))), CNOT[1])
    with pytest.raises(NotImplementedError):
        F.getitem(CNOT, (slice(0, 1),))
    assert F.as_scalar(np.array(1, np.uint8)) and not F.as_scalar(np.array(0, np.uint8))


def test_prod_inputs_relation():
    # cells 102/104: inputs("01?") == prod([bit_0, bit_1, bit_unk])
    p = F.prod([G.bit_0, G.bit_1, G.bit_unk])
    assert sorted(map(tuple, np.argwhere(p).tolist())) == [tuple(x) for x in GOLD["inputs_relation"]["set"]]


def test_brute_force_counts():
    # #sat prototype: ((a|b)|c) & ((~c|~b)|~a) has 6 models over 3 variables
    clauses = [(1, 2, 3), (-3, -2, -1)]
    assert brute.count(3, clauses) == GOLD["sat_count"]["count"]
    m = brute.satisfying_mask(3, clauses)
    assert m.astype(int).tolist() == GOLD["sat_open_vector"]["vector"]


def test_full_adder_relation_by_oracle():
    # full_adder wiring (cells 61/65) contracted with the oracle == cell 92's relation
    fa = GOLD["full_adder_wiring"]
    boxes = [getattr(G, b) for b in fa["boxes"]]
    t = boxes[0]
    wires = list(fa["slot_wires_list"][0])
    for b, ws in zip(boxes[1:], fa["slot_wires_list"][1:]):
        out = sorted(set(wires) | set(ws))
        t = F.contract2(t, wires, b, ws, out, "GT0")
        wires = out
    res = F.rewire(t, [wires.index(w) for w in fa["out_wires"]], "GT0")
    assert sorted(map(tuple, np.argwhere(res).tolist())) == [tuple(x) for x in GOLD["full_adder_relation"]["set"]]


def test_oracle_record_decoder_matches_product_decoder():
    """The oracle decodes step and job records from the documented byte layout with its
    own code; on real programs (both encodings' cube jobs) it agrees field by field with
    the product's decoder, so neither can misread a field silently."""
    from oracle import records
    from tensorsat_amd import program as PG
    from tensorsat_amd.workloads import batched_3sat

    prog = PG.sat_batch(batched_3sat(24), width=8, max_cubes=3)
    steps = records.decode_steps(prog.steps)
    jobs = records.decode_jobs(prog.jobs)
    assert len(steps) == len(prog.steps) > 100 and len(jobs) == len(prog.jobs)
    for st, rec in zip(steps, prog.steps):
        lab_l, lab_r, lab_o = PG.step_labels(rec)
        assert (st.lhs_off, st.rhs_off, st.out_off) == (int(rec["lhs_off"]), int(rec["rhs_off"]), int(rec["out_off"]))
        assert (st.lhs_rank, st.rhs_rank, st.out_rank, st.sum_rank) == tuple(
            int(rec[k]) for k in ("lhs_rank", "rhs_rank", "out_rank", "sum_rank"))
        assert list(st.lhs_labels) == lab_l and list(st.out_labels) == lab_o
        assert st.norhs == bool(int(rec["flags"]) & PG.STEP_NORHS)
        if not st.norhs:
            assert list(st.rhs_labels) == lab_r
    for jb, rec in zip(jobs, prog.jobs):
        for f in ("instance", "cube", "step_begin", "num_steps", "leaf_begin", "num_leaf_bytes", "result_off",
                  "result_rank", "free_vars", "arena_elems"):
            assert getattr(jb, f) == int(rec[f]), f
    # a corrupted position byte is caught, not decoded into a wrong einsum
    bad = prog.steps.copy()
    k = int(np.flatnonzero(bad["lhs_rank"] >= 2)[0])
    bad["o_lpos"][k, 0] = bad["lhs_rank"][k] + 3 if bad["out_rank"][k] else bad["o_lpos"][k, 0]
    bad["c_lpos"][k, 0] = bad["lhs_rank"][k] + 3 if bad["sum_rank"][k] and not bad["out_rank"][k] else bad["c_lpos"][k, 0]
    if bad["out_rank"][k] or bad["sum_rank"][k]:
        with pytest.raises(ValueError):
            records.decode_steps(bad[k : k + 1])


def test_sharp_counter_matches_brute_force():
    """oracle/sharp.c (component-caching model counter) against enumeration on random
    formulas with unit, duplicate-literal, tautological and empty clauses."""
    from oracle import sharp

    rng = np.random.default_rng(0)
    for t in range(300):
        n = int(rng.integers(1, 15))
        m = int(rng.integers(0, 3 * n + 4))
        cls = []
        for _ in range(m):
            k = int(rng.integers(0, 5)) if rng.random() < 0.1 else int(rng.integers(1, 4))
            cls.append(tuple(int(rng.integers(1, n + 1)) * (1 if rng.random() < 0.5 else -1) for _ in range(k)))
        assert sharp.count(n, cls) == brute.count(n, cls), (n, cls)
    with pytest.raises(ValueError):
        sharp.count(3, [(1, 4)])


def test_sharp_counter_notebook_count_and_tensor_oracle():
    """The #SAT notebook's count 6, and agreement with the oracle's tensor-contraction
    count (a different algorithm) on cube programs of 30-40 variable instances."""
    from oracle import sharp
    from tensorsat_amd import program as PG
    from tensorsat_amd.sat import CNFInstance

    # #sat prototype: ((a|b)|c) & ((~c|~b)|~a) has 6 models over 3 variables
    assert sharp.count(3, [(1, 2, 3), (-3, -2, -1)]) == GOLD["sat_count"]["count"] == 6
    cnfs = [CNFInstance.random(3, n, int(2.5 * n), rng=900 + n) for n in (30, 34, 40)]
    prog = PG.sat_batch(cnfs, width=12)
    assert OX.count(prog) == [sharp.count(c.num_vars, c.clauses) for c in cnfs]
