"""
Binary circuits (``tensorsat/lib/bincirc.py:91-177``) on the host: the builder's flat
wirings and nested counts against the reference's printed values, and the contracted
relations (oracle box class) against the circuits' arithmetic.
"""

import json
from pathlib import Path

import numpy as np
import pytest

from oracle_box import OracleRel
from tensorsat_amd import circuits as C
from tensorsat_amd import gates as G
from tensorsat_amd.contractions import PlannedContraction, SimpleContraction

GOLD = json.loads((Path(__file__).parent / "golden" / "reference_goldens.json").read_text())


def _madd(num_bits: int) -> C.Circuit:
    """``notebooks/01-introduction.ipynb`` cell 77, as written there (the library adder's
    ports taken in the notebook's interleaved order; the recipe adds no outputs)."""
    b = C.CircuitBuilder()
    x = b.add_inputs(num_bits)
    y = b.add_inputs(num_bits)
    z = b.add_inputs(num_bits)
    m = b.apply(C.wallace_multiplier(num_bits), *y, *z)
    (carry0,) = b.apply("bit_0")
    ins = [carry0]
    for xb, mb in zip(x, m):
        ins += [xb, mb]
    b.apply(C.rc_adder(num_bits), *ins)
    return b.build("madd")


def _contract(diagram, kind="auto"):
    if kind == "simple":
        c = SimpleContraction.using_planner(OracleRel, diagram.wiring)
    else:
        c = PlannedContraction(OracleRel, diagram.wiring, optimize=kind)
    return np.asarray(c.contract(diagram).tensor)


def test_full_adder_wiring_is_the_reference_recipe():
    fa = GOLD["full_adder_wiring"]
    assert [list(s) for s in C.full_adder.wiring.slot_wires_list] == fa["slot_wires_list"]
    assert list(C.full_adder.wiring.out_wires) == fa["out_wires"]
    assert list(C.full_adder.diagram.leaf_kinds) == fa["boxes"]


def test_nested_counts_match_the_reference():
    g = GOLD["circuit_summaries"]
    assert C.rc_adder(3).summary() == g["rc_adder_3"]
    assert C.wallace_multiplier(3).summary() == g["wallace_multiplier_3"]
    assert _madd(3).summary() == g["madd_3"]


def test_closed_form_adder_is_the_full_adder_golden():
    """The arithmetic relation of a 1-bit ripple-carry adder (ports a, b, c_in, s, c_out)
    is the reference's full-adder relation (cell 92)."""
    got = sorted(map(tuple, np.argwhere(C.adder_relation(1)).tolist()))
    assert got == [tuple(x) for x in GOLD["full_adder_relation"]["set"]]


@pytest.mark.parametrize("n", [1, 2, 3, 4])
@pytest.mark.parametrize("kind", ["auto", "greedy", "simple"])
def test_adder_relation(n, kind):
    d = C.rc_adder(n).diagram
    np.testing.assert_array_equal(_contract(d, kind), C.adder_relation(n))


@pytest.mark.parametrize("n", [1, 2, 3])
def test_multiplier_relation(n):
    circ, weights = C.wallace_columns(n)
    assert weights == list(range(len(weights)))
    assert len(weights) >= 2 * n - (n == 1)
    np.testing.assert_array_equal(_contract(circ.diagram), C.multiplier_relation(n))


def test_multiplier_inverted_gives_factor_pairs():
    """Fixing the product ports of a 4-bit multiplier to 143 leaves the relation of the
    factor pairs 11 x 13 and 13 x 11."""
    circ, weights = C.wallace_columns(4)
    prod = {8 + k: (143 >> w) & 1 for k, w in enumerate(weights)}
    rel = _contract(C.constrain(circ, prod))
    pairs = sorted(
        (sum(b << i for i, b in enumerate(t[:4])), sum(b << i for i, b in enumerate(t[4:])))
        for t in map(tuple, np.argwhere(rel).tolist())
    )
    assert pairs == [(11, 13), (13, 11)]


def test_depth_two_composition():
    """x + (y*z mod 2^n) through a multiplier nested inside the builder next to an adder
    (depth 2), against the arithmetic."""
    n = 2
    b = C.CircuitBuilder()
    x = b.add_inputs(n)
    y = b.add_inputs(n)
    z = b.add_inputs(n)
    m = b.apply(C.wallace_multiplier(n), *y, *z)
    (carry0,) = b.apply("bit_0")
    outs = b.apply(C.rc_adder(n), *x, *m[:n], carry0)
    b.add_outputs(outs)
    circ = b.build("madd2")
    assert circ.depth == 2 and circ.num_ports == 3 * n + n + 1
    rel = _contract(circ.diagram)
    want = np.zeros((2,) * circ.num_ports, dtype=np.uint8)
    for xv in range(1 << n):
        for yv in range(1 << n):
            for zv in range(1 << n):
                s = xv + ((yv * zv) % (1 << n))
                bits = [(v >> i) & 1 for v in (xv, yv, zv) for i in range(n)]
                bits += [(s >> i) & 1 for i in range(n + 1)]
                want[tuple(bits)] = 1
    np.testing.assert_array_equal(rel, want)


def test_ports_joined_inside_a_sub_circuit_share_one_wire():
    """A sub-circuit whose ports are all one inner wire (a three-way copy) joins the outer
    wires it is applied to: the lowest of them names the joined wire, and contracting
    gives the equality relation."""
    inner = C.CircuitBuilder()
    (w,) = inner.add_inputs(1)
    inner.add_outputs([w, w])
    copy3 = inner.build("copy3")
    assert copy3.wiring.out_wires == (0, 0, 0) and copy3.wiring.num_slots == 0
    b = C.CircuitBuilder()
    p, q = b.add_inputs(2)
    (r,) = b.apply(copy3, q, p)
    (s,) = b.apply("not_", r)
    b.add_outputs([s])
    circ = b.build("joined")
    # q and p (wires 1 and 0) and the new wire 2 become wire 0; not_'s output is wire 1
    assert circ.wiring.out_wires == (0, 0, 1)
    assert circ.wiring.slot_wires_list == ((0, 1),)
    rel = _contract(circ.diagram)
    want = np.zeros((2,) * 3, dtype=np.uint8)
    want[0, 0, 1] = want[1, 1, 0] = 1
    np.testing.assert_array_equal(rel, want)


def test_errors():
    with pytest.raises(ValueError):
        C.rc_adder(0)
    with pytest.raises(ValueError):
        C.wallace_multiplier(-1)
    b = C.CircuitBuilder()
    a = b.add_inputs(2)
    with pytest.raises(ValueError):
        b.apply("and_", *a, 0, 1)
    with pytest.raises(ValueError):
        b.apply("nand_", *a)
    with pytest.raises(ValueError):
        b.add_outputs([5])
    with pytest.raises(ValueError):
        C.constrain(C.full_adder, {7: 1})
    with pytest.raises(ValueError):
        C.constrain(C.full_adder, {0: 2})
    assert G.GATES["and_"].ndim == 3
