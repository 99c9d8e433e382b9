"""
Binary circuits (``tensorsat/lib/bincirc.py:122-177``) contracted on the GPU through the
reference's call shape (``Contraction(box_class, wiring).contract(diagram)``) with the
device box class: the ripple-carry adder up to 8 bits (a 2^26-entry relation) and the
Wallace multiplier up to 5 bits, against the circuits' arithmetic and the oracle box
class, in both epilogues; and a multiplier run backwards (product fixed) to factor.
"""

import numpy as np
import pytest

from oracle_box import OracleRel
from tensorsat_amd import circuits as C
from tensorsat_amd.contractions import PlannedContraction, SimpleContraction

pytestmark = pytest.mark.gpu


def _device(diagram, kind="auto", mode="GT0"):
    from tensorsat_amd.rel import DeviceRel

    with DeviceRel.mode_scope(mode):
        if kind == "simple":
            c = SimpleContraction.using_planner(DeviceRel, diagram.wiring)
        else:
            c = PlannedContraction(DeviceRel, diagram.wiring, optimize=kind)
        return np.asarray(c.contract(diagram).tensor)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 6, 8])
def test_adder_on_device(n):
    d = C.rc_adder(n).diagram
    want = C.adder_relation(n)
    print(f"rc_adder({n}): {d}", flush=True)
    for mode in ("GT0", "MOD256"):
        np.testing.assert_array_equal(_device(d, mode=mode), want, err_msg=mode)
    if n <= 4:
        np.testing.assert_array_equal(_device(d, "simple"), want)
        ref = np.asarray(PlannedContraction(OracleRel, d.wiring, optimize="auto").contract(d).tensor)
        np.testing.assert_array_equal(ref, want)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_multiplier_on_device(n):
    d = C.wallace_multiplier(n).diagram
    want = C.multiplier_relation(n)
    print(f"wallace_multiplier({n}): {d}", flush=True)
    for mode in ("GT0", "MOD256"):
        np.testing.assert_array_equal(_device(d, mode=mode), want, err_msg=mode)
    if n <= 3:
        ref = np.asarray(PlannedContraction(OracleRel, d.wiring, optimize="auto").contract(d).tensor)
        np.testing.assert_array_equal(ref, want)


def test_factoring_by_the_inverted_multiplier():
    """6-bit factors of 3233 = 53 x 61: the product ports fixed by bit leaves, the relation
    of the free factor ports contracted on the device."""
    circ, weights = C.wallace_columns(6)
    prod = {12 + k: (3233 >> w) & 1 for k, w in enumerate(weights)}
    rel = _device(C.constrain(circ, prod))
    pairs = sorted(
        (sum(b << i for i, b in enumerate(t[:6])), sum(b << i for i, b in enumerate(t[6:])))
        for t in map(tuple, np.argwhere(rel).tolist())
    )
    assert pairs == [(53, 61), (61, 53)]
