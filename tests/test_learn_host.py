"""CPU checks of the learned planner's host side: incidence graphs and parameter layout."""

import numpy as np

from tensorsat_amd.learn import graph as G
from tensorsat_amd.learn.model import HID, LAYERS, init_params, param_layout
from tensorsat_amd.sat import CNFInstance


def test_incidence_csr_is_consistent_and_transposed():
    cnfs = [CNFInstance.random(k, 12 + 3 * k, 5 * k + 20, rng=k) for k in (3, 4, 5)]
    g = G.build(cnfs)
    assert g.num_vars == sum(c.num_vars for c in cnfs)
    assert g.num_clauses == sum(len(c.clauses) for c in cnfs)
    edges = set()
    ci = 0
    for i, cnf in enumerate(cnfs):
        for cl in cnf.clauses:
            got = [(int(g.c2v_col[e]), int(g.c2v_sign[e])) for e in range(g.c2v_rowptr[ci], g.c2v_rowptr[ci + 1])]
            want = [(abs(l) - 1 + int(g.var_offset[i]), 1 if l > 0 else -1) for l in cl]
            assert got == want
            edges.update((ci, v, s) for v, s in want)
            ci += 1
    tr = set()
    for v in range(g.num_vars):
        for e in range(g.v2c_rowptr[v], g.v2c_rowptr[v + 1]):
            tr.add((int(g.v2c_col[e]), v, int(g.v2c_sign[e])))
    assert tr == edges
    assert g.var_feat.shape == (g.num_vars, G.FEAT) and g.clause_feat.shape == (g.num_clauses, G.FEAT)
    # occurrence features count literals
    deg = np.diff(g.v2c_rowptr)
    np.testing.assert_allclose(g.var_feat[:, 2] * 20.0, deg, rtol=1e-6)


def test_parameter_layout():
    layout, total = param_layout()
    assert total == 2 * (32 * HID + HID) + LAYERS * 2 * (3 * HID * HID + HID) + HID + 1
    p = init_params(0)
    assert p.shape == (total,) and np.isfinite(p).all()
    offs = sorted(off for off, _ in layout.values())
    assert offs[0] == 0 and len(set(offs)) == len(offs)
