"""The independent DPLL checker (oracle/dpll.c) against brute force and the reference's
verdict expectations; then contraction verdicts against DPLL on the batched workload."""

import json
from pathlib import Path

import numpy as np

from oracle import brute, dpll
from oracle import executor as OX
from tensorsat_amd import program as PG
from tensorsat_amd.sat import CNFInstance
from tensorsat_amd.workloads import batched_3sat

HERE = Path(__file__).parent
GOLD = json.loads((HERE / "golden" / "reference_goldens.json").read_text())


def test_dpll_vs_brute_force():
    rng = np.random.default_rng(0)
    for i in range(200):
        n = int(rng.integers(3, 16))
        k = int(rng.integers(1, 4))
        m = int(rng.integers(1, 6 * n))
        k = min(k, n)
        from math import comb

        cnf = CNFInstance.random(k, n, min(m, comb(n, k) * 2**k), rng=i)
        assert dpll.solve(n, cnf.clauses) == brute.is_sat(n, cnf.clauses)


def test_dpll_reference_verdicts():
    for name, expected in GOLD["verdicts"].items():
        if name.startswith("_"):
            continue
        cnf = CNFInstance.from_dimacs((HERE / "golden" / "kissat" / "cnf" / f"{name}.cnf").read_text())
        assert dpll.solve(cnf.num_vars, cnf.clauses) == expected, name


def test_cube_verdicts_agree_with_dpll_on_batched_workload():
    cnfs = batched_3sat(96)
    prog = PG.sat_batch(cnfs, width=10, plan_trials=1)
    got = OX.decide(prog, "GT0")
    ref = np.array([dpll.solve(c.num_vars, c.clauses) for c in cnfs], dtype=np.uint8)
    np.testing.assert_array_equal(got, ref)
