"""
Programs of one contraction step whose summed bits are not innermost: two leaves of
``n`` bits over the same wires, contracted to ``out_wires``.  The compiler lays out leaves
so that such a step is K-innermost; ``permute_rhs`` then permutes the bits of the rhs
leaf (its bytes and the step's rhs positions together), which keeps the step's value and
makes it a general step (KINNER cleared) with ``n - len(out_wires)`` summed bits.
"""

import numpy as np

from tensorsat_amd import program as PG
from tensorsat_amd.network import Wiring


def permute_rhs(prog: PG.Program, rng) -> np.ndarray:
    steps = prog.steps
    assert len(steps) == 1 and len(prog.jobs) == 1 and int(prog.jobs["leaf_begin"][0]) == 0
    rr, off = int(steps["rhs_rank"][0]), int(steps["rhs_off"][0])
    assert off + (1 << rr) <= int(prog.jobs["num_leaf_bytes"][0])
    pi = rng.permutation(rr)
    while (pi == np.arange(rr)).all():
        pi = rng.permutation(rr)
    idx = np.arange(1 << rr)
    new_idx = np.zeros_like(idx)
    for k in range(rr):
        new_idx |= ((idx >> k) & 1) << int(pi[k])
    old = prog.leaves[off : off + (1 << rr)].copy()
    prog.leaves[off + new_idx] = old
    for name, cnt in (("o_rpos", int(steps["out_rank"][0])), ("c_rpos", int(steps["sum_rank"][0]))):
        for i in range(cnt):
            p = int(steps[name][0, i])
            if p != PG.NOPOS:
                steps[name][0, i] = pi[p]
    steps["flags"][0] &= ~np.uint8(PG.STEP_KINNER)
    return pi


def wide_sum_program(seed: int, out_wires=(0, 1), n: int = 10, density: float | None = None,
                     permute: bool = True) -> tuple[PG.Program, list[np.ndarray]]:
    rng = np.random.default_rng(seed)
    w = Wiring([2] * n, [list(range(n)), [int(x) for x in rng.permutation(n)]], list(out_wires))
    p = rng.uniform(0.01, 0.06) if density is None else density
    leaves = [(rng.random(1 << n) < p).astype(np.uint8) for _ in range(2)]
    prog = PG.compile_network(w, leaves, PG.elim(w, trials=1))
    if permute:
        permute_rhs(prog, rng)
    return prog, leaves
