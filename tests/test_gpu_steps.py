"""
GPU parity of the step kernels outside the MFMA route, through ts_contract2, in both
epilogues, bit-exact against the oracle: steps with few outputs and many summed terms
(one workgroup per output), with the summed wires innermost in both operands or
scattered -- all-ones operands make every count a multiple of 256, where MOD256 gives 0
and GT0 gives 1 -- and wide steps (summed wires innermost, at least 2^12 outputs) over
out orders that keep the operands' rows in runs (the vector path) or scatter them.
"""

import numpy as np
import pytest

from oracle import finrel as F

pytestmark = pytest.mark.gpu


def _run(lhs, lw, rhs, rw, ow, mode):
    import torch

    from tensorsat_amd import _native as N

    a = torch.from_numpy(lhs.reshape(-1).copy()).cuda()
    b = torch.from_numpy(rhs.reshape(-1).copy()).cuda()
    out = torch.full((1 << len(ow),), 7, dtype=torch.uint8, device="cuda")
    N.check(N.lib().ts_contract2(N.dptr(a), len(lw), N.i32arr(lw), N.dptr(b), len(rw), N.i32arr(rw), N.dptr(out),
                                 len(ow), N.i32arr(ow), N.MODES[mode], N.stream_ptr()), "ts_contract2")
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape((2,) * len(ow))


DOT_CASES = [
    # summed bits, lhs-only out bits, rhs-only out bits, shared out bits, density, summed innermost
    (17, 0, 0, 0, 0.5, True),   # a scalar closing a tree: one workgroup, 2^17 terms
    (12, 2, 1, 1, 0.3, True),
    (10, 3, 3, 2, 1.0, True),   # every count 1024: MOD256 gives 0
    (8, 4, 4, 4, 1.0, True),    # 2^12 outputs of count 256
    (9, 2, 2, 0, 0.01, True),   # sparse: many zero counts
    (11, 2, 2, 1, 0.5, False),  # summed bits scattered through both operands
    (8, 3, 2, 2, 1.0, False),
    (14, 0, 0, 0, 0.5, False),
]


@pytest.mark.parametrize("mode", ["GT0", "MOD256"])
@pytest.mark.parametrize("case", DOT_CASES, ids=[str(c) for c in DOT_CASES])
def test_few_outputs_many_terms(case, mode):
    kbits, mbits, nbits, bbits, density, inner = case
    rng = np.random.default_rng(kbits * 131 + mbits * 17 + nbits * 7 + bbits)
    K = list(range(kbits))
    M = [100 + i for i in range(mbits)]
    Nw = [200 + i for i in range(nbits)]
    B = [300 + i for i in range(bbits)]
    if inner:
        lw = [int(x) for x in rng.permutation(M + B)] + K[::-1]
        rw = [int(x) for x in rng.permutation(Nw + B)] + K[::-1]
    else:
        lw = [int(x) for x in rng.permutation(M + B + K)]
        rw = [int(x) for x in rng.permutation(Nw + B + K)]
    ow = [int(x) for x in rng.permutation(M + Nw + B)]
    if density >= 1.0:
        lhs = np.ones((2,) * len(lw), np.uint8)
        rhs = np.ones((2,) * len(rw), np.uint8)
    else:
        lhs = (rng.random((2,) * len(lw)) < density).astype(np.uint8)
        rhs = (rng.random((2,) * len(rw)) < density).astype(np.uint8)
    got = _run(lhs, lw, rhs, rw, ow, mode)
    want = F.contract2(lhs, lw, rhs, rw, ow, mode)
    np.testing.assert_array_equal(got, want)
    if density >= 1.0:
        assert want.all() == (mode == "GT0") and (want.any() == (mode == "GT0"))


WIDE_CASES = [
    # summed bits, lhs-only, rhs-only, shared out bits, density, out order
    (1, 4, 4, 6, 0.5, "lhs"),     # the n=120 cubes' common shape: one summed bit
    (1, 4, 4, 6, 0.5, "random"),
    (1, 4, 4, 6, 0.5, "shared"),  # low out bits: rows in runs in both operands (vector path)
    (1, 4, 5, 6, 0.5, "rhs-only"),  # low out bits: a run in rhs, absent from lhs
    (0, 6, 6, 2, 0.5, "lhs"),     # outer product with shared bits
    (0, 5, 6, 4, 0.5, "shared"),
    (2, 4, 4, 5, 0.5, "shared"),
    (3, 5, 3, 6, 0.5, "rhs"),     # 8-byte rows: eight outputs' loads per trip
    (3, 3, 3, 6, 0.5, "shared"),
    (5, 2, 7, 5, 0.3, "random"),
    (6, 3, 3, 8, 0.2, "lhs"),     # 64 terms, three lhs-only bits (not MFMA-shaped)
    (2, 0, 8, 6, 0.5, "rhs"),     # lhs holds only shared bits
    (4, 7, 0, 7, 0.5, "lhs"),     # rhs holds only shared bits
    (1, 1, 1, 12, 0.5, "random"),  # batch-heavy
    (2, 6, 6, 4, 1.0, "lhs"),     # all ones: every output 1 in both epilogues
    (0, 9, 9, 0, 0.05, "random"),  # sparse outer product, 2^18 outputs
]


@pytest.mark.parametrize("mode", ["GT0", "MOD256"])
@pytest.mark.parametrize("case", WIDE_CASES, ids=[str(c) for c in WIDE_CASES])
def test_wide_steps(case, mode):
    """Summed wires innermost in both operands, at least 2^12 outputs, out orders that keep
    the operands' rows in runs (one load per operand for 2..16 consecutive outputs) or
    scatter them."""
    kbits, mbits, nbits, bbits, density, order = case
    rng = np.random.default_rng(kbits * 1009 + mbits * 101 + nbits * 11 + bbits + len(order))
    K = list(range(kbits))
    M = [100 + i for i in range(mbits)]
    Nw = [200 + i for i in range(nbits)]
    B = [300 + i for i in range(bbits)]
    lhead = [int(x) for x in rng.permutation(M + B)]
    rhead = [int(x) for x in rng.permutation(Nw + B)]
    if order == "shared":  # shared bits lowest in both operands, same order, lowest in out
        lhead, rhead = M + B, Nw + B
        ow = M + Nw + B
    elif order == "rhs-only":  # rhs-only bits lowest in rhs and in out
        rhead = B + Nw
        ow = M + B + Nw
    elif order == "lhs":  # low out bits are the lhs bits right above the summed ones, in order
        ow = [w for w in rhead if w not in lhead] + lhead
    elif order == "rhs":
        ow = [w for w in lhead if w not in rhead] + rhead
    else:
        ow = [int(x) for x in rng.permutation(M + Nw + B)]
    lw = lhead + K[::-1]
    rw = rhead + K[::-1]
    if density >= 1.0:
        lhs = np.ones((2,) * len(lw), np.uint8)
        rhs = np.ones((2,) * len(rw), np.uint8)
    else:
        lhs = (rng.random((2,) * len(lw)) < density).astype(np.uint8)
        rhs = (rng.random((2,) * len(rw)) < density).astype(np.uint8)
    got = _run(lhs, lw, rhs, rw, ow, mode)
    np.testing.assert_array_equal(got, F.contract2(lhs, lw, rhs, rw, ow, mode))
