"""
GPU parity of the int8 MFMA route for dense Boolean steps (ts_contract2 with the summed
wires innermost in both operands; 6 or more summed bits on the pipelined kernel, 0..5 on
the one-stage kernel): bit-exact against the oracle in both epilogues, over
interleaved row / column / batch bits, partial tiles, sparse and all-ones operands (the
count 256 and 512 cases where MOD256 differs from GT0), and against the VALU route on
larger shapes.
"""

import numpy as np
import pytest

from oracle import finrel as F

pytestmark = pytest.mark.gpu


def _case(rng, kbits, mbits, nbits, bbits, density):
    K = list(range(kbits))
    M = [100 + i for i in range(mbits)]
    Nw = [200 + i for i in range(nbits)]
    B = [300 + i for i in range(bbits)]
    tail = K[::-1]  # summed wires innermost, same order in both operands
    lw = list(rng.permutation(M + B)) + tail
    rw = list(rng.permutation(Nw + B)) + tail
    ow = list(rng.permutation(M + Nw + B))
    lw, rw, ow = [int(x) for x in lw], [int(x) for x in rw], [int(x) for x in ow]
    if density >= 1.0:
        lhs = np.ones((2,) * len(lw), np.uint8)
        rhs = np.ones((2,) * len(rw), np.uint8)
    else:
        lhs = (rng.random((2,) * len(lw)) < density).astype(np.uint8)
        rhs = (rng.random((2,) * len(rw)) < density).astype(np.uint8)
    return lhs, lw, rhs, rw, ow


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


CASES = [
    # kbits, mbits, nbits, bbits, density
    (6, 4, 4, 0, 0.5),
    (8, 6, 6, 2, 0.3),
    (9, 7, 5, 1, 1.0),   # every count is 512: MOD256 gives 0, GT0 gives 1
    (8, 5, 6, 1, 1.0),   # every count is 256
    (6, 5, 9, 3, 0.5),   # partial row tiles
    (10, 8, 8, 0, 0.02),  # sparse: many zero counts
    (7, 4, 6, 4, 0.6),   # batch bits interleaved with row/column bits
    (9, 6, 7, 2, 0.97),  # counts near 2^9
    (6, 7, 8, 1, 0.5),   # 128 x 128 workgroup tiles
    (7, 8, 7, 0, 0.3),
    (8, 7, 7, 0, 1.0),   # 128 x 128 tiles, every count 256
    (12, 7, 8, 0, 0.05),  # 64 K stages: several unrolled loop trips, then the tail
    # one-stage kernel: K of 1..32 bytes zero-padded to 64
    (0, 4, 4, 0, 0.5),   # outer product
    (1, 5, 6, 2, 0.5),
    (2, 7, 7, 1, 0.5),   # 128 x 128 tiles
    (3, 6, 5, 3, 0.4),   # partial tiles, interleaved batch bits
    (4, 10, 11, 0, 0.3),  # 128 x 256 tiles
    (5, 8, 9, 1, 1.0),   # every count 32
    (5, 4, 4, 4, 0.05),  # sparse: many zero counts
]


@pytest.mark.parametrize("mode", ["GT0", "MOD256"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_dense_route_matches_oracle(case, mode):
    rng = np.random.default_rng(100 + case)
    kb, mb, nb, bb, dens = CASES[case]
    lhs, lw, rhs, rw, ow = _case(rng, kb, mb, nb, bb, dens)
    print(f"case {CASES[case]} {mode}", flush=True)
    from tensorsat_amd import _native as N

    if kb < 6:  # the one-stage kernel is opt-in: route the step there for the test
        N.check(N.lib().ts_dense_config(1, 0), "ts_dense_config")
    try:
        got = _run(lhs, lw, rhs, rw, ow, mode)
    finally:
        N.check(N.lib().ts_dense_config(1, 6), "ts_dense_config")
    ref = F.contract2(lhs, lw, rhs, rw, ow, mode)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("shape", [(12, 8, 8, 0), (10, 7, 7, 3), (6, 10, 10, 0), (4, 10, 10, 2), (2, 11, 9, 1)])
def test_dense_route_matches_valu_route(shape):
    from tensorsat_amd import _native as N

    rng = np.random.default_rng(7)
    lhs, lw, rhs, rw, ow = _case(rng, *shape, 2.0 ** (-shape[0] / 2))
    print(f"shape {shape}", flush=True)
    for mode in ("GT0", "MOD256"):
        N.check(N.lib().ts_dense_config(1, 0), "ts_dense_config")  # every eligible step on MFMA
        try:
            mfma = _run(lhs, lw, rhs, rw, ow, mode)
        finally:
            N.check(N.lib().ts_dense_config(1, 6), "ts_dense_config")
        N.check(N.lib().ts_dense_config(0, 6), "ts_dense_config")
        try:
            valu = _run(lhs, lw, rhs, rw, ow, mode)
        finally:
            N.check(N.lib().ts_dense_config(1, 6), "ts_dense_config")
        np.testing.assert_array_equal(mfma, valu)
        assert 0 < int(mfma.sum()) < mfma.size


@pytest.mark.parametrize("tile", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("case", [1, 4, 6, 8, 11])
def test_every_tile_matches_oracle(case, tile):
    """Each workgroup tile of the MFMA kernel (64x64, 128x64, 64x128, 128x128, 256x128,
    128x256), forced for the case's shape, gives the oracle's result (tiles larger than a
    side included)."""
    from tensorsat_amd import dense_tune as DT

    rng = np.random.default_rng(100 + case)
    kb, mb, nb, bb, dens = CASES[case]
    lhs, lw, rhs, rw, ow = _case(rng, kb, mb, nb, bb, dens)
    DT.set_tile(mb, nb, bb, kb, tile)
    try:
        assert DT.tile_of(mb, nb, bb, kb) == tile
        print(f"case {CASES[case]} tile {tile}", flush=True)
        for mode in ("GT0", "MOD256"):
            np.testing.assert_array_equal(_run(lhs, lw, rhs, rw, ow, mode), F.contract2(lhs, lw, rhs, rw, ow, mode))
    finally:
        DT.set_tile(mb, nb, bb, kb, 0)


def test_tile_autotuner_registers_the_fastest():
    from tensorsat_amd import dense_tune as DT

    shape = (8, 8, 1, 10)
    try:
        table = DT.autotune_tiles([shape, (3, 8, 0, 8)], reps=3)
        assert list(table) == [shape]  # the second shape is not on the MFMA route
        row = table[shape]
        assert set(row["ms"]) == set(DT.TILES) and row["tile"] == min(row["ms"], key=row["ms"].get)
        assert DT.tile_of(*shape) == row["tile"]
        print(f"tile table {row}", flush=True)
    finally:
        DT.set_tile(*shape, 0)
