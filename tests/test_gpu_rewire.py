"""
GPU parity of the rewire kernels: the LDS-staged bit permutation (kernel (a)) over random
permutations of ranks 0..20, and the wave-level reduction over 6..12 discarded axes
(kernel (b)) in both epilogues, against the oracle's exact rewire (and the reference's
literal rewire where the kept ports are increasing, where the two agree).
"""

import numpy as np
import pytest

from oracle import finrel as F

pytestmark = pytest.mark.gpu


def _rewire(x, out_ports, mode):
    import torch

    from tensorsat_amd import _native as N

    r = x.ndim
    a = torch.from_numpy(x.reshape(-1).copy()).cuda()
    out = torch.full((1 << len(out_ports),), 7, dtype=torch.uint8, device="cuda")
    N.check(N.lib().ts_rewire(N.dptr(a), r, N.dptr(out), len(out_ports), N.i32arr(out_ports), N.MODES[mode],
                              N.stream_ptr()), "ts_rewire")
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape((2,) * len(out_ports))


@pytest.mark.parametrize("rank", [0, 1, 3, 6, 7, 11, 14, 20])
def test_permutation_matches_transpose(rank):
    rng = np.random.default_rng(rank)
    for trial in range(3):
        perm = [int(p) for p in rng.permutation(rank)]
        x = rng.integers(0, 2, size=(2,) * rank).astype(np.uint8)
        print(f"rank {rank} perm {perm}", flush=True)
        got = _rewire(x, perm, "GT0")
        np.testing.assert_array_equal(got, np.transpose(x, perm) if rank else x)


@pytest.mark.parametrize("mode", ["GT0", "MOD256"])
@pytest.mark.parametrize("rank,keep", [(8, 2), (12, 4), (14, 8), (16, 4), (18, 10)])
def test_reduction_matches_oracle(rank, keep, mode):
    rng = np.random.default_rng(rank * 31 + keep)
    for trial, dens in enumerate((0.5, 0.01, 1.0)):
        kept = sorted(int(p) for p in rng.choice(rank, size=keep, replace=False))
        order = [int(p) for p in rng.permutation(kept)] if trial == 1 else kept
        x = (rng.random((2,) * rank) < dens).astype(np.uint8)
        print(f"rank {rank} out_ports {order} density {dens} {mode}", flush=True)
        got = _rewire(x, order, mode)
        cnt = F.rewire_exact(x, order)
        want = (cnt > 0) if mode == "GT0" else (cnt % 256 != 0)
        np.testing.assert_array_equal(got, want.astype(np.uint8))
        if order == kept:
            np.testing.assert_array_equal(got, F.rewire(x, order, mode))


@pytest.mark.parametrize("rank,kept", [
    (10, [4, 5, 6, 7]),                      # no in-register bit kept, all lane bits
    (12, [0, 1, 2, 3, 10, 11]),              # all four in-register bits kept
    (13, [1, 3, 6, 12]),                     # two in-register bits kept
    (20, [0, 2, 3, 5, 9, 15, 19]),           # three kept, discarded high bits
    (22, [1, 4, 8, 13, 17, 21]),             # 16 discarded: term chunks across waves
    (24, [2, 7, 11, 12, 16, 20, 23]),        # 17 discarded
])
def test_wide_reduction_matches_oracle(rank, kept):
    """The 16-byte-lane existential reduction (tensors of >= 10 bits, GT0): every split
    of the four in-register bits into kept and discarded, term chunks across waves, and
    bytes other than 0/1 (any nonzero byte is true)."""
    rng = np.random.default_rng(rank)
    for trial, dens in enumerate((0.3, 0.002, 0.0)):
        order = [int(p) for p in rng.permutation(kept)] if trial == 1 else list(kept)
        x = ((rng.random((2,) * rank) < dens) * rng.integers(1, 256, size=(2,) * rank)).astype(np.uint8)
        print(f"rank {rank} out_ports {order} density {dens}", flush=True)
        got = _rewire(x, order, "GT0")
        want = F.rewire_exact(x != 0, order) > 0
        np.testing.assert_array_equal(got, want.astype(np.uint8))
