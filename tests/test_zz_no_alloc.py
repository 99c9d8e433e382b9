"""
The allocation-free search entry point (``ts_search_decide_noalloc``, continuation lists
in the library's static device array) against the allocating path (``ts_search_run``
rounds with caller buffers, ``SearchSolver``): the same verdicts and round structure on
two calls in a row, the second on a smaller batch, with no device memory taken.
"""

import ctypes

import numpy as np
import pytest

from tensorsat_amd.workloads import batched_3sat

pytestmark = pytest.mark.gpu


def _noalloc(cnfs, width, split_target, budget, spill_capacity=0):
    import torch

    from tensorsat_amd import _native as N
    from tensorsat_amd import search as S

    words, inst = S.pack_batch(cnfs)
    jobs = S.plan_jobs(inst, split_target=split_target)
    S.validate(words, inst, jobs, width)
    d_words = torch.from_numpy(words.view(np.int32)).cuda()
    d_inst = torch.from_numpy(inst.view(np.int32).reshape(-1)).cuda()
    d_jobs = torch.from_numpy(jobs.view(np.int32).reshape(-1)).cuda()
    d_ver = torch.zeros(len(cnfs), dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    args = S._Search(N.dptr(d_words), N.dptr(d_inst), len(inst), N.dptr(d_jobs), len(jobs), N.dptr(d_ver), None,
                     width, 0, 1, spill_capacity, budget, None, None)
    rounds = N.i32(0)
    free0 = torch.cuda.mem_get_info()[0]
    N.check(N.lib().ts_search_decide_noalloc(ctypes.byref(args), ctypes.byref(rounds), N.stream_ptr()),
            "ts_search_decide_noalloc")
    free1 = torch.cuda.mem_get_info()[0]
    v = d_ver.cpu().numpy().astype(np.int64)
    return np.where(v & 1, 1, np.where(v & 2, -1, 0)), int(rounds.value), free0 - free1


def _reference(cnfs, width, split_target, budget):
    from tensorsat_amd.search import SearchSolver

    s = SearchSolver(width=width, split_target=split_target, budget=budget, queue=False)
    v = s.run(cnfs)
    return v, s.rounds


def test_noalloc_matches_allocating_search_twice():
    big = batched_3sat(4096)
    small = big[:300]
    for cnfs in (big, small):
        want, want_rounds = _reference(cnfs, 6, 8.0, 8)
        got, rounds, taken = _noalloc(cnfs, 6, 8.0, 8)
        print(f"{len(cnfs)} instances: rounds {rounds} (allocating path {want_rounds}), device bytes taken {taken}",
              flush=True)
        np.testing.assert_array_equal(got, want)
        assert (got >= 0).all()
        assert rounds > 1
        assert taken == 0


def test_noalloc_small_list_and_bad_arguments():
    from tensorsat_amd import _native as N
    from tensorsat_amd import search as S

    cnfs = batched_3sat(512)
    want, _ = _reference(cnfs, 6, None, 4)
    got, rounds, _ = _noalloc(cnfs, 6, None, 4, spill_capacity=32)  # hand-overs overflow: searched in place
    np.testing.assert_array_equal(got, want)
    too_long = S._Search(None, None, 0, None, 0, None, None, 6, 0, 1, 70000, 8, None, None)  # > static list
    assert N.lib().ts_search_decide_noalloc(ctypes.byref(too_long), None, N.stream_ptr()) == N.TS_EINVAL
    zero_budget = S._Search(None, None, 0, None, 0, None, None, 6, 0, 1, 0, 0, None, None)
    assert N.lib().ts_search_decide_noalloc(ctypes.byref(zero_budget), None, N.stream_ptr()) == N.TS_EINVAL
