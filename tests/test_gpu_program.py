"""GPU parity: compiled programs on the HIP executors against the oracle's step loop."""

import numpy as np
import pytest

from tensorsat_amd.network import bintree_network, clause_network
from tensorsat_amd.sat import CNFInstance
from tensorsat_amd import program as PG

pytestmark = pytest.mark.gpu


def _instances():
    out = []
    for seed in range(12):
        n = 8 + seed % 9
        out.append(CNFInstance.random(3, n, int(round(4.26 * n)), rng=seed))
    return out


@pytest.mark.parametrize("mode", ["GT0", "MOD256"])
@pytest.mark.parametrize("enc", [bintree_network, clause_network])
def test_single_job_intermediates_bit_exact(mode, enc):
    from oracle import executor as OX
    from tensorsat_amd.executor import DeviceProgram, run_job

    for cnf in _instances():
        d = enc(cnf)
        prog = PG.compile_network(d.wiring, d.blocks, PG.elim(d.wiring, trials=4))
        print(f"case enc={enc.__name__} mode={mode} n={cnf.num_vars} steps={len(prog.steps)}", flush=True)
        ref, ref_inter = OX.run_job(prog, 0, mode, keep_intermediates=True)
        dp = DeviceProgram(prog)
        res, inter = run_job(dp, 0, mode, keep_intermediates=True)
        assert len(inter) == len(ref_inter)
        for s, (a, b) in enumerate(zip(inter, ref_inter)):
            np.testing.assert_array_equal(a.cpu().numpy(), b, err_msg=f"step {s}")
        np.testing.assert_array_equal(res.cpu().numpy(), ref)


def test_open_output_relation():
    from oracle import executor as OX
    from tensorsat_amd.executor import DeviceProgram, run_job

    cnf = CNFInstance.random(4, 10, 95, rng=0)
    d = bintree_network(cnf, "_" * 10)
    prog = PG.compile_network(d.wiring, d.blocks, PG.elim(d.wiring, trials=4))
    res, _ = run_job(DeviceProgram(prog), 0, "GT0")
    ref, _ = OX.run_job(prog, 0, "GT0")
    np.testing.assert_array_equal(res.cpu().numpy(), ref)
    sols = {tuple(int(b) for b in np.binary_repr(i, 10)) for i in np.flatnonzero(res.cpu().numpy())}
    # notebooks/02-random-ksat.ipynb cell 25
    assert sols == {(1, 0, 0, 1, 0, 0, 0, 0, 0, 0), (1, 0, 0, 1, 1, 0, 0, 0, 0, 0),
                    (1, 0, 0, 1, 1, 1, 0, 1, 0, 0), (1, 0, 0, 1, 1, 1, 1, 1, 0, 0)}


@pytest.mark.parametrize("mode", ["GT0", "MOD256"])
def test_batch_verdicts(mode):
    from oracle import brute, executor as OX
    from tensorsat_amd.executor import BatchExecutor

    rng = np.random.default_rng(5)
    cnfs = []
    for i in range(48):
        n = int(rng.integers(14, 23))
        ratio = 4.26 if i % 2 == 0 else 3.5
        cnfs.append(CNFInstance.random(3, n, int(round(ratio * n)), rng=1000 + i))
    prog = PG.sat_batch(cnfs, width=8)
    print(f"batch jobs={len(prog.jobs)} steps={len(prog.steps)}", flush=True)
    ex = BatchExecutor(prog, track_jobs=True)
    got = ex.run(mode, early_exit=False)
    ref = OX.decide(prog, mode)
    np.testing.assert_array_equal(got, ref)
    truth = np.array([brute.is_sat(c.num_vars, c.clauses) for c in cnfs], dtype=np.uint8)
    np.testing.assert_array_equal(got, truth)
    got2 = ex.run(mode, early_exit=True)
    np.testing.assert_array_equal(got2, ref)


def test_counts_exact():
    from oracle import brute
    from tensorsat_amd.executor import DeviceProgram, count_program

    cnfs = [CNFInstance.random(3, n, int(3.0 * n), rng=200 + n) for n in range(12, 22)]
    prog = PG.sat_batch(cnfs, width=8)
    got = count_program(DeviceProgram(prog))
    assert got == [brute.count(c.num_vars, c.clauses) for c in cnfs]
