"""
Writes ``config2_w26_digests.json``: for the single-instance decision workload (random
3-SAT n=120, m=510, seed 0) sliced into cubes of width 26 by the product's planner, the
xxh64 digest of every intermediate tensor of every cube as the oracle's step loop
(``oracle.executor.run_job``) computes it, folded per cube into one chain digest
(xxh64 over the concatenated 8-byte step digests, then the result's digest), in GT0
for every cube and in MOD256 for the first cubes.  A digest of the compiled program
(jobs, steps, leaves) is stored beside them, so a test can tell a changed plan from a
wrong kernel; the cube search runs on a fixed number of host threads (its parallel
split seeds each subtree's planner, so the program depends on it).

Run from the repository root: ``python tests/golden/make_config2_digests.py`` (a few
minutes: ~3.5 s of numpy per cube).
"""

import json
import sys
from pathlib import Path

import xxhash

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import executor as OX  # noqa: E402
from tensorsat_amd import program as PG  # noqa: E402
from tensorsat_amd.workloads import single_3sat  # noqa: E402

SEED, WIDTH, TRIALS, THREADS, MOD256_CUBES = 0, 26, 2, 8, 6


def program_digest(prog) -> str:
    h = xxhash.xxh64()
    for a in (prog.jobs, prog.steps, prog.leaves):
        h.update(a.tobytes())
    return h.hexdigest()


def chain_digest(step_outputs, result) -> str:
    h = xxhash.xxh64()
    for t in step_outputs:
        h.update(xxhash.xxh64(t.tobytes()).digest())
    h.update(xxhash.xxh64(result.tobytes()).digest())
    return h.hexdigest()


def main() -> None:
    prog = PG.sat_batch([single_3sat(SEED)], width=WIDTH, plan_trials=TRIALS, threads=THREADS)
    out = {"seed": SEED, "width": WIDTH, "plan_trials": TRIALS, "threads": THREADS, "program": program_digest(prog),
           "jobs": len(prog.jobs), "GT0": [], "MOD256": []}
    for j in range(len(prog.jobs)):
        for mode in ("GT0", "MOD256"):
            if mode == "MOD256" and j >= MOD256_CUBES:
                continue
            res, inter = OX.run_job(prog, j, mode, keep_intermediates=True)
            out[mode].append(chain_digest(inter, res))
        print(j, out["GT0"][-1], flush=True)
    (Path(__file__).parent / "config2_w26_digests.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
