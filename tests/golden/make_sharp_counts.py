"""
Writes ``sharp_n100_counts.json``: exact model counts of the #SAT workload's instances
(random 3-SAT, n=100, m=300, seeds 0..9) computed by the oracle's independent model
counter (``oracle/sharp.c``: DPLL with connected components and a component cache),
which shares no code with the product's cube front-end or counting kernels.

Run from the repository root: ``python tests/golden/make_sharp_counts.py``.
"""

import json
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import sharp  # noqa: E402
from tensorsat_amd.workloads import count_3sat  # noqa: E402


def main() -> None:
    out = {"_what": "exact #SAT of CNFInstance.random(3, 100, 300, rng=seed) by oracle/sharp.c", "counts": {}}
    for seed in range(10):
        cnf = count_3sat(seed)
        c, nodes = sharp.count(cnf.num_vars, cnf.clauses, with_nodes=True)
        out["counts"][str(seed)] = str(c)
        print(seed, c, nodes, flush=True)
    (Path(__file__).parent / "sharp_n100_counts.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
