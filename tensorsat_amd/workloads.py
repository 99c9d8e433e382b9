"""
Synthetic workloads with fixed seeds (the measurement configurations).

* :func:`batched_3sat` — 4096 random 3-SAT instances, instance ``i`` drawn from
  ``default_rng(10000 + i)``: ``n ~ U{20..80}``, clause ratio 4.26 for even ``i`` and
  3.5 for odd ``i``, ``m = round(ratio * n)``, clauses by ``CNFInstance.random`` with the
  same generator.
* :func:`single_3sat` — one random 3-SAT instance, n=120, m=510, seeds 0..19.
* :func:`count_3sat` — #SAT instances, n=100, m=300, seeds 0..9.
* :func:`plain_cpu_3sat` — n=40, m=170, seeds 0..99.
* :func:`learned_ksat` — k ~ U{3,4,5}, n ~ U[150, 300], ratio 0.9 x threshold.
"""

from __future__ import annotations

import numpy as np

from .sat import CNFInstance

THRESHOLD = {3: 4.267, 4: 9.931, 5: 21.117}


def batched_3sat(num: int = 4096, start: int = 0) -> list[CNFInstance]:
    out = []
    for i in range(start, start + num):
        rng = np.random.default_rng(10_000 + i)
        n = int(rng.integers(20, 81))
        ratio = 4.26 if i % 2 == 0 else 3.5
        out.append(CNFInstance.random(3, n, int(round(ratio * n)), rng=rng))
    return out


def single_3sat(seed: int = 0) -> CNFInstance:
    return CNFInstance.random(3, 120, 510, rng=seed)


def count_3sat(seed: int = 0) -> CNFInstance:
    return CNFInstance.random(3, 100, 300, rng=seed)


def plain_cpu_3sat(seed: int = 0) -> CNFInstance:
    return CNFInstance.random(3, 40, 170, rng=seed)


def learned_ksat(i: int, base: int = 1_000_000) -> CNFInstance:
    rng = np.random.default_rng(base + i)
    k = int(rng.integers(3, 6))
    n = int(rng.integers(150, 301))
    m = int(round(0.9 * THRESHOLD[k] * n))
    return CNFInstance.random(k, n, m, rng=rng)
