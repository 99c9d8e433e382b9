"""
ORACLE — TEST INFRASTRUCTURE ONLY.

CPU restatement of the reference's Boolean-relation arithmetic and contraction loop,
used as the checker for the HIP executor.  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import this package, and only to
check or time; the product (``tensorsat_amd``) never imports it.

Modules:

* :mod:`oracle.finrel` — ``FinRel`` arithmetic (``lang/fin_rel.py``, ``_utils/misc.py``)
  in two epilogue modes: ``"MOD256"`` (what the reference's numpy code computes) and
  ``"GT0"`` (the mathematically intended ``count > 0``), plus exact counting.
* :mod:`oracle.executor` — the reference-equivalent step loop over a contraction
  plan (``contractions/simple.py:251-292`` / cotengra's pairwise loop), timed as the
  CPU baseline.
* :mod:`oracle.brute` — brute-force SAT / #SAT enumeration for small ``n``.
* ``oracle/dpll.c`` — an independent DPLL decision procedure (kissat is absent),
  built by ``oracle/Makefile`` into ``oracle/_build``.

Pinning: every function is checked in ``tests/test_oracle.py`` against the golden
vectors held by the reference (notebook outputs, DIMACS parser expectations, verdict
tests), closed forms and brute force; see the header of each module.
"""
