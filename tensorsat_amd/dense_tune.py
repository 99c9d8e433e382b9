"""
Tile autotuner of the dense int8 MFMA route (``csrc/dense.hip``).

A dense step is a batched product of 0/1 int8 matrices: 2^m rows held by lhs only, 2^n
columns held by rhs only, 2^b shared batch entries and 2^k contracted entries (K innermost
in both operands).  For each shape the tuner times the six workgroup tiles of the MFMA
kernel (64x64, 128x64, 64x128, 128x128, 256x128, 128x256 rows x columns) on synthetic operands of that
shape, through the ordinary ``ts_contract2`` entry point, and registers the fastest with
``ts_dense_tiles``; every later step of that shape (any operands, any caller) uses it.
Tiles change the schedule only: counts are exact int32 sums either way, so results are
identical for every tile.
"""

from __future__ import annotations

import ctypes

from . import _native as N

TILES = {1: (64, 64), 2: (128, 64), 3: (64, 128), 4: (128, 128), 5: (256, 128), 6: (128, 256)}


def tile_of(m: int, n: int, b: int, k: int) -> int:
    """The tile (1..6) steps of this shape use now."""
    t = N.i32(0)
    N.check(N.lib().ts_dense_tile_of(m, n, b, k, ctypes.byref(t)), "ts_dense_tile_of")
    return int(t.value)


def set_tile(m: int, n: int, b: int, k: int, tile: int) -> None:
    """Register a tile (1..6) for the shape, or 0 for the built-in rule."""
    N.check(N.lib().ts_dense_tiles(m, n, b, k, int(tile)), "ts_dense_tiles")


def _step(m: int, n: int, b: int, k: int, seed: int = 0):
    """A K-innermost dense step of the shape on synthetic 0/1 operands (density 2^(-k/2));
    returns a callable launching it on the current stream, and its buffers."""
    import torch

    from . import _native as N_

    L = N_.lib()
    K = list(range(k))[::-1]
    B = [300 + i for i in range(b)]
    lw = B + [100 + i for i in range(m)] + K
    rw = B + [200 + i for i in range(n)] + K
    ow = B + [100 + i for i in range(m)] + [200 + i for i in range(n)]
    g = torch.Generator(device="cuda").manual_seed(seed)
    dens = 2.0 ** (-k / 2)
    lhs = (torch.rand(1 << (b + m + k), device="cuda", generator=g) < dens).to(torch.uint8)
    rhs = (torch.rand(1 << (b + n + k), device="cuda", generator=g) < dens).to(torch.uint8)
    out = torch.empty(1 << (b + m + n), dtype=torch.uint8, device="cuda")
    la, ra, oa = N_.i32arr(lw), N_.i32arr(rw), N_.i32arr(ow)

    def run(mode: int = 0):
        N_.check(L.ts_contract2(N_.dptr(lhs), len(lw), la, N_.dptr(rhs), len(rw), ra, N_.dptr(out), len(ow), oa, mode,
                                N_.stream_ptr()), "ts_contract2")

    return run, (lhs, rhs, out)


def autotune_tiles(shapes, *, reps: int = 5, tiles=tuple(TILES)) -> dict:
    """
    Time every tile on each (m, n, b, k) shape (median of ``reps`` launches after one warm
    launch, HIP events) and register the fastest.  Returns {shape: {"tile": t,
    "ms": {tile: ms}}}.  Shapes the route does not take (k < 6, m or n < 4) are skipped.
    """
    import torch

    table = {}
    for shape in shapes:
        m, n, b, k = (int(x) for x in shape)
        if k < 6 or m < 4 or n < 4:
            continue
        run, _ = _step(m, n, b, k)
        times = {}
        for t in tiles:
            set_tile(m, n, b, k, t)
            run()
            torch.cuda.synchronize()
            ev = []
            for _ in range(reps):
                a, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record()
                run()
                e.record()
                ev.append((a, e))
            torch.cuda.synchronize()
            ms = sorted(a.elapsed_time(e) for a, e in ev)
            times[t] = ms[len(ms) // 2]
        best = min(times, key=times.get)
        set_tile(m, n, b, k, best)
        table[(m, n, b, k)] = {"tile": best, "ms": times}
    return table


def step_shapes(prog) -> list[tuple[int, int, int, int]]:
    """Distinct (m, n, b, k) shapes of the dense-route steps of a compiled program
    (K-innermost, >= 6 summed bits, >= 4 lhs-only and rhs-only out bits)."""
    from . import program as PG

    shapes = set()
    for st in prog.steps:
        fl = int(st["flags"])
        if fl & PG.STEP_NORHS or not fl & PG.STEP_KINNER or int(st["sum_rank"]) < 6:
            continue
        m = n = b = 0
        for j in range(int(st["out_rank"])):
            lp, rp = int(st["o_lpos"][j]) != PG.NOPOS, int(st["o_rpos"][j]) != PG.NOPOS
            b += lp and rp
            m += lp and not rp
            n += rp and not lp
        if m >= 4 and n >= 4:
            shapes.add((m, n, b, int(st["sum_rank"])))
    return sorted(shapes)
