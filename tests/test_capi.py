"""The C-ABI libraries load on the CPU and export every symbol their headers declare."""

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def _declared(header: str) -> set[str]:
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_\s\*]*?\b(ts[a-z]*_[a-z0-9_]+)\s*\(", text, flags=re.M))
    return names


def test_headers_declare_expected_entry_points():
    d = _declared("tensorsat_amd.h")
    assert {"ts_contract2", "ts_rewire", "ts_batch_run", "ts_program_count"} <= d
    assert {"tsp_greedy", "tsp_hyper", "tsp_tree_cost", "tsp_slice"} <= _declared("tensorsat_plan.h")
    assert {"tsc_compile", "tsc_sat_batch", "tsp_elim"} <= _declared("tensorsat_prog.h")


def test_planner_library_exports_everything():
    lib = ctypes.CDLL(str(ROOT / "tensorsat_amd" / "libtsplan.so"))
    for name in _declared("tensorsat_plan.h") | _declared("tensorsat_prog.h"):
        assert hasattr(lib, name), name


def test_hip_library_exports_everything():
    path = ROOT / "tensorsat_amd" / "libtensorsat_amd.so"
    if not path.exists():
        pytest.fail("libtensorsat_amd.so is not built (python -m tensorsat_amd.build)")
    from tensorsat_amd import _native as N

    lib = N.load()  # no GPU call
    declared = _declared("tensorsat_amd.h")
    for name in declared:
        assert hasattr(lib, name), name
    assert set(N.EXPORTS) == declared
    assert lib.ts_version().startswith(b"tensorsat_amd")


def test_dense_tile_table_host_side():
    """The dense tile table is host state (no GPU needed): set, read back, reset to the
    built-in rule, bad arguments."""
    import ctypes

    from tensorsat_amd import _native as N

    L = N.load()

    def tile_of(*shape):
        t = ctypes.c_int32(0)
        assert L.ts_dense_tile_of(*shape, ctypes.byref(t)) == 0
        return t.value

    assert tile_of(8, 8, 0, 8) == 4 and tile_of(5, 9, 0, 8) == 1 and tile_of(10, 12, 0, 8) == 6  # built-in rule
    assert L.ts_dense_tiles(8, 8, 0, 8, 2) == 0
    assert tile_of(8, 8, 0, 8) == 2
    assert L.ts_dense_tiles(8, 8, 0, 8, 0) == 0
    assert tile_of(8, 8, 0, 8) == 4
    assert L.ts_dense_tiles(8, 8, 0, 8, 7) == N.TS_EINVAL
    assert L.ts_dense_tiles(-1, 8, 0, 8, 1) == N.TS_EINVAL
    assert L.ts_dense_tile_of(8, 8, 0, 28, None) == N.TS_EINVAL
