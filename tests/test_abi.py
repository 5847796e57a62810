"""The C-ABI library loads and exports every entry point the headers declare;
argument validation answers with status codes before touching the GPU."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(lufqhe_\w+)\s*\(", text))
    return sorted(names)


def test_headers_declare_entry_points():
    names = declared_symbols()
    assert "lufqhe_det_batched" in names
    assert len(names) >= 4


def test_library_exports_declared_symbols(lc):
    lib = lc.load()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_status_strings(lc):
    lib = lc.load()
    assert lib.lufqhe_status_string(0) == b"success"
    assert lib.lufqhe_version() >= 100


@pytest.mark.parametrize("n,ld,stride,batch,policy,arith,want", [
    (0, 1, 1, 1, 0, 0, 1),       # n < 1
    (65, 65, 65 * 65, 1, 0, 0, 1),  # n > 64
    (8, 7, 64, 1, 0, 0, 1),      # ld < n
    (8, 8, -1, 1, 0, 0, 1),      # negative stride
    (8, 8, 64, -1, 0, 0, 1),     # negative batch
    (8, 8, 64, 1, 1, 1, 2),      # ABS pivots with reference arithmetic: unsupported
    (8, 8, 64, 1, 7, 0, 2),      # unknown policy
    (8, 8, 64, 0, 0, 0, 0),      # empty batch: no-op success
])
def test_argument_validation(lc, n, ld, stride, batch, policy, arith, want):
    lib = lc.load()
    st = lib.lufqhe_det_batched(None, batch, n, ld, stride, None, None, None, None, None, None,
                                policy, arith, None)
    assert st == want


def test_null_input_rejected(lc):
    lib = lc.load()
    assert lib.lufqhe_det_batched(None, 4, 8, 8, 64, None, None, None, None, None, None,
                                  0, 0, None) == 1


def test_epilogue_argument_validation(lc):
    lib = lc.load()
    assert lib.lufqhe_det_inplace(None, 0, 8, 8, 64, 1, 0, None, None) == 0   # empty batch
    assert lib.lufqhe_det_inplace(None, 4, 8, 8, 64, 1, 0, None, None) == 1   # null A
    assert lib.lufqhe_det_inplace(None, 4, 0, 8, 64, 1, 0, None, None) == 1   # bad n
    assert lib.lufqhe_det_inplace(None, 4, 8, 8, 64, 1, 9, None, None) == 2   # bad policy
    assert lib.lufqhe_global_reduction(None, 4, 64, None, None, None) == 1     # null output
    assert lib.lufqhe_global_reduction(None, 4, 64, None, 1, None) == 1        # null A, pairs > 0
    assert lib.lufqhe_pair_sum(None, None, 4, None, None, None, None) == 1
    assert lib.lufqhe_det_inplace_workspace_bytes(10) == 200
    assert lib.lufqhe_pair_sum_workspace_bytes() > 0


def test_shape_table_api(lc):
    lib = lc.load()
    for n in range(1, 65):
        c = lib.lufqhe_shape_count(n)
        assert 1 <= c <= 4
        for sh in range(c):
            lanes, cols, waves = lc.shape_info(n, sh)
            assert lanes in (4, 8, 16, 32, 64) and cols in (1, 2, 4) and lanes % cols == 0
            assert waves in (1, 2)
    assert lib.lufqhe_shape_count(0) == 0 and lib.lufqhe_shape_count(65) == 0
    assert lib.lufqhe_set_shape(0, 0, 32, 7) == 1          # shape out of range
    assert lib.lufqhe_set_shape(0, 1, 32, 1) == 4          # reference arithmetic: shape 0 only
    assert lib.lufqhe_set_shape(1, 1, 32, 0) == 2          # ABS + reference arithmetic
    assert lib.lufqhe_get_shape(0, 0, 0) == -1
    old = lc.get_shape("crout_ref", "fused", 32)
    assert lc.set_shape("crout_ref", "fused", 32, 2) and lc.get_shape("crout_ref", "fused", 32) == 2
    assert lc.set_shape("crout_ref", "fused", 32, old)


def test_fqhe_argument_validation(lc):
    lib = lc.load()
    args = (None, None, None, None, None, None, 0, 0, None)
    assert lib.lufqhe_fqhe_det(None, 0, 8, 8, None, *args) == 0     # empty batch
    assert lib.lufqhe_fqhe_det(None, 4, 8, 5, None, *args) == 1     # 0 < z_stride < n
    assert lib.lufqhe_fqhe_det(None, 4, 65, 65, None, *args) == 1   # n > 64
    assert lib.lufqhe_fqhe_det(None, 4, 8, 8, None, *args) == 1     # null positions
    assert lib.lufqhe_fqhe_det(None, 4, 8, 0, None, None, None, None, None, None, None,
                               1, 1, None) == 2                      # ABS + reference arithmetic
    assert lib.lufqhe_fqhe_build(None, 4, 8, 8, None, None, None) == 1
    assert lib.lufqhe_fqhe_build(None, 0, 8, 8, None, None, None) == 0


def test_shape_table_loads_from_c(lc, tmp_path):
    """The C loader applies the committed table exactly as the Python one."""
    import json
    lib = lc.load()
    from importlib import import_module
    core = import_module("lion-codes_amd.lucore")
    entries = json.loads(core.SHAPE_TABLE.read_text())["entries"]
    saved = {(e["policy"], e["n"]): lc.get_shape(e["policy"], "fused", e["n"]) for e in entries}
    try:
        for (pol, n) in saved:
            assert lc.set_shape(pol, "fused", n, 0)
        applied = lib.lufqhe_load_shape_table(str(core.SHAPE_TABLE).encode())
        assert applied == len(entries)
        for e in entries:
            assert lc.get_shape(e["policy"], "fused", e["n"]) == e["shape"]
        # an entry whose geometry no longer matches its index is skipped
        bad = dict(entries[0], lanes=entries[0]["lanes"] * 2 if entries[0]["lanes"] < 64 else 4)
        p = tmp_path / "t.json"
        p.write_text(json.dumps({"entries": [bad]}))
        assert lib.lufqhe_load_shape_table(str(p).encode()) == 0
        assert lib.lufqhe_load_shape_table(str(tmp_path / "missing.json").encode()) == -1
    finally:
        for (pol, n), sh in saved.items():
            lc.set_shape(pol, "fused", n, sh)
