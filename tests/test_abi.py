"""The C-ABI library loads on a machine without a GPU and exports every
entry point that include/vfsgpu.h declares (no compute calls here)."""
import ctypes as C
import os
import re

import vfsgpu as V

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "vfsgpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(vfs_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = V.lib()
    names = declared_symbols()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_workload_rng_is_deterministic_and_uniform():
    a = [V.uniform(1001, q) for q in range(20000)]
    assert a == [V.uniform(1001, q) for q in range(20000)]
    assert -1.0 <= min(a) and max(a) < 1.0
    m = sum(a) / len(a)
    assert abs(m) < 0.02


def test_config_layout_matches_header():
    # vfs_config: 3 ints + 3 ints + 6 ints + 8 ints, then 4 doubles
    assert C.sizeof(V.Config) == 20 * 4 + 4 * 8
