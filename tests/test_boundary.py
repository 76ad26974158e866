"""The C ABI boundary on the CPU: the library loads and exports every symbol that
include/zxtalori.h declares, including the whole table the reference's Python adapter
resolves at load time (python/zxtalori/api.py:29-53).  No compute calls (no GPU here)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from tests.conftest import GROUPS, ROOT

HEADER = ROOT / "include" / "zxtalori.h"
# The names the reference's api.py binds (python/zxtalori/api.py:31-50).
REFERENCE_TABLE = [
    "zo_create", "zo_create_error", "zo_destroy", "zo_generate", "zo_random", "zo_set", "zo_count", "zo_order",
    "zo_quaternions", "zo_operations", "zo_measure", "zo_spacing", "zo_relax", "zo_relax_neighbourhood",
    "zo_point_values", "zo_identity", "zo_cover", "zo_cover_optimize", "zo_cover_kicks", "zo_mesh",
    "zo_mesh_vertices", "zo_mesh_tets", "zo_mesh_vertex_count", "zo_mesh_tet_count", "zo_generate_exact",
    "zo_stats", "zo_error", "zo_cancel",
]


def declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(zo_[a-z_0-9]+)\s*\(", text)))


def library():
    from zxtalori_amd.api import library_path
    path = library_path()
    if not path.is_file():
        pytest.skip("libzxtalori.so not built (python -m zxtalori_amd.build)")
    return C.CDLL(str(path))


def test_header_declares_reference_table():
    names = declared()
    for name in REFERENCE_TABLE:
        assert name in names, name


def test_library_exports_every_declared_symbol():
    lib = library()
    for name in declared():
        assert hasattr(lib, name), name


def test_python_surface_matches_reference():
    import zxtalori_amd as zo
    assert zo.GROUPS == GROUPS
    for name in ("Grid", "generate", "from_quaternions", "load", "haar"):
        assert callable(getattr(zo, name))
    for method in ("relax", "measure", "spacing", "point_values", "set", "identity", "save", "close",
                   "covering", "triangulation", "optimize_covering"):
        assert callable(getattr(zo.Grid, method)), method
    for prop in ("quaternions", "operations", "stats"):
        assert isinstance(getattr(zo.Grid, prop), property)


def test_operator_table_matches_fixture(operator_fixture):
    """groups_table.h holds the fixture's operators rounded to f32, in the same order."""
    text = (ROOT / "zxtalori_amd" / "csrc" / "groups_table.h").read_text()
    rows = re.findall(r"\{(0x[^}]*)\}", text)
    values = np.array([[float.fromhex(v.strip().rstrip("f")) for v in r.split(",")] for r in rows], np.float64)
    offsets = [int(x) for x in re.search(r"kGroupOffset\[kGroupCount\] = \{([^}]*)\}", text).group(1).split(",")]
    for gid, group in enumerate(GROUPS):
        ops = operator_fixture[group]
        got = values[offsets[gid]:offsets[gid] + len(ops)]
        np.testing.assert_array_equal(got, ops.astype(np.float32).astype(np.float64))


def test_operator_tables_are_groups(operator_fixture):
    """Each table is closed under products up to sign (a group of unit quaternions)."""
    from oracle.dense_f64 import multiply
    for group in GROUPS:
        ops = operator_fixture[group]
        prod = multiply(ops[:, None, :], ops[None, :, :]).reshape(-1, 4)
        d = np.minimum(np.abs(prod[:, None, :] - ops[None]).max(-1), np.abs(prod[:, None, :] + ops[None]).max(-1))
        assert d.min(1).max() < 1e-12, group
