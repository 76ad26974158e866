"""On-disk samples: save()/load()/from_quaternions() (python/zxtalori/api.py:336-362).

JSON "zxtalori-1" (shared with the browser app) and NPZ with a JSON `metadata` string.
Loading imports the rows through zo_set, which folds them as the reference does, so a
loaded sample equals the oracle's fold of the saved rows bit for bit (and the saved rows
to the last place of a renormalisation).  The browser app's own export
("zxtalori-orientations", web/app/app.js:162) loads through the same path.
"""
import json

import numpy as np
import pytest

import oracle
import zxtalori_amd as zo
from tests.conftest import GROUPS, ops32

pytestmark = pytest.mark.gpu
GROUP_ID = {g: i for i, g in enumerate(GROUPS)}


@pytest.mark.parametrize("suffix", [".json", ".npz"])
@pytest.mark.parametrize("group", ["O", "D6", "I"])
def test_round_trip(tmp_path, operator_fixture, suffix, group):
    path = tmp_path / f"sample{suffix}"
    with zo.haar(group, 700, 21) as grid:
        grid.relax(3)
        saved = grid.quaternions
        stats = grid.stats
        assert grid.save(path) == path
    if suffix == ".json":
        doc = json.loads(path.read_text(encoding="utf-8"))
        assert doc["format"] == "zxtalori-1" and doc["group"] == group
        assert doc["metadata"]["iteration"] == 3 and doc["metadata"]["count"] == 700
        np.testing.assert_array_equal(np.asarray(doc["quaternions"], np.float32), saved)
    else:
        with np.load(path, allow_pickle=False) as data:
            assert json.loads(str(data["metadata"]))["group"] == group
            np.testing.assert_array_equal(data["quaternions"], saved)
    with zo.load(path) as again:
        q = again.quaternions
        assert again.stats["group"] == group and len(again) == 700
        assert again.stats["generator"] == "Imported quaternions"
    np.testing.assert_array_equal(q, oracle.fold(saved, ops32(operator_fixture, group), GROUP_ID[group]))
    assert np.abs(q - saved).max() <= 2.4e-7
    assert stats["group"] == group


def test_generated_grid_round_trip(tmp_path):
    with zo.generate("T", h=4) as grid:
        grid.save(tmp_path / "grid.json")
        q = grid.quaternions
    with zo.load(tmp_path / "grid.json") as again:
        assert np.abs(again.quaternions - q).max() <= 2.4e-7


def test_browser_export_loads(tmp_path, operator_fixture):
    rows = oracle.fold(oracle.haar(64, 3), ops32(operator_fixture, "O"), GROUP_ID["O"])
    doc = {"format": "zxtalori-orientations", "version": 1, "convention": "scalar-first, right group action",
           "group": "O", "quaternions": rows.tolist(), "provenance": {}, "measures": {}, "history": []}
    path = tmp_path / "zxtalori-O-64.json"
    path.write_text(json.dumps(doc, indent=2), encoding="utf-8")
    with zo.load(path) as grid:
        assert grid.stats["group"] == "O" and len(grid) == 64
        np.testing.assert_array_equal(grid.quaternions, oracle.fold(rows, ops32(operator_fixture, "O"), GROUP_ID["O"]))


def test_from_quaternions_rejects_bad_rows():
    with pytest.raises(RuntimeError, match="NonFiniteQuaternion"):
        zo.from_quaternions(np.array([[np.inf, 0, 0, 0]], np.float32), "O")
    with pytest.raises(RuntimeError, match="ZeroQuaternion"):
        zo.from_quaternions(np.zeros((2, 4), np.float32), "O")
