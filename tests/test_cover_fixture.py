"""The covering oracle (full-orbit Qhull) against the reference's covering-reference.json.

The covering radius depends only on the orbits, so the unfolded host grids (inverse Hopf
f = 2, zo_hopf_points) and the reference's own h = 2 transports (reference-small.npz)
must reproduce the fixture's radii; this pins the Qhull oracle the GPU tests use and the
counts the GPU tests expect.
"""
import json

import numpy as np
import pytest

import zxtalori_amd as zo
from tests.conftest import GOLDEN
from tests.test_gpu_cover import full_orbit_covering

REFERENCE = json.loads((GOLDEN / "covering-reference.json").read_text())


@pytest.fixture(scope="module")
def operators():
    with np.load(GOLDEN / "pair-operators.npz", allow_pickle=False) as d:
        return {k: d[k].copy() for k in d.files}


@pytest.mark.parametrize("group", list(REFERENCE))
def test_fixture_radius(group, operators):
    expected = REFERENCE[group]
    if group in ("T", "O", "I"):
        with np.load(GOLDEN / "reference-small.npz", allow_pickle=False) as d:
            q = d[group].copy()
    else:
        q = zo.hopf_points(group, 2).astype(np.float64)
    assert len(q) == expected["count"]
    radius = full_orbit_covering(q, operators[group])
    # The fixture's transports were measured on the reference's FP32 device grids, which
    # lie within 3e-6 of these f64 rows (tests/test_grid_shaders.py:39-64): about 5e-6 deg
    # of radius.  The Hopf grids are the same f64 rows on both sides.
    tolerance = 2e-5 if group in ("T", "O", "I") else 2e-6
    assert abs(radius - expected["radius_deg"]) < tolerance, (group, radius, expected["radius_deg"])
