"""bench.py's contract on the CPU: one GPU only, and the figures its roofline fields are
computed from (the per-pair work it prices, the committed counter summaries it reads)."""
import json
import subprocess
import sys

import pytest

import bench
from conftest import ROOT


def test_more_than_one_gpu_is_refused():
    out = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 2
    line = json.loads(out.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


@pytest.mark.parametrize("terms,valu", [(2, 95.375), (4, 97.375)])
def test_priced_pair_work(terms, valu):
    """The reference's per-pair near-field work on the counting type: the window runs for
    7/8 of the pairs, and one arctangent term is one fused multiply-add."""
    got_valu, got_trans = bench.pair_op_counts(terms)
    assert got_valu == pytest.approx(valu)
    assert got_trans == pytest.approx(7.75)


def test_counter_summaries_parse():
    """The roofline's traffic and VALU-busy fields come from the committed summaries: a
    malformed edit of either would silently drop them from the bench line."""
    assert bench.traffic_from_profiles() == 465813088
    counts = bench.counters_from_profile("near_force_i1e6")
    assert counts == {"valu_per_launch": 9506401256.0, "trans_per_launch": 253788840.0,
                      "salu_per_launch": 3233402398.0}
    # 64 % of the published issue peak over the kernel's 12.99 ms (2 cycles per VALU, 8 per
    # transcendental, 1024 SIMDs at 2.4 GHz).
    busy = bench.valu_busy(counts, 12.99e-3)
    expected = ((9506401256 - 253788840) * 2 + 253788840 * 8) / (1024 * 2.4e9 * 12.99e-3)
    assert busy == pytest.approx(expected)
    assert busy == pytest.approx(0.643, abs=0.001)


def test_spacing_counter_summary_parses():
    counts = bench.counters_from_profile("spacing_i1e7")
    assert counts == {"valu_per_launch": 778408599.0, "trans_per_launch": 3769623.0, "salu_per_launch": 125594259.0}
    assert bench.valu_busy(counts, 1.26e-3) == pytest.approx(0.51, abs=0.005)
