"""Diagnostics on an MI355X: the measured packed-int16 VALU peak is in the
range the CU count and reported clock allow (the roofline of the batched
kernel is quoted against both)."""
import pytest

import eschalign_amd as ea

pytestmark = pytest.mark.gpu


def test_valu_peak_probe_is_within_the_spec_bound():
    p = ea.valu_peak16(0)
    spec = p["n_cu"] * p["clock_khz"] * 1e3 * 128
    assert p["n_cu"] > 0 and p["clock_khz"] > 0
    # a VALU microbenchmark cannot beat one packed op per SIMD per 4 clocks
    # at the top clock, and a healthy device reaches most of it
    assert 0.5 * spec < p["lane_ops_per_s"] <= 1.02 * spec, (p, spec)
