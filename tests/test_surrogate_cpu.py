"""Host logic of the surrogate's error-aware gate (no GPU)."""
import numpy as np

from stonefish_amd import surrogate as S


def test_fit_gate_walks_up_from_the_shallowest_bin():
    rng = np.random.default_rng(0)
    d = rng.uniform(-1.0, 1.0, 20000)
    # large errors below -0.4 m, small above: the band ends at the first bin
    # whose median error is at most tau, within one bin of -0.4
    e = np.where(d < -0.4, 0.3, 0.02) * rng.uniform(0.8, 1.2, d.size)
    band = S.fit_gate(d, e, in_lo9=-1.0, tau=0.1, bins=20)
    assert abs((-1.0 + band) - (-0.4)) <= 0.11
    # nothing above tau: off; everything above tau: the whole range
    assert S.fit_gate(d, np.full(d.size, 0.01), -1.0, 0.1) == 0.0
    assert abs(S.fit_gate(d, np.full(d.size, 0.5), -1.0, 0.1) - (d.max() + 1.0)) < 1e-9
    # a bin below tau ends the band even if deeper bins are above it again
    e2 = np.where(d < -0.6, 0.3, np.where(d < -0.2, 0.01, 0.3))
    assert -1.0 + S.fit_gate(d, e2, -1.0, 0.1) < -0.45
    # too few samples: off
    assert S.fit_gate(d[:50], e[:50], -1.0, 0.1) == 0.0


def test_delivered_summary_counts_flagged_bodies_as_exact():
    e = np.array([[0.5, 0.5], [0.001, 0.002], [0.2, 0.3], [0.0, 0.0]], np.float32)
    flagged = np.array([True, False, True, False])
    pos = np.array([2, 1, 2, 0])
    net = S.error_summary(e, flagged, pos)
    dl = S.delivered_summary(e, flagged, pos)
    assert net["n"] == 1 and dl["n"] == 3
    assert dl["F_median"] == 0.0 and dl["F_within_1pct"] == 1.0 and dl["T_median"] == 0.0
    assert abs(dl["T_p99"] - float(np.quantile(np.float32([0.0, 0.002, 0.0]), 0.99))) < 1e-9
    assert S.Norm(*(np.zeros(13),) * 4, np.zeros(6), np.ones(6)).c().gate_band == 0.0
