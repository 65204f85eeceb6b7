"""Host-side parts of the surrogate warm start (no GPU): the input scaling,
the parameter sampler's ranges, and the orbits-to-steady-torque measure."""
import math

import numpy as np

from dustyfargoadsg_amd import surrogate


def test_features_span_the_ranges():
    ps = surrogate.sample_params(500, 3)
    f = surrogate.features(ps)
    assert f.shape == (500, 4) and f.min() >= -1.0 and f.max() <= 1.0
    assert f.min(axis=0).max() < -0.9 and f.max(axis=0).min() > 0.9
    lo = surrogate.features([dict(q=1e-5, alpha=1e-4, h=0.03, flaring=0.0)])
    hi = surrogate.features([dict(q=3e-3, alpha=1e-2, h=0.08, flaring=0.25)])
    assert np.allclose(lo, -1.0) and np.allclose(hi, 1.0)


def test_orbits_to_steady():
    # a torque relaxing as exp(-t / 5 orbits) towards 1: the one-orbit mean
    # stays within 5 % from about t = 5 ln(1/0.05) orbits on
    t = np.arange(800) / 20.0
    tq = 1.0 + np.exp(-t / 5.0)
    o = surrogate.orbits_to_steady(tq, 1.0)
    assert abs(o - (5.0 * math.log(1 / 0.05) + 0.5)) < 0.6
    assert surrogate.orbits_to_steady(np.ones(100), 1.0) == 1.0
    assert math.isnan(surrogate.orbits_to_steady(np.full(100, 2.0), 1.0))
