"""Host-side request and ensemble generators of the server and the ensemble bench (no GPU)."""
import math

import numpy as np

from mhdturbulence_amd.serve import ENSEMBLE_SIZES, ensemble_box, poisson_stream, request_params, splitmix64


def test_splitmix64_known_values():
    # reference outputs of the published SplitMix64 generator (Vigna) seeded with 0: the
    # first draw is splitmix64(0), the second splitmix64(0 + golden gamma), ...
    assert splitmix64(0) == 0xE220A8397B1DCDAF
    assert splitmix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4


def test_ensemble_boxes_deterministic_and_in_range():
    seen = set()
    for b in range(96):
        shape, ic = ensemble_box(b)
        assert ensemble_box(b) == (shape, ic)
        assert all(n in ENSEMBLE_SIZES for n in shape)
        assert 0.02 <= ic["a"] <= 0.1 and 0.0 <= ic["bx0"] <= 0.2
        assert 1e-3 <= ic["amp"] <= 1e-2 and 0.5 <= ic["p0"] <= 5.0
        seen.update(shape)
    assert seen == set(ENSEMBLE_SIZES)  # 288 draws cover every size
    assert ensemble_box(0, seed=1) != ensemble_box(0)


def test_request_params_ranges():
    for rid in range(200):
        p = request_params(rid)
        assert 0.02 <= p["a"] <= 0.1 and 0.0 <= p["bx0"] <= 0.2
        assert 1e-3 <= p["amp"] <= 1e-2 and 0.5 <= p["p0"] <= 5.0


def test_poisson_stream_shape():
    reqs = poisson_stream(rate=50.0, count=400, seed=3, n_range=(32, 256), t_range=(1.0, 6.0))
    arr = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(arr) >= 0)
    gaps = np.diff(np.concatenate([[0.0], arr]))
    assert abs(gaps.mean() - 1 / 50.0) < 0.25 / 50.0  # exponential gaps of mean 1/rate
    ns = np.array([r.n for r in reqs])
    assert np.all(ns % 8 == 0) and ns.min() >= 32 and ns.max() <= 256
    # log-uniform: the median sits near the geometric mean of the range
    assert abs(math.log(np.median(ns)) - 0.5 * (math.log(32) + math.log(256))) < 0.3
    assert all(1.0 <= r.t_end <= 6.0 for r in reqs)
    assert all(r.arrival == 0.0 for r in poisson_stream(rate=0.0, count=5))
