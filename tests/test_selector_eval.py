"""Selector evaluation bookkeeping (CPU): the fixed blend is the one variant
that applies everywhere with the lowest training total, and the report
prices every choice with the measured timings."""
import numpy as np

from vldb_voila_amd import explore, selector


class _Net:
    """Stands in for the trained network: always confident in `pick`."""
    def __init__(self, pick):
        self.pick = pick

    def predict(self, x, mask):
        p = np.zeros((len(x), selector.N_VARIANTS), dtype=np.float32)
        p[:, self.pick] = 1.0
        return p


def _timings(rows, rng):
    T = rng.uniform(1.0, 2.0, (rows, selector.N_VARIANTS + 1)).astype(np.float32)
    return T


def test_fixed_blend_skips_variants_that_do_not_always_apply():
    rng = np.random.default_rng(1)
    tr, ev = _timings(50, rng), _timings(20, rng)
    tr[:, 3] = 0.1                 # fastest in training ...
    ev[5, 3] = -1.0                # ... but does not apply to one evaluated pipeline
    tr[:, 7] = 0.2
    assert explore.fixed_blend(tr, ev) == 7


def test_report_prices_choices_with_measured_times():
    rng = np.random.default_rng(2)
    T = _timings(40, rng)
    T[:, 9] = 0.5                  # the selector's pick and the best variant
    T[:, 11] = 1.0                 # the fixed blend
    X = np.zeros((40, selector.N_FEATURES), dtype=np.float32)
    params = [{"theta": 1.0 if i % 2 else 0.0} for i in range(40)]
    rep = explore.evaluate(_Net(9), np.zeros(selector.N_FEATURES), np.ones(selector.N_FEATURES), X, T, params,
                           fixed=11)
    assert rep["fallback_fraction"] == 0.0
    assert abs(rep["selector_vs_best_total"] - 1.0) < 1e-6
    fb = rep["fixed_blend"]
    assert fb["variant"] == 11
    assert abs(fb["over_selector_total"] - 2.0) < 1e-5 and abs(fb["over_selector_skewed"] - 2.0) < 1e-5
    assert abs(rep["total_ms"]["default"] - float(T[:, selector.N_VARIANTS].sum())) < 1e-3
