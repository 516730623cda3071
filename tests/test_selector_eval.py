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
    S = np.full(40, 0.2, dtype=np.float32)   # 200 ms of search per pipeline
    X = np.zeros((40, selector.N_FEATURES), dtype=np.float32)
    params = [{"theta": 1.0 if i % 2 else 0.0} for i in range(40)]
    rep = explore.evaluate(_Net(9), np.zeros(selector.N_FEATURES), np.ones(selector.N_FEATURES), X, T, S, params,
                           fixed=11, runs=10, baseline_net=_Net(11))
    assert rep["fallback_fraction"] == 0.0
    assert abs(rep["selector_run_over_best"] - 1.0) < 1e-6
    fb = rep["fixed_blend"]
    assert fb["variant"] == 11
    assert abs(fb["run_over_selector"] - 2.0) < 1e-5 and abs(fb["run_over_selector_skewed"] - 2.0) < 1e-5
    assert abs(rep["run_ms"]["default"] - float(T[:, selector.N_VARIANTS].sum())) < 1e-3
    # policies over 10 runs: the exhaustive search pays 200 ms per pipeline,
    # the confident selector none
    assert abs(rep["policy_total_ms"]["selector"] - 40 * 10 * 0.5) < 1e-3
    assert abs(rep["policy_total_ms"]["exhaustive"] - (40 * 10 * 0.5 + 40 * 200)) < 1e-2
    assert abs(rep["baseline_net"]["run_ms_no_fallback"] - 40.0) < 1e-4


class _Unsure(_Net):
    def predict(self, x, mask):
        p = super().predict(x, mask)
        p[:, self.pick] = 0.3          # below the confidence threshold
        return p


def test_fallback_is_charged_its_search():
    rng = np.random.default_rng(3)
    T = _timings(10, rng)
    T[:, 4] = 0.25
    S = np.linspace(0.1, 1.0, 10).astype(np.float32)
    X = np.zeros((10, selector.N_FEATURES), dtype=np.float32)
    params = [{"theta": 0.0}] * 10
    rep = explore.evaluate(_Unsure(7), np.zeros(6), np.ones(6), X, T, S, params, runs=3)
    assert rep["fallback_fraction"] == 1.0
    want = float((3 * 0.25 + S * 1e3).sum())
    assert abs(rep["policy_total_ms"]["selector"] - want) < 1e-2


def test_features_are_measured_from_the_data():
    """Selectivity from the predicate on a sample, the group domain from the
    metadata, key skew from the sample: uniform keys give ~0, Zipf 1.5 keys
    a clearly positive estimate."""
    from vldb_voila_amd import datagen, workloads
    for theta in (0.0, 1.5):
        prm = {"probe_rows": 200_000, "build_rows": 50_000, "selectivity": 0.3, "groups": 1000, "theta": theta}
        b, p = datagen.gen_batch_pipeline(prm, seed=1, index=0)
        d = workloads.batch_pipeline(prm, b, p).desc()
        f = selector.features(d, p, "fk")
        assert abs(f[0] - np.log10(200_000)) < 1e-5 and abs(f[1] - np.log10(50_000)) < 1e-5
        assert abs(f[2] - 0.3) < 0.03
        assert abs(f[3] - 3.0) < 0.01
        assert f[5] == 4
        if theta == 0.0:
            assert f[4] < 0.3
        else:
            assert f[4] > 0.8
