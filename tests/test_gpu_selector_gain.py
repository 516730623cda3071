"""What the learned selector gains, measured on the GPU: the explorer times
all 48 variants (and the built-in default) of 300 training and 60 held-out
pipelines drawn like the batched workload (smaller sizes so the exploration
takes about two minutes), with the features measured from each pipeline's
data; the network is trained on those timings with the HIP kernels.  On the
held-out pipelines:

  * its choices run faster in total than the best fixed blend (the one
    variant with the lowest training total, used for every pipeline);
  * its picks run faster in total than the untrained network's;
  * run 100 times per pipeline, the selector policy (its choice, or the
    exhaustive search with that search's measured wall time charged when
    its confidence is below 0.5) costs less in total than searching every
    pipeline.

Only the kernel choice changes; every variant's rows are the same."""
import json

import numpy as np
import pytest

from vldb_voila_amd import datagen, explore, runtime, selector

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def explored(monkeypatch_module=None):
    print("selector gain: explore 300 + 60 pipelines", flush=True)
    runtime.init(0)
    saved = (explore.PROBE_RANGE, explore.BUILD_RANGE, explore.GROUP_RANGE)
    explore.PROBE_RANGE, explore.BUILD_RANGE, explore.GROUP_RANGE = (1e4, 2e5), (1e3, 5e4), (1, 1e5)
    try:
        tr = explore.explore(300, seed=7, reps=1, log_every=30)
        ho = explore.explore(60, seed=77, reps=1, log_every=30)
    finally:
        explore.PROBE_RANGE, explore.BUILD_RANGE, explore.GROUP_RANGE = saved
    return tr, ho


def test_selector_beats_fixed_blend_untrained_network_and_exhaustive_search(explored):
    (Xtr, Ttr, Str, ptr), (Xho, Tho, Sho, pho) = explored
    print("selector gain: train", flush=True)
    net, mu, sd, losses = selector.train(Xtr, Ttr[:, :selector.N_VARIANTS], epochs=300)
    assert losses[-1] < losses[0]
    untrained = selector.MLP(selector._he_init(np.random.default_rng(0)))
    rep = explore.evaluate(net, mu, sd, Xho, Tho, Sho, pho, fixed=explore.fixed_blend(Ttr, Tho), runs=100,
                           baseline_net=untrained)
    print(json.dumps(rep), flush=True)
    assert rep["run_ms"]["selector"] < rep["fixed_blend"]["run_ms"]
    assert rep["run_ms"]["selector_no_fallback"] < rep["baseline_net"]["run_ms_no_fallback"]
    assert rep["policy_total_ms"]["selector"] < rep["policy_total_ms"]["exhaustive"]
