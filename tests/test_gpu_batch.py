"""Batched persistent launch: many uneven select -> join -> group-by
pipelines in one kernel, each checked bit-exact against the oracle, also
when replayed from a captured hipGraph."""
import numpy as np
import pytest

import oracle
from vldb_voila_amd import abi, datagen, runtime, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    print("gpu batch: init device 0", flush=True)
    runtime.init(0)


def small_batch(count, seed=7):
    params = datagen.batch_params(count, seed=seed, probe_range=(1e3, 2e5), build_range=(1e2, 5e4),
                                  group_range=(1, 2e4))
    data, wants = [], []
    for i, prm in enumerate(params):
        b, p = datagen.gen_batch_pipeline(prm, seed=seed, index=i)
        data.append((prm, b, p))
        wants.append(oracle.sorted_rows(oracle.bruteforce(workloads.batch_pipeline(prm, b, p).desc())))
    return data, wants


def test_batch_matches_oracle_and_sequential():
    print("batch of 24 pipelines", flush=True)
    data, wants = small_batch(24)
    devs, descs, cfgs = [], [], []
    modes = [abi.VH_GROUP_AUTO, abi.VH_GROUP_GLOBAL, abi.VH_GROUP_HASH, abi.VH_GROUP_LDSHASH]
    for i, (prm, b, p) in enumerate(data):
        db, dp = runtime.upload(b), runtime.upload(p)
        devs.append((db, dp))
        descs.append(workloads.batch_pipeline(prm, b, p, runtime.addrs(db), runtime.addrs(dp)).desc())
        cfgs.append(runtime.make_config(group_mode=modes[i % len(modes)], tile_shift=-(i % 3)))
    batch = runtime.Batch(descs, cfgs)
    s = runtime.Stream()
    batch.run(s)
    for i in range(len(data)):
        assert oracle.sorted_rows(batch.result(i, s)) == wants[i], f"pipeline {i}"
    # graph replays, with eager work in between, every replay checked
    for rep in range(3):
        batch.run_graph(s)
        s.synchronize()
        for i in range(0, len(data), 5):
            assert oracle.sorted_rows(batch.result(i, s)) == wants[i], f"replay {rep} pipeline {i}"
        q = runtime.Query(descs[1])
        q.run(s)
        assert oracle.sorted_rows(q.result(s)) == wants[1]
        q.free()
    batch.free()


def test_batch_with_empty_pipelines():
    print("batch with empty inputs", flush=True)
    data, wants = small_batch(4, seed=3)
    prm, b, p = data[0]
    b0 = {k: v[:0] for k, v in b.items()}
    p0 = {k: v[:0] for k, v in p.items()}
    cases = [(prm, b0, p), (prm, b, p0), data[1], data[2]]
    wants2 = [oracle.sorted_rows(oracle.bruteforce(workloads.batch_pipeline(x, y, z).desc())) for x, y, z in cases]
    devs, descs = [], []
    for prm_, bb, pp in cases:
        db, dp = runtime.upload(bb), runtime.upload(pp)
        devs.append((db, dp))
        descs.append(workloads.batch_pipeline(prm_, bb, pp, runtime.addrs(db), runtime.addrs(dp)).desc())
    batch = runtime.Batch(descs)
    batch.run()
    for i in range(4):
        assert oracle.sorted_rows(batch.result(i)) == wants2[i]


def test_batch_mixed_kernels_modes_and_repeated_keys():
    """Generated and generic probe launches in one batch, every dense
    group-by strategy, and builds with repeated keys (the partitioned builds'
    overflow slots), over two runs and a graph replay."""
    print("batch: mixed kernels, modes, repeated build keys", flush=True)
    data, _ = small_batch(12, seed=21)
    rng = np.random.default_rng(21)
    cases, wants = [], []
    for i, (prm, b, p) in enumerate(data):
        if i % 3 == 0 and len(b["key"]) > 10:
            b = dict(b)
            k = b["key"].copy()
            k[rng.integers(0, len(k), max(1, len(k) // 20))] = k[0]     # ~5% of rows repeat one key
            b["key"] = k
        cases.append((prm, b, p))
        wants.append(oracle.sorted_rows(oracle.bruteforce(workloads.batch_pipeline(prm, b, p).desc())))
    modes = [abi.VH_GROUP_LDS, abi.VH_GROUP_PART, abi.VH_GROUP_AUTO, abi.VH_GROUP_HASH]
    devs, descs, cfgs = [], [], []
    for i, (prm, b, p) in enumerate(cases):
        db, dp = runtime.upload(b), runtime.upload(p)
        devs.append((db, dp))
        descs.append(workloads.batch_pipeline(prm, b, p, runtime.addrs(db), runtime.addrs(dp)).desc())
        mode = modes[i % len(modes)]
        if mode == abi.VH_GROUP_LDS and prm["groups"] > 1000:
            mode = abi.VH_GROUP_AUTO       # the dense LDS table holds ~1,100 groups of these sums
        kernel = abi.VH_KERNEL_GENERIC if i % 2 else abi.VH_KERNEL_AUTO
        cfgs.append(runtime.make_config(group_mode=mode, kernel=kernel))
    batch = runtime.Batch(descs, cfgs)
    s = runtime.Stream()
    for run in range(2):
        batch.run(s)
        for i in range(len(cases)):
            assert oracle.sorted_rows(batch.result(i, s)) == wants[i], f"run {run} pipeline {i}"
    batch.run_graph(s)
    s.synchronize()
    for i in range(len(cases)):
        assert oracle.sorted_rows(batch.result(i, s)) == wants[i], f"graph pipeline {i}"
    batch.free()
