"""Learned flavor selector and the exhaustive autotuner it falls back to.

The reference picks flavors by exploring them all (explorer.cpp:112-427)
and choosing the fastest offline (print_base_flavors.py:32-40).  Here the
autotuner times the 48 kernel variants of a pipeline on the GPU
(vh_tune_time), and a small MLP trained on those timings (vh_mlp_*: forward,
backward and Adam are HIP kernels) predicts the fastest variant from six
pipeline features.  Predictions below 0.5 confidence fall back to the
exhaustive search.  Only the choice of kernel changes: results stay
bit-exact.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

from . import abi
from .runtime import DeviceBuffer, check, init, lib

N_VARIANTS = 48
N_FEATURES = 6
N_PARAMS = 7728
CONFIDENCE = 0.5


def _bind():
    L = lib()
    L.vh_variant_config.argtypes = [C.c_int32, C.c_void_p]
    L.vh_tune_time.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p]
    L.vh_mlp_new.argtypes = [C.POINTER(C.c_void_p), C.c_int32, C.c_void_p]
    L.vh_mlp_free.argtypes = [C.c_void_p]
    L.vh_mlp_params.argtypes = [C.c_void_p, C.c_void_p]
    L.vh_mlp_predict.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]
    L.vh_mlp_train_step.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_float,
                                    C.c_void_p, C.POINTER(C.c_float), C.c_void_p]
    L.vh_tune_best.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
    return L


def variant_configs():
    L = _bind()
    out = []
    for v in range(N_VARIANTS):
        c = abi.vh_config()
        check(L.vh_variant_config(v, C.byref(c)))
        out.append(c)
    return out


def time_variants(desc, reps=3, stream=None, configs=None) -> np.ndarray:
    """Fastest run (ms) of each variant; -1 where the variant does not apply."""
    init()
    L = _bind()
    cfgs = configs or variant_configs()
    arr = (abi.vh_config * len(cfgs))(*cfgs)
    ms = np.zeros(len(cfgs), dtype=np.float32)
    check(L.vh_tune_time(C.byref(desc), C.addressof(arr), len(cfgs), reps, ms.ctypes.data,
                         stream.ptr if stream else None))
    return ms


SAMPLE = 8192
_TYPE_BYTES = {abi.VH_U8: 1, abi.VH_I8: 1, abi.VH_U16: 2, abi.VH_I16: 2, abi.VH_U32: 4, abi.VH_I32: 4,
               abi.VH_U64: 8, abi.VH_I64: 8}
_CMP = {abi.VH_LT: np.less, abi.VH_LE: np.less_equal, abi.VH_GT: np.greater, abi.VH_GE: np.greater_equal,
        abi.VH_EQ: np.equal, abi.VH_NE: np.not_equal}


def zipf_theta(keys: np.ndarray) -> float:
    """Skew estimate of a key sample: minus the slope of log(count) against
    log(rank) over its most frequent keys (those seen at least twice, at
    most 64 ranks); 0 when no key repeats."""
    _, counts = np.unique(keys, return_counts=True)
    c = np.sort(counts)[::-1]
    c = c[c >= 2][:64]
    if len(c) < 3:
        return 0.0
    x = np.log(np.arange(1, len(c) + 1, dtype=np.float64))
    y = np.log(c.astype(np.float64))
    slope = np.polyfit(x, y, 1)[0]
    return float(min(max(-slope, 0.0), 3.0))


def features(desc, probe: dict, key_col: str) -> np.ndarray:
    """The selector's six features, measured from the pipeline and its host
    columns: log10 probe rows, log10 build rows of the (last) join, the
    selectivity of the probe predicates on an evenly strided sample of
    SAMPLE rows, log10 of the group-key domain (the keys' metadata, as the
    lowering sizes the group-by), the skew of the probe join key on the
    same sample (zipf_theta), and the probe key's width in bytes."""
    n = int(desc.probe.rows)
    idx = np.linspace(0, max(n - 1, 0), num=min(SAMPLE, max(n, 1))).astype(np.int64) if n else np.zeros(0, np.int64)
    names = list(probe.keys())
    keep = np.ones(len(idx), dtype=bool)
    for k in range(desc.probe.npreds):
        pr = desc.probe.preds[k]
        keep &= _CMP[pr.cmp](probe[names[pr.col]][idx], pr.value)
    sel = float(keep.mean()) if len(idx) else 0.0
    build_rows = int(desc.joins[desc.njoins - 1].build.rows) if desc.njoins else 0
    domain = 1.0
    for k in range(desc.nkeys):
        v = desc.keys[k]
        c = desc.probe.cols[v] if v < desc.probe.ncols else None
        if c is None:   # a join payload: its build column's range
            base = desc.probe.ncols
            for j in range(desc.njoins):
                J = desc.joins[j]
                if v < base + J.npayload:
                    c = J.build.cols[J.payload[v - base]]
                    break
                base += J.npayload
        domain *= float(c.max) - float(c.min) + 1
    theta = zipf_theta(probe[key_col][idx]) if len(idx) else 0.0
    width = _TYPE_BYTES.get(desc.probe.cols[names.index(key_col)].type, 8)
    return np.array([math.log10(max(n, 1)), math.log10(max(build_rows, 1)), sel, math.log10(max(domain, 1)),
                     theta, width], dtype=np.float32)


class MLP:
    """vh_mlp: the selector network on the GPU."""

    def __init__(self, init_params: np.ndarray, batch_capacity=1024):
        init()
        self.L = _bind()
        p = C.c_void_p()
        ip = np.ascontiguousarray(init_params, dtype=np.float32)
        assert ip.size == N_PARAMS
        check(self.L.vh_mlp_new(C.byref(p), batch_capacity, ip.ctypes.data))
        self.ptr = p.value
        self.cap = batch_capacity

    def params(self) -> np.ndarray:
        out = np.empty(N_PARAMS, dtype=np.float32)
        check(self.L.vh_mlp_params(self.ptr, out.ctypes.data))
        return out

    def predict(self, x: np.ndarray, mask: np.ndarray | None = None) -> np.ndarray:
        """Probabilities (B, 48); rows go through in chunks of the network's
        batch capacity."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        B = x.shape[0]
        out = np.empty((B, N_VARIANTS), dtype=np.float32)
        for s in range(0, B, self.cap):
            xs = np.ascontiguousarray(x[s:s + self.cap])
            b = xs.shape[0]
            dx = DeviceBuffer.from_numpy(xs)
            dm = (DeviceBuffer.from_numpy(np.ascontiguousarray(mask[s:s + self.cap], dtype=np.uint8))
                  if mask is not None else None)
            dp = DeviceBuffer(b * N_VARIANTS * 4)
            check(self.L.vh_mlp_predict(self.ptr, dx.ptr, dm.ptr if dm else None, b, dp.ptr, None))
            out[s:s + b] = dp.to_numpy(np.float32, b * N_VARIANTS).reshape(b, N_VARIANTS)
        return out

    def train_step(self, x, y, mask=None, lr=0.0, want_grads=False):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.int32)
        B = x.shape[0]
        dx, dy = DeviceBuffer.from_numpy(x), DeviceBuffer.from_numpy(y)
        dm = DeviceBuffer.from_numpy(np.ascontiguousarray(mask, dtype=np.uint8)) if mask is not None else None
        grads = np.empty(N_PARAMS, dtype=np.float32) if want_grads else None
        loss = C.c_float()
        check(self.L.vh_mlp_train_step(self.ptr, dx.ptr, dm.ptr if dm else None, dy.ptr, B, lr,
                                       grads.ctypes.data if grads is not None else None, C.byref(loss), None))
        return loss.value, grads

    def free(self):
        if self.ptr:
            self.L.vh_mlp_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def normalise(X: np.ndarray):
    """Feature scaling (per-feature mean / std of the training set)."""
    mu = X.mean(0)
    sd = X.std(0)
    sd[sd == 0] = 1
    return mu.astype(np.float32), sd.astype(np.float32)


def train(X, times, epochs=200, lr=3e-3, batch=256, seed=0, init_params=None):
    """Train on the explorer's timing table: label = fastest applicable
    variant, mask = applicable variants.  Returns (MLP, mu, sd, losses)."""
    X = np.asarray(X, dtype=np.float32)
    times = np.asarray(times, dtype=np.float32)
    mask = (times > 0).astype(np.uint8)
    y = np.where(times > 0, times, np.inf).argmin(1).astype(np.int32)
    mu, sd = normalise(X)
    Xn = (X - mu) / sd
    if init_params is None:
        from_seed = np.random.default_rng(seed)
        init_params = _he_init(from_seed)
    net = MLP(init_params, batch_capacity=max(batch, 1))
    rng = np.random.default_rng(seed + 1)
    losses = []
    n = len(Xn)
    for ep in range(epochs):
        order = rng.permutation(n)
        tot = 0.0
        for s in range(0, n, batch):
            idx = order[s:s + batch]
            l, _ = net.train_step(Xn[idx], y[idx], mask[idx], lr=lr)
            tot += l * len(idx)
        losses.append(tot / n)
    return net, mu, sd, losses


def _he_init(rng):
    parts = []
    for fan_in, shape in ((6, (6, 64)), (None, (64,)), (64, (64, 64)), (None, (64,)), (64, (64, 48)),
                          (None, (48,))):
        if fan_in is None:
            parts.append(np.zeros(shape))
        else:
            lim = math.sqrt(6.0 / fan_in)
            parts.append(rng.uniform(-lim, lim, shape))
    return np.concatenate([p.reshape(-1) for p in parts]).astype(np.float32)


def choose(net, mu, sd, feats, mask, fallback):
    """Selector decision: predicted variant when its probability >= 0.5,
    else fallback() (the exhaustive autotuner).  Returns (variant, used_fallback)."""
    p = net.predict(((np.asarray(feats, dtype=np.float32) - mu) / sd)[None, :],
                    np.asarray(mask, dtype=np.uint8)[None, :])[0]
    v = int(p.argmax())
    if p[v] >= CONFIDENCE and mask[v]:
        return v, False
    return int(fallback()), True


MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "selector_model.npz")


def load_model(path=MODEL):
    """(network, mu, sd) of a trained selector saved by the explorer
    (model.npz: params, mu, sd)."""
    z = np.load(path)
    return MLP(z["params"]), z["mu"].astype(np.float32), z["sd"].astype(np.float32)


def tune_best(desc, reps=2, stream=None, configs=None):
    """The autotuner through its cache (vh_tune_best): (variant index, cached)."""
    init()
    L = _bind()
    cfgs = configs or variant_configs()
    arr = (abi.vh_config * len(cfgs))(*cfgs)
    best, cached = C.c_int32(), C.c_int32()
    check(L.vh_tune_best(C.byref(desc), C.addressof(arr), len(cfgs), reps, C.byref(best), C.byref(cached),
                         stream.ptr if stream else None))
    return best.value, bool(cached.value)


def applicable(desc, configs=None, batch=False) -> np.ndarray:
    """Which variants lower for the pipeline (host-side lowering only, as a
    single query or as a batch's pipeline): the selector's mask."""
    L = lib()
    fn = L.vh_generate_batch_source if batch else L.vh_generate_source
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]
    cfgs = configs or variant_configs()
    n = C.c_size_t()
    return np.array([fn(C.byref(desc), C.byref(c), None, 0, C.byref(n)) == 0 for c in cfgs], dtype=np.uint8)


def pick_config(net, mu, sd, desc, probe: dict, key_col: str, configs=None, reps=2, stream=None, batch=False):
    """The selector's variant for a pipeline among those that apply, or the
    autotuner's (exhaustive timing, cached per signature) when the network
    is below 0.5 confidence.  Returns (config, variant, used_fallback)."""
    cfgs = configs or variant_configs()
    x = features(desc, probe, key_col)
    mask = applicable(desc, cfgs, batch=batch)
    p = net.predict(((x - mu) / sd)[None, :], mask[None, :])[0]
    v = int(p.argmax())
    if p[v] >= CONFIDENCE and mask[v]:
        return cfgs[v], v, False
    v, _ = tune_best(desc, reps=reps, stream=stream, configs=cfgs)
    if batch and not mask[v]:
        v = int(np.argmax(np.where(mask > 0, p, -1.0)))   # the fastest alone does not apply in a batch
    return cfgs[v], v, True
