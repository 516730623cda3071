"""Seeded synthetic inputs for the benchmark workloads.

PRNG (counter based, so every column is reproducible in any order and in C):

    mix64(z)      = splitmix64 finaliser:
                    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9
                    z = (z ^ (z >> 27)) * 0x94D049BB133111EB
                    z ^ (z >> 31)                              (all mod 2^64)
    base(s, c)    = mix64(s * 0x9E3779B97F4A7C15 + c)          seed s, column id c
    r(s, c, i)    = mix64(base(s, c) + (i + 1) * 0x9E3779B97F4A7C15)
    uniform_int(lo, hi) [lo, hi) = lo + r % (hi - lo)
    unit double   = (r >> 11) * 2^-53
    permutation   = stable argsort of r over [0, n)
    zipf(theta,n) = rank k in [1, n] by inverse CDF of k^-theta (theta 0 = uniform)

The same arrays feed the GPU path and the CPU oracle.
"""
from __future__ import annotations

import numpy as np

M64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15
SPAN_START = 2448623          # Julian day number of 1992-01-01 (types::Date)
SPAN_DAYS = 2526


def mix64_int(z: int) -> int:
    z &= M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def _mix64(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def rand_u64(seed: int, column: int, n: int, chunk: int = 1 << 24) -> np.ndarray:
    """r(seed, column, i) for i in [0, n)."""
    base = mix64_int((seed * GOLDEN + column) & M64)
    out = np.empty(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        for s in range(0, n, chunk):
            e = min(n, s + chunk)
            i = np.arange(s + 1, e + 1, dtype=np.uint64)
            out[s:e] = _mix64(np.uint64(base) + i * np.uint64(GOLDEN))
    return out


def uniform_int(seed, column, n, lo, hi, dtype):
    r = rand_u64(seed, column, n)
    return (np.int64(lo) + (r % np.uint64(hi - lo)).astype(np.int64)).astype(dtype)


def unit_double(seed, column, n):
    return (rand_u64(seed, column, n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def permutation(seed, column, n):
    return np.argsort(rand_u64(seed, column, n), kind="stable")


def zipf_ranks(seed, column, n, items, theta):
    """Ranks in [1, items]; theta == 0 is uniform."""
    if theta == 0:
        return uniform_int(seed, column, n, 1, items + 1, np.int64)
    w = np.arange(1, items + 1, dtype=np.float64) ** (-float(theta))
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    u = unit_double(seed, column, n)
    return np.minimum(np.searchsorted(cdf, u, side="right"), items - 1).astype(np.int64) + 1


def _col(arr, lo=None, hi=None):
    """(array, min, max) with min/max taken from the data (Import.cpp:175-186)."""
    if arr.size == 0:
        return arr, 0, 0
    return arr, int(arr.min()), int(arr.max())


# --------------------------------------------------------------------------
# Workload inputs
# --------------------------------------------------------------------------

def gen_filter_groupby(rows=6_000_000, seed=42):
    """Filter + low-cardinality group-by (TPC-H Q1 shape)."""
    t = {
        "flag": uniform_int(seed, 1, rows, 0, 3, np.int32),
        "status": uniform_int(seed, 2, rows, 0, 2, np.int32),
        "qty": (uniform_int(seed, 3, rows, 1, 51, np.int64) * 100).astype(np.int64),
        "price": uniform_int(seed, 4, rows, 90000, 10500000, np.int64),
        "disc": uniform_int(seed, 5, rows, 0, 11, np.int64),
        "date": uniform_int(seed, 6, rows, SPAN_START, SPAN_START + SPAN_DAYS, np.int32),
    }
    return t


def gen_join_aggregate(build_rows=1_500_000, probe_rows=6_000_000, seed=42):
    """Single hash join + 5-group aggregate."""
    build = {
        "key": (permutation(seed, 11, build_rows) + 1).astype(np.int32),
        "nation": uniform_int(seed, 12, build_rows, 0, 25, np.int32),
        "segment": uniform_int(seed, 13, build_rows, 0, 5, np.int32),
    }
    probe = {
        "fk": uniform_int(seed, 14, probe_rows, 1, build_rows + 1, np.int32),
        "date": uniform_int(seed, 15, probe_rows, SPAN_START, SPAN_START + SPAN_DAYS, np.int32),
        "price": uniform_int(seed, 16, probe_rows, 90000, 10500000, np.int64),
    }
    return build, probe


def sparse_keys(seed, column, n, bits=31):
    """n distinct keys over [0, 2^bits): a bijection of [0, 2^bits) (odd
    multiplies and xorshifts modulo 2^bits) applied to a seeded offset plus
    0..n-1."""
    m = np.uint64((1 << bits) - 1)
    c = np.uint64(mix64_int((seed * GOLDEN + column) & M64) & ((1 << bits) - 1))
    x = (np.arange(n, dtype=np.uint64) + c) & m
    with np.errstate(over="ignore"):
        x = (x * np.uint64(0x2545F491)) & m
        x ^= x >> np.uint64(bits // 2)
        x = (x * np.uint64(0x9E3779B1)) & m
        x ^= x >> np.uint64(bits // 3)
        x = (x * np.uint64(0x85EBCA6B)) & m
    return x.astype(np.int64)


def gen_join_sparse(build_rows=1_500_000, probe_rows=6_000_000, seed=42):
    """The join of gen_join_aggregate with build keys drawn without
    repetition over [0, 2^31) instead of 1..build_rows (a key range far too
    wide for a direct table: the join takes a hash table); probe fks are
    uniform over the build keys."""
    keys = sparse_keys(seed, 17, build_rows).astype(np.int32)
    build = {
        "key": keys,
        "nation": uniform_int(seed, 12, build_rows, 0, 25, np.int32),
        "segment": uniform_int(seed, 13, build_rows, 0, 5, np.int32),
    }
    pick = uniform_int(seed, 14, probe_rows, 0, build_rows, np.int64)
    probe = {
        "fk": keys[pick],
        "date": uniform_int(seed, 15, probe_rows, SPAN_START, SPAN_START + SPAN_DAYS, np.int32),
        "price": uniform_int(seed, 16, probe_rows, 90000, 10500000, np.int64),
    }
    return build, probe


def _years_of_days(days: np.ndarray) -> np.ndarray:
    """Calendar year of Julian day numbers (proleptic Gregorian)."""
    a = days.astype(np.int64) + 32044
    b = (4 * a + 3) // 146097
    c = a - 146097 * b // 4
    d = (4 * c + 3) // 1461
    e = c - 1461 * d // 4
    m = (5 * e + 2) // 153
    month = m + 3 - 12 * (m // 10)
    return (100 * b + d - 4800 + m // 10).astype(np.int32) + 0 * month


def gen_star(fact_rows=60_000_000, zipf_part=0.0, seed=42, dims_scale=1.0):
    """Star schema: fact with four FKs; date (2,556 days from 1992-01-01,
    year 1992..1998), customer / supplier (region U5, nation = 5 * region +
    U5), part (category U25, brand U1000)."""
    n_date = 2556
    n_cust = max(1, int(300_000 * dims_scale))
    n_supp = max(1, int(20_000 * dims_scale))
    n_part = max(1, int(1_400_000 * dims_scale))
    dkey = np.arange(SPAN_START, SPAN_START + n_date, dtype=np.int32)
    date = {"d_key": dkey, "d_year": _years_of_days(dkey)}
    c_region = uniform_int(seed, 21, n_cust, 0, 5, np.int32)
    customer = {"c_key": (permutation(seed, 22, n_cust) + 1).astype(np.int32),
                "c_region": c_region,
                "c_nation": (c_region * 5 + uniform_int(seed, 23, n_cust, 0, 5, np.int32)).astype(np.int32)}
    s_region = uniform_int(seed, 24, n_supp, 0, 5, np.int32)
    supplier = {"s_key": (permutation(seed, 25, n_supp) + 1).astype(np.int32),
                "s_region": s_region,
                "s_nation": (s_region * 5 + uniform_int(seed, 26, n_supp, 0, 5, np.int32)).astype(np.int32)}
    part = {"p_key": (permutation(seed, 27, n_part) + 1).astype(np.int32),
            "p_category": uniform_int(seed, 28, n_part, 0, 25, np.int32),
            "p_brand": uniform_int(seed, 29, n_part, 0, 1000, np.int32)}
    fact = {
        "f_date": uniform_int(seed, 31, fact_rows, SPAN_START, SPAN_START + n_date, np.int32),
        "f_cust": uniform_int(seed, 32, fact_rows, 1, n_cust + 1, np.int32),
        "f_supp": uniform_int(seed, 33, fact_rows, 1, n_supp + 1, np.int32),
        "f_part": zipf_ranks(seed, 34, fact_rows, n_part, zipf_part).astype(np.int32),
        "f_revenue": uniform_int(seed, 35, fact_rows, 100_000, 10_000_000, np.int64),
        "f_cost": uniform_int(seed, 36, fact_rows, 10_000, 1_000_000, np.int64),
    }
    return fact, date, customer, supplier, part


def batch_params(count=256, seed=42, probe_range=(1e4, 1e7), build_range=(1e3, 2e6),
                 group_range=(1, 1e6)):
    """Parameters of `count` uneven select -> join -> group-by pipelines:
    probe / build rows and group counts log-uniform, selectivity uniform in
    [0.01, 0.99], key skew theta from {0, 0.5, 1.0, 1.5}."""
    u = [unit_double(seed, 41 + k, count) for k in range(5)]

    def logu(x, lo, hi):
        return np.exp(np.log(lo) + x * (np.log(hi) - np.log(lo)))
    out = []
    thetas = [0.0, 0.5, 1.0, 1.5]
    for i in range(count):
        out.append({
            "probe_rows": int(logu(u[0][i], *probe_range)),
            "build_rows": int(logu(u[1][i], *build_range)),
            "selectivity": 0.01 + 0.98 * float(u[2][i]),
            "groups": max(1, int(logu(u[3][i], *group_range))),
            "theta": thetas[int(u[4][i] * 4) % 4],
        })
    return out


def gen_batch_pipeline(params: dict, seed=42, index=0):
    """Inputs of one batched pipeline: build (key unique 1..m, w U[0,100)),
    probe (fk Zipf(theta) over [1, 2m] so about half the probes match under
    no skew, a U[0,1000) for the selection, g U[0,G), v int64)."""
    m, n = params["build_rows"], params["probe_rows"]
    s = seed * 1000 + index
    build = {"key": (permutation(s, 51, m) + 1).astype(np.int32),
             "w": uniform_int(s, 52, m, 0, 100, np.int32)}
    probe = {"fk": zipf_ranks(s, 53, n, 2 * m, params["theta"]).astype(np.int32),
             "a": uniform_int(s, 54, n, 0, 1000, np.int32),
             "g": uniform_int(s, 55, n, 0, params["groups"], np.int32),
             "v": uniform_int(s, 56, n, -10**9, 10**12, np.int64)}
    return build, probe


# The colour words dbgen draws p_name from (TPC-H specification 4.2.3).
P_NAME_WORDS = (
    "almond antique aquamarine azure beige bisque black blanched blue blush brown burlywood burnished "
    "chartreuse chiffon chocolate coral cornflower cornsilk cream cyan dark deep dim dodger drab firebrick "
    "floral forest frosted gainsboro ghost goldenrod green grey honeydew hot indian ivory khaki lace "
    "lavender lawn lemon light lime linen magenta maroon medium metallic midnight mint misty moccasin "
    "navajo navy olive orange orchid pale papaya peach peru pink plum powder puff purple red rose rosy "
    "royal saddle salmon sandy seashell sienna sky slate smoke snow spring steel tan thistle tomato "
    "turquoise violet wheat white yellow").split()


def gen_part_names(rows: int, seed=42, distinct=1 << 16) -> np.ndarray:
    """p_name-like varchar(55) column, (rows, 56) uint8 slots: `distinct`
    names of five different colour words, each row one of them uniformly."""
    from .varchar import pack
    rng = np.random.default_rng(seed)
    names = [" ".join(P_NAME_WORDS[c] for c in rng.choice(len(P_NAME_WORDS), 5, replace=False))
             for _ in range(distinct)]
    base = pack(names, 55)
    return base[(rand_u64(seed, 41, rows) % np.uint64(distinct)).astype(np.int64)]
