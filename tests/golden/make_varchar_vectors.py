"""Writes tests/golden/varchar_vectors.json from the reference's own runtime
(oracle/_ref/libvoila_ref.so, built from /root/reference/runtime.hpp by
`make -C oracle ref`): voila_hash<varchar>::hash and rehash (MurmurHash64A),
varchar::contains and varchar::operator== on seeded strings, read from
(1 + max_len)-byte slots through varchar(void* data, idx, max_len).

Run:  python tests/golden/make_varchar_vectors.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import ref  # noqa: E402
from vldb_voila_amd.varchar import pack  # noqa: E402

MAX_LEN = 40


def strings(rng):
    out = [b"", b"a", b"ab", b"green", b"forest green", b"\x00", b"ab\x00cd", b"ab\x00ce",
           b"BUILDING", b"BUILDING ", b" BUILDING"]
    out += [bytes(rng.integers(0, 256, n, dtype=np.uint8)) for n in range(0, 18)]
    out += [bytes(rng.choice(np.frombuffer(b"ab", dtype=np.uint8), n)) for n in (3, 7, 8, 9, 15, 16, 17, 33, 40)]
    return out


def main():
    assert ref.available(), "oracle/_ref could not be built"
    L = ref.lib()
    rng = np.random.default_rng(20240917)
    vals = strings(rng)
    col = pack(vals, MAX_LEN)
    p = col.ctypes.data
    out = {"max_len": MAX_LEN, "strings": [v.hex() for v in vals]}
    out["hash"] = [str(L.ref_hash_varchar(p, i, MAX_LEN)) for i in range(len(vals))]
    seeds = [int(rng.integers(0, 2**63)) * 2 + 1 for _ in vals]
    out["rehash"] = [[str(h), str(L.ref_rehash_varchar(h, p, i, MAX_LEN))] for i, h in enumerate(seeds)]
    needles = [b"", b"a", b"b", b"ab", b"aab", b"green", b"\x00", b"ING", b"BUILDING", b"x" * 41]
    out["contains"] = [[n.hex(), [L.ref_varchar_contains(p, i, MAX_LEN, n, len(n)) for i in range(len(vals))]]
                       for n in needles]
    eq = []
    for i in range(len(vals)):
        for j in range(len(vals)):
            eq.append(L.ref_varchar_eq(p, i, MAX_LEN, p, j, MAX_LEN))
    out["eq"] = eq
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "varchar_vectors.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
