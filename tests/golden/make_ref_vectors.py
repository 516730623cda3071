"""Writes tests/golden/ref_vectors.json from the reference's own runtime
(oracle/_ref/libvoila_ref.so, compiled from /root/reference/runtime.cpp and
runtime.hpp by `make -C oracle ref`): voila_hash / rehash values,
voila_cast::to_cstr strings, select_true / select_false and
ComplexFuncs::selunion outputs on seeded inputs.

Run:  python tests/golden/make_ref_vectors.py
"""
import ctypes as C
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import ref  # noqa: E402


def main():
    assert ref.available(), "oracle/_ref could not be built"
    L = ref.lib()
    rng = np.random.default_rng(20240601)
    out = {}
    vals32 = [0, 1, -1, 2, 7, 42, 65535, 65536, 2**31 - 1, -2**31, 123456789, -987654321]
    vals32 += [int(x) for x in rng.integers(-2**31, 2**31, 40)]
    out["hash_i32"] = [[v, L.ref_hash_i32(v)] for v in vals32]
    valsu32 = [v & 0xFFFFFFFF for v in vals32]
    out["hash_u32"] = [[v, L.ref_hash_u32(v)] for v in valsu32]
    vals64 = [0, 1, -1, 2**63 - 1, -2**63, 10**18, -(10**17)] + [int(x) for x in rng.integers(-2**63, 2**63 - 1, 40)]
    out["hash_i64"] = [[v, L.ref_hash_i64(v)] for v in vals64]
    out["rehash_i32"] = []
    for _ in range(40):
        h = int(rng.integers(0, 2**63)) * 2 + int(rng.integers(0, 2))
        v = int(rng.integers(-2**31, 2**31))
        out["rehash_i32"].append([h, v, L.ref_rehash_i32(h, v)])
    buf = C.create_string_buffer(64)
    i128s = [0, 1, -1, 10, -10, 2**64, -(2**64), 2**127 - 1, -(2**127), 2**127 - 12345,
             123456789012345678901234567890, -98765432109876543210]
    i128s += [int(x) * 2**40 + int(y) for x, y in zip(rng.integers(-2**62, 2**62, 20), rng.integers(0, 2**40, 20))]
    out["to_cstr_i128"] = []
    for v in i128s:
        u = v & ((1 << 128) - 1)
        L.ref_to_cstr_i128(u & (2**64 - 1), u >> 64, buf)
        out["to_cstr_i128"].append([str(v), buf.value.decode()])
    out["to_cstr_i64"] = []
    for v in [0, -1, 2**63 - 1, -2**63, 42, -42]:
        L.ref_to_cstr_i64(v, buf)
        out["to_cstr_i64"].append([v, buf.value.decode()])
    out["to_cstr_u64"] = []
    for v in [0, 1, 2**64 - 1, 2**63, 12345678901234567890]:
        L.ref_to_cstr_u64(v, buf)
        out["to_cstr_u64"].append([str(v), buf.value.decode()])

    sel_cases = []
    for n, density, use_sel in [(0, 0.5, False), (1, 1.0, False), (17, 0.5, False), (1024, 0.3, False),
                                (1024, 0.0, True), (1024, 1.0, True), (1000, 0.5, True), (33, 0.9, True)]:
        pred = (rng.random(1024) < density).astype(np.uint8)
        if use_sel:
            sel = np.sort(rng.choice(1024, size=n, replace=False)).astype(np.int32)
        else:
            sel = None
        res_t = np.zeros(1024, dtype=np.uint32)
        res_f = np.zeros(1024, dtype=np.uint32)
        sp = sel.ctypes.data if sel is not None else None
        kt = L.ref_select_true(sp, n, res_t.ctypes.data, pred.ctypes.data)
        kf = L.ref_select_false(sp, n, res_f.ctypes.data, pred.ctypes.data)
        sel_cases.append({"n": n, "sel": None if sel is None else sel.tolist(), "pred": pred.tolist(),
                          "true": res_t[:kt].tolist(), "false": res_f[:kf].tolist()})
    out["select"] = sel_cases

    union_cases = []
    for ln, rn in [(0, 0), (0, 5), (5, 0), (10, 10), (300, 700), (1024, 1024), (1, 1000)]:
        l = np.sort(rng.choice(2048, size=ln, replace=False)).astype(np.int32)
        r = np.sort(rng.choice(2048, size=rn, replace=False)).astype(np.int32)
        res = np.zeros(ln + rn + 1, dtype=np.int32)
        k = L.ref_selunion(res.ctypes.data, l.ctypes.data if ln else None, ln,
                           r.ctypes.data if rn else None, rn) if (ln or rn) else 0
        union_cases.append({"l": l.tolist(), "r": r.tolist(), "union": res[:k].tolist()})
    out["selunion"] = union_cases

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_vectors.json")
    with open(path, "w") as f:
        json.dump(out, f)
    print("wrote", path)


if __name__ == "__main__":
    main()
