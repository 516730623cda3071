"""Build recipe for libvoila_hip.so (gfx950).  hipcc cross-compiles without a GPU."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libvoila_hip.so")
SOURCES = ["vh_pipeline.hip", "vh_primitives.hip", "vh_varchar.hip", "vh_batch.hip", "vh_mlp.hip", "vh_engine.cpp",
           "vh_capi.cpp", "vh_capi_prims.cpp", "vh_batch_host.cpp", "vh_capi_tune.cpp", "vh_relalg.cpp",
           "vh_jit.cpp", "vh_voila.cpp"]
# vh_modes.hip holds the generic kernels of one group-by strategy and is
# compiled once per strategy (GroupMode values of vh_plan.hpp), in parallel
MODES = {"reg": 0, "lds": 1, "global": 2, "hash": 3, "ldshash": 4, "part": 6}
# Headers the generated kernels include, embedded in the library as text
# (name as #included, path): hipRTC compiles against these copies.
JIT_HEADERS = [("vh_fused.hpp", os.path.join(CSRC, "vh_fused.hpp")),
               ("vh_device.hpp", os.path.join(CSRC, "vh_device.hpp")),
               ("vh_plan.hpp", os.path.join(CSRC, "vh_plan.hpp")),
               ("../../include/voila_hip.h", os.path.join(HERE, "..", "include", "voila_hip.h"))]
JIT_INC = os.path.join(CSRC, "vh_jit_headers.inc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def write_jit_headers():
    """vh_jit_headers.inc: the device headers as raw string literals."""
    delim = "VHJIT"
    parts = ["// generated by build.py: device headers for hipRTC\n",
             f"static const int kJitHeaders = {len(JIT_HEADERS)};\n",
             "static const char* const kJitHeaderNames[] = {" +
             ", ".join(f'"{n}"' for n, _ in JIT_HEADERS) + "};\n",
             "static const char* const kJitHeaderText[] = {\n"]
    for _, path in JIT_HEADERS:
        with open(path) as f:
            text = f.read()
        if f"){delim}\"" in text:
            raise RuntimeError(f"{path} contains the raw-string delimiter")
        parts.append(f'R"{delim}({text}){delim}",\n')
    parts.append("};\n")
    data = "".join(parts)
    old = open(JIT_INC).read() if os.path.exists(JIT_INC) else None
    if old != data:
        with open(JIT_INC, "w") as f:
            f.write(data)


def _units():
    """(source, object, extra flags) of every translation unit."""
    units = [(src, os.path.join(CSRC, os.path.splitext(src)[0] + ".o"), []) for src in SOURCES]
    for name, mode in MODES.items():
        units.append(("vh_modes.hip", os.path.join(CSRC, f"vh_modes_{name}.o"), [f"-DVH_MODE={mode}"]))
    return units


def _stale(obj, src):
    """An object older than its source or any header of csrc/ or include/."""
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [os.path.join(CSRC, src)] + [os.path.join(CSRC, f) for f in os.listdir(CSRC)
                                         if f.endswith((".hpp", ".inc"))]
    deps.append(os.path.join(HERE, "..", "include", "voila_hip.h"))
    return any(os.path.getmtime(d) > t for d in deps)


def build(verbose=False, incremental=False):
    """Compile every translation unit and link the library.  With
    `incremental` (a developer convenience), objects newer than their sources
    and headers are kept and reused."""
    write_jit_headers()
    objs, procs = [], []
    for src, obj, extra in _units():
        path = os.path.join(CSRC, src)
        objs.append(obj)
        if incremental and not _stale(obj, src):
            continue
        cmd = [HIPCC, *FLAGS, *extra, "-x", "hip", "-c", path, "-o", obj]
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = []
    for src, p in procs:
        out = p.communicate()[0]
        if p.returncode:
            sys.stderr.write(out)
            failed.append(src)
    if failed:
        raise RuntimeError(f"hipcc failed on {failed}")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs,
           "-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    if not incremental:
        for o in objs:
            os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(incremental="--incremental" in sys.argv))
