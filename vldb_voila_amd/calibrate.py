"""Counter calibration: one plain device-to-device copy of 256 MiB.

    rocprofv3 --pmc FETCH_SIZE -- python3 -m vldb_voila_amd.calibrate
    rocprofv3 --pmc WRITE_SIZE -- python3 -m vldb_voila_amd.calibrate

The copy reads and writes exactly 268,435,456 bytes, so what FETCH_SIZE and
WRITE_SIZE report for its dispatch (in KiB) gives the factor that turns the
counters of the probe kernels into HBM bytes.  PyTorch only provides the
buffers and the copy here; nothing of the operator path runs.

    python3 -m vldb_voila_amd.calibrate --bandwidth

times 20 such copies from four rotating sources (1 GiB, past the 256 MB
Infinity Cache) and prints the streaming-copy rate of the device, read plus
written bytes per second: the practical ceiling the roofline figures sit
under."""
from __future__ import annotations

import json
import sys

import torch

BYTES = 256 * 2**20


def main():
    src = torch.full((BYTES // 8,), 7, dtype=torch.int64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    # rotate through a 1 GiB buffer first so the copy's source is not in the
    # 256 MB Infinity Cache
    flush = torch.empty(2**30, dtype=torch.uint8, device="cuda")
    flush.fill_(1)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    assert int(dst[-1]) == 7
    print(f"copied {BYTES} bytes", flush=True)


def bandwidth(steps=20):
    srcs = [torch.full((BYTES // 8,), i, dtype=torch.int64, device="cuda") for i in range(4)]
    dst = torch.empty_like(srcs[0])
    for i in range(4):
        dst.copy_(srcs[i])
    torch.cuda.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    times = []
    for i in range(steps):
        t0.record()
        dst.copy_(srcs[i % 4])
        t1.record()
        t1.synchronize()
        times.append(t0.elapsed_time(t1) / 1e3)
    times.sort()
    best, median = times[0], times[len(times) // 2]
    print(json.dumps({"copy_bytes": BYTES, "read_plus_write_GBps_best": 2 * BYTES / best / 1e9,
                      "read_plus_write_GBps_median": 2 * BYTES / median / 1e9}), flush=True)


if __name__ == "__main__":
    if "--bandwidth" in sys.argv[1:]:
        bandwidth()
    else:
        main()
