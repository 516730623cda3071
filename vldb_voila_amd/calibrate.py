"""Counter calibration: one plain device-to-device copy of 256 MiB.

    rocprofv3 --pmc FETCH_SIZE -- python3 -m vldb_voila_amd.calibrate
    rocprofv3 --pmc WRITE_SIZE -- python3 -m vldb_voila_amd.calibrate

The copy reads and writes exactly 268,435,456 bytes, so what FETCH_SIZE and
WRITE_SIZE report for its dispatch (in KiB) gives the factor that turns the
counters of the probe kernels into HBM bytes.  PyTorch only provides the
buffers and the copy here; nothing of the operator path runs."""
from __future__ import annotations

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


if __name__ == "__main__":
    main()
