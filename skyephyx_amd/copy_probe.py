"""Calibration kernel for HBM counters: one device copy of a known size.

rocprofv3 --pmc FETCH_SIZE (or WRITE_SIZE) over this script reads the counter
for a copy that moves exactly BYTES in and BYTES out, which gives the scale
factor between the counter and real HBM bytes on this device.
"""
import sys

import torch

BYTES = 512 * 1024 * 1024  # 512 MiB read + 512 MiB written


def main():
    n = BYTES // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print(f"copied {BYTES} bytes", file=sys.stderr)


if __name__ == "__main__":
    main()
