"""Plain device-to-device copy of a known size, used to calibrate rocprofv3's
FETCH_SIZE / WRITE_SIZE counters against bytes actually moved (the factor
that converts the score gate's counter readings into HBM bytes).

Usage: python -m hspz_amd.copyprobe [MiB]   (default 1024 MiB, cuda:0)
"""
import sys

import torch


def main():
    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 1024
    n = mib << 20
    a = torch.empty(n, dtype=torch.uint8, device="cuda:0")
    a.fill_(1)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print("copied %d bytes" % n, flush=True)


if __name__ == "__main__":
    main()
