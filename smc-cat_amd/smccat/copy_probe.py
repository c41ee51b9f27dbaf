"""Plain device-to-device copy of a known size, used to calibrate the rocprofv3
FETCH_SIZE / WRITE_SIZE counters against bytes actually moved.

usage: python smc-cat_amd/smccat/copy_probe.py [MiB] [repeats]
Each repeat is one copy kernel dispatch reading and writing exactly MiB MiB.
"""
import sys

import torch


def main():
    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 1024
    reps = int(sys.argv[2]) if len(sys.argv) > 2 else 3
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print("copied %d MiB x %d" % (mib, reps), flush=True)


if __name__ == "__main__":
    main()
