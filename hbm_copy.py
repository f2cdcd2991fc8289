"""Plain device copy of a known size, for calibrating HBM byte counters.

    python hbm_copy.py [MiB] [repeats]

Copies a float32 buffer of MiB mebibytes into another one `repeats` times
(torch's copy kernel): each launch reads and writes exactly MiB MiB, so the
FETCH_SIZE / WRITE_SIZE counters of one launch, divided by MiB*1024 KiB,
give the counters' scale on this device.
"""
import sys

import torch


def main():
    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 256
    reps = int(sys.argv[2]) if len(sys.argv) > 2 else 3
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB x {reps}", flush=True)


if __name__ == "__main__":
    main()
