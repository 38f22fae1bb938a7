"""Device-to-device copy of a known size, used to calibrate rocprofv3's
FETCH_SIZE / WRITE_SIZE counters against bytes actually moved: the copy
reads and writes exactly `--mib` MiB per launch."""
import argparse

import torch


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=5)
    a = ap.parse_args()
    n = a.mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(a.reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copied {a.mib} MiB x {a.reps}")


if __name__ == "__main__":
    main()
