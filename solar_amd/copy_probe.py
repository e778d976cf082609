"""Copies a 1 GiB FP64 buffer on the GPU a few times with a plain device
copy (torch's copy kernel): the reference transfer for calibrating the
rocprofv3 FETCH_SIZE / WRITE_SIZE counters against a known byte count
(each copy reads and writes exactly 2**30 bytes).

    rocprofv3 --pmc FETCH_SIZE -- python3 solar_amd/copy_probe.py
"""
import torch


def main(reps=4, nbytes=1 << 30):
    x = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(f"copy_probe: {reps} copies of {nbytes} bytes", flush=True)


if __name__ == "__main__":
    main()
