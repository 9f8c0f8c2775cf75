"""Plain device-to-device copy of a known size, run under
`rocprofv3 --pmc FETCH_SIZE` (or WRITE_SIZE) to calibrate what the counters
read for a known number of HBM bytes.

    rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT -o copy -- \
        python3 warp-reactor_amd/prof/copy_calibration.py

Three copies of 2^27 doubles (1 GiB each: 1 GiB read + 1 GiB written).
"""
import torch


def main():
    n = 1 << 27
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {n * 8} bytes 3 times", flush=True)


if __name__ == "__main__":
    main()
