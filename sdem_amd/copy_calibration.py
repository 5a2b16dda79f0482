"""A plain device-to-device copy of known size, used to calibrate the HBM
byte counters (FETCH_SIZE, WRITE_SIZE) that rocprofv3 reports for the DEM
kernels: run it under the same `rocprofv3 --pmc` command and divide the
counter by the bytes moved.  Uses torch only for the allocation and copy.

    python3 -m sdem_amd.copy_calibration [MiB]
"""
import sys


def main(mib: int = 512) -> None:
    import torch

    n = mib * 1024 * 1024 // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB five times: {n * 8} bytes read and written per copy", flush=True)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 512)
