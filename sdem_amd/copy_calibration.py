"""A plain device-to-device copy of known size.

It calibrates the HBM byte counters (FETCH_SIZE, WRITE_SIZE) that rocprofv3
reports for the DEM kernels: run it under the same `rocprofv3 --pmc` command
and divide the counter by the bytes moved.  It also prints the copy's own
bandwidth (bytes read + written per second).  The runtime's copy is not the
fastest streaming kernel on this part: `sdem_amd/hbm_probe` measures the
rates of plain copy, read and write kernels, the ceiling the DEM kernels'
traffic is compared with.  Uses torch only for the allocation and the
copy.

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
    reps = 20
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        b.copy_(a)
    t1.record()
    torch.cuda.synchronize()
    sec = t0.elapsed_time(t1) / 1e3 / reps
    print(f"copied {mib} MiB {5 + reps} times: {n * 8} bytes read and written per copy; "
          f"{2 * n * 8 / sec / 1e9:.1f} GB/s (read + written)", flush=True)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 512)
