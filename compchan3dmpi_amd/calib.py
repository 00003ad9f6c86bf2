"""Copy kernel of known size, run under rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE
to calibrate the HBM-traffic counters (bytes the counters report for a
plain device-to-device copy of N bytes read + N bytes written).

    python -m compchan3dmpi_amd.calib [MiB]
"""
import sys

import torch


def main():
    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 1024
    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB x3", flush=True)


if __name__ == "__main__":
    main()
