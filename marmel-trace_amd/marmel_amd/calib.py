"""Copy-kernel calibration for HBM counters.

    python -m marmel_amd.calib [MiB]

Runs one device-to-device copy of a known size through a plain elementwise
copy kernel (torch's copy_), so that rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
readings of the render kernels can be scaled by what the same counters report
for exactly N bytes read and N bytes written.
"""
import sys


def main():
    import torch

    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 1024
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print(f"calib: copied {n * 4} bytes (read {n * 4}, write {n * 4})", flush=True)


if __name__ == "__main__":
    main()
