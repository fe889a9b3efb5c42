"""Copy-kernel calibration for HBM counters.

    python -m marmel_amd.calib [MiB]

Streams a buffer of known size through two plain kernels, a device copy
(torch's copy_, the runtime's copyBuffer kernel) and an elementwise multiply
(y = 2x), each reading exactly N bytes and writing exactly N bytes, so that
rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE readings of the render kernels can be
scaled by what the same counters report for known traffic.
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
    torch.mul(x, 2.0, out=y)
    torch.cuda.synchronize()
    print(f"calib: copy and multiply, each read {n * 4} bytes and wrote {n * 4} bytes", flush=True)


if __name__ == "__main__":
    main()
