"""Plain device-to-device copy of a known size, for calibrating HBM traffic
counters: under `rocprofv3 --pmc FETCH_SIZE` (or WRITE_SIZE) the copy
kernel's counter per dispatch, in KiB, divided by the bytes it must move
gives the factor that turns the same counter of another kernel into bytes.

    python -m pixelforge_amd.copy_probe [MiB]
"""
import sys

import torch


def main():
    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 256
    n = mib * (1 << 20) // 4
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {n * 4} bytes x3", flush=True)


if __name__ == "__main__":
    main()
