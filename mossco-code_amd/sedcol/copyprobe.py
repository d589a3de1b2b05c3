"""Plain device-to-device copy of a known size, for calibrating HBM byte
counters: run under ``rocprofv3 --pmc FETCH_SIZE`` (then ``WRITE_SIZE``) and
divide the counter by the copied bytes to get the scale applied to the
counters of the sediment kernels.

    python -m sedcol.copyprobe [MiB]        (from mossco-code_amd/)
"""
from __future__ import annotations

import sys


def main(argv=None):
    import torch

    argv = sys.argv[1:] if argv is None else argv
    mib = int(argv[0]) if argv else 256
    n = mib * (1 << 20) // 8
    a = torch.arange(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copyprobe: 3 copies of {mib} MiB ({n * 8} bytes read and written each)", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
