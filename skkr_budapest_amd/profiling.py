"""Small workloads for rocprofv3 runs of single kernels.

  python -m skkr_budapest_amd.profiling copy     # plain device copy, 256 MiB
  python -m skkr_budapest_amd.profiling gemm     # 1024 complex GEMMs 128^3
  python -m skkr_budapest_amd.profiling inverse  # 1024 inversions 128 x 128
  python -m skkr_budapest_amd.profiling split    # the same GEMMs: FP64, then split-BF16 with 2..5 slices
  python -m skkr_budapest_amd.profiling int8     # the same GEMMs: FP64, then INT8 digits, 4..6 per operand

The copy moves a known number of bytes, so its FETCH_SIZE / WRITE_SIZE
counters calibrate those of the other kernels.  Each workload runs 3 times.
"""
from __future__ import annotations

import sys

import numpy as np

COPY_BYTES = 256 << 20


def run_copy():
    import torch
    src = torch.empty(COPY_BYTES // 4, dtype=torch.float32, device="cuda:0").fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copy: {COPY_BYTES} bytes read and written per dispatch", flush=True)


def _mats(n, nb, seed):
    rng = np.random.default_rng(seed)
    h = rng.normal(size=(nb, n, n)) + 1j * rng.normal(size=(nb, n, n))
    return h


def run_gemm(n=128, nb=1024):
    from . import gpu
    a, b = _mats(n, nb, 1), _mats(n, nb, 2)
    with gpu.Context(lk=n, sites_per_pl=1, n_pl=1) as ctx:
        for _ in range(3):
            ctx.zgemm(a, b)
    print(f"gemm: {nb} x complex {n}^3, algorithmic bytes per dispatch "
          f"{3 * nb * n * n * 16} (A, B read, C written)", flush=True)


def run_split(n=128, nb=1024, slices=(0, 2, 3, 4, 5)):
    from . import gpu
    a, b = _mats(n, nb, 1), _mats(n, nb, 2)
    with gpu.Context(lk=n, sites_per_pl=1, n_pl=1) as ctx:
        for s in slices:
            ctx.set_option("zgemm_split", s)
            for _ in range(3):
                ctx.zgemm(a, b)
            print(f"split: {nb} x complex {n}^3 with {s} slices per operand (0: FP64 kernel)", flush=True)


def run_int8(n=128, nb=1024, digits=(0, 4, 5, 6), tiles=(128, 64)):
    from . import gpu
    a, b = _mats(n, nb, 1), _mats(n, nb, 2)
    with gpu.Context(lk=n, sites_per_pl=1, n_pl=1) as ctx:
        for t in tiles:
            ctx.set_option("int8_tile", t)
            for s in digits:
                if s == 0 and t != tiles[0]:
                    continue
                ctx.set_option("zgemm_int8", s)
                for _ in range(3):
                    ctx.zgemm(a, b)
                print(f"int8: {nb} x complex {n}^3 with {s} digits per operand (0: FP64 kernel), "
                      f"INT8 tile {t}", flush=True)


def run_inverse(n=128, nb=1024):
    from . import gpu
    h = _mats(n, nb, 3)
    a = (0.3 + 0.05j) * np.eye(n) - (h + np.conj(np.swapaxes(h, 1, 2))) / 2
    with gpu.Context(lk=n, sites_per_pl=1, n_pl=1) as ctx:
        for _ in range(3):
            ctx.gjinv(a)
    print(f"inverse: {nb} x {n} x {n}, algorithmic flops per dispatch {8 * n ** 3 * nb}",
          flush=True)


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "gemm"
    {"copy": run_copy, "gemm": run_gemm, "inverse": run_inverse, "split": run_split,
     "int8": run_int8}[what]()
