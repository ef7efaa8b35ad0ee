"""Short driver for rocprofv3 runs of the matvec passes.

Runs `reps` device matvecs in each precision on the config-2 box
(100^3 inhomogeneous cube, 200^3 FFT grid) and, for counter calibration, a
plain device copy of a known number of bytes (torch's copy kernel), so that
FETCH_SIZE / WRITE_SIZE per dispatch can be scaled by what the counters read
for the copy:

    rocprofv3 --kernel-trace --stats -d OUT -o run -- python3 if-dda_amd/ifdda/prof.py
    rocprofv3 --pmc FETCH_SIZE -d OUT -o fetch -- python3 if-dda_amd/ifdda/prof.py
    rocprofv3 --pmc WRITE_SIZE -d OUT -o write -- python3 if-dda_amd/ifdda/prof.py
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=100)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--copy-mb", type=int, default=512)
    args = ap.parse_args()
    import torch

    import ifdda
    from ifdda import scene

    torch.cuda.set_device(0)
    s, k0, alpha, E = scene.config2(args.n)
    sv = ifdda.Solver(s.nx, s.ny, s.nz)
    sv.greencalculfft(s.d, k0)
    sv.set_dipoles(s.tabdip)
    sv.set_polarizability(alpha)
    for prec in (64, 32):
        cdt = np.complex128 if prec == 64 else np.complex64
        rdt = np.float64 if prec == 64 else np.float32
        x = torch.from_numpy(E.astype(cdt).view(rdt).copy()).cuda()
        y = torch.empty_like(x)
        for _ in range(args.reps):
            sv.matvec_device(x.data_ptr(), x.data_ptr(), y.data_ptr(), prec=prec)
        torch.cuda.synchronize()
        print(f"matvec fp{prec} x{args.reps} done", flush=True)
    nbytes = args.copy_mb << 20
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"calibration copy {nbytes} bytes x3 done", flush=True)
    sv.close()


if __name__ == "__main__":
    main()
