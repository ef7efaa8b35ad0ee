"""Benchmark of the FFT-accelerated DDA solve on one MI355X.

Workload (one "step" = one full solve): the 100^3-dipole inhomogeneous cube
(side 2 um, d = 20 nm, eps 2.25+0.01i with a Gaussian random real part,
std 0.1, correlation 100 nm), lambda 632.8 nm, circular plane wave at 30 deg,
padded FFT grid 200^3, GPBICG from x0 = 0 to 1e-4 relative residual.

The headline value is mixed-precision solves/s (FP32 matvec + FP64
refinement, true residual checked with the FP64 matvec); the plain FP64
solve, the matvec rates, solve latency percentiles, the roofline of the
dominant kernel (HIP-event durations of every matvec pass measured inside the
timed solves) and a CPU baseline (the reference's own produitfftmatvect3
compiled from source, timed on this host) are reported beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "if-dda_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "solves/s & matvec HBM GB/s, mixed vs FP64 at 1e-4 residual; p50/p99 latency"


def pass_bytes(nx, ny, nz, L, nd, s, alpha_words=1, tab=True):
    """Algorithmic HBM bytes of the five matvec passes (s bytes per complex).

    P1 reads x (3 nd), alpha and the dipole map, writes the x-transformed
    planes A (nz*ny*L1*3); P2 reads A, writes B (nz*L2*L1*3); P3 reads and
    writes B and reads the folded Green tensor ((L/2+1)^3 * 6); P4 reads B,
    writes A; P5 reads A, the map and b (3 nd), writes y (3 nd)."""
    L1, L2, L3 = L
    A = nz * ny * L1 * 3 * s
    B = nz * L2 * L1 * 3 * s
    Gf = (L1 // 2 + 1) * (L2 // 2 + 1) * (L3 // 2 + 1) * 6 * s
    t = 4 * nx * ny * nz if tab else 0
    return [3 * nd * s + alpha_words * nd * s + t + A, A + B, 2 * B + Gf, B + A,
            A + t + 2 * 3 * nd * s]


def cpu_baseline(sv, nx, ny, nz, tabdip, nd, ncompte_fp64, budget_s=15.0):
    """The reference's produitfftmatvect3 (oracle/_ref, compiled from the
    reference's Fortran) on this host for a bounded number of calls; scaled
    to solves/s with the FP64 solve's matvec count.  Falls back to the
    oracle's C restatement when the reference build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle, Ref  # the checker, timed only as the CPU baseline

    G = sv.green_get()  # reference layout, FP64 (equal to greencalculfft's output)
    rng = np.random.default_rng(0)
    x = rng.standard_normal(3 * nd) + 1j * rng.standard_normal(3 * nd)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    if Ref.available():
        r = Ref()
        kind, run = "reference", lambda: r.matvec3(G, tabdip, nx, ny, nz, x, x)
    else:
        o = Oracle()
        threads = 1
        kind, run = "port", lambda: o.matvec(G, tabdip, nx, ny, nz, x, x)
    times = []
    t_start = time.perf_counter()
    while len(times) < 2 or (time.perf_counter() - t_start < budget_s and len(times) < 20):
        t0 = time.perf_counter()
        run()
        times.append(time.perf_counter() - t0)
    tmv = float(np.median(times))
    return {"value": 1.0 / (tmv * ncompte_fp64), "unit": "solves/s", "cores": threads,
            "kind": kind,
            "sample": f"{len(times)} produitfftmatvect3 calls on the {sv.L[0]}x{sv.L[1]}x{sv.L[2]}"
                      f" grid (median {tmv:.3f} s), scaled by the FP64 GPU solve's {ncompte_fp64}"
                      f" matvecs per solve; Krylov vector updates not included",
            "matvec_s": tmv}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=100, help="cube edge in dipoles")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(independent solves shard trivially across GPUs; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch

    import ifdda
    from ifdda import scene

    torch.cuda.set_device(0)
    n = args.n
    s, k0, alpha, E = scene.config2(n)
    sv = ifdda.Solver(s.nx, s.ny, s.nz)
    t0 = time.perf_counter()
    sv.greencalculfft(s.d, k0)
    green_s = time.perf_counter() - t0
    sv.set_dipoles(s.tabdip)
    sv.set_polarizability(alpha)
    nd = s.nd
    Ed = torch.from_numpy(E.view(np.float64).copy()).cuda()  # inputs resident in HBM
    FFd = torch.empty_like(Ed)
    FLd = torch.empty_like(Ed)

    def solve(prec):
        info = sv.solve_device(Ed.data_ptr(), FFd.data_ptr(), FLd.data_ptr(), "GPBICG1", 1e-4,
                               precision=prec, x0_zero=True)
        assert info["nstop"] == 0 and info["tol1"] <= 1e-4, info
        return info

    for _ in range(args.warmup):
        solve(32)
    info64 = solve(64)
    E64 = FLd.cpu().numpy().view(np.complex128).copy()

    # timed region: K mixed-precision solves, per-pass HIP events on
    sv.timing_enable(200000)
    lat = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t1 = time.perf_counter()
        info32 = solve(32)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t1)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_sum, cnt = sv.timing_read()
    sv.timing_enable(0)
    E32 = FLd.cpu().numpy().view(np.complex128)
    field_err = float(np.linalg.norm(E32 - E64) / np.linalg.norm(E64))

    # plain FP64 solves (same count, untimed by events)
    lat64 = []
    for _ in range(max(2, args.steps // 2)):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        info64 = solve(64)
        torch.cuda.synchronize()
        lat64.append(time.perf_counter() - t1)

    # matvec rates
    mv = {}
    for prec, dt in ((64, torch.float64), (32, torch.float32)):
        xd = torch.from_numpy(E.astype(np.complex128 if prec == 64 else np.complex64)
                              .view(np.float64 if prec == 64 else np.float32).copy()).cuda()
        yd = torch.empty_like(xd)
        for _ in range(3):
            sv.matvec_device(xd.data_ptr(), xd.data_ptr(), yd.data_ptr(), prec=prec)
        reps = 20
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(reps):
            sv.matvec_device(xd.data_ptr(), xd.data_ptr(), yd.data_ptr(), prec=prec)
        torch.cuda.synchronize()
        tm = (time.perf_counter() - t1) / reps
        byts = sum(pass_bytes(s.nx, s.ny, s.nz, sv.L, nd, 16 if prec == 64 else 8))
        mv[f"fp{prec}"] = {"matvec_per_s": 1.0 / tm, "ms": tm * 1e3, "alg_bytes": byts,
                           "GB_per_s": byts / tm / 1e9, "frac_hbm_peak": byts / tm / 1e9 / HBM_PEAK_GBS}

    # dominant kernel of the timed region (FP32 passes dominate mixed solves)
    names = ["P1 scatter+FFT-x", "P2 FFT-y", "P3 FFT-z.G.IFFT-z", "P4 IFFT-y", "P5 IFFT-x+gather"]
    kern = ["ifdda::k_xpass<float2, 3, 1, 1>", "ifdda::k_strided<float2, 3, 1>",
            "ifdda::k_zconv<float2", "ifdda::k_strided<float2, 3, -1>",
            "ifdda::k_xpass<float2, 3, -1, 2>"]
    b32 = pass_bytes(s.nx, s.ny, s.nz, sv.L, nd, 8)
    b64 = pass_bytes(s.nx, s.ny, s.nz, sv.L, nd, 16)
    passes = {}
    for pi, tag in ((0, "fp64"), (1, "fp32")):
        bb = b64 if pi == 0 else b32
        for p in range(5):
            if cnt[pi, p]:
                avg = ms_sum[pi, p] / cnt[pi, p]
                passes[f"{tag} {names[p]}"] = {"avg_us": avg * 1e3, "launches": int(cnt[pi, p]),
                                               "total_ms": float(ms_sum[pi, p]),
                                               "GB_per_s": bb[p] / (avg * 1e-3) / 1e9}
    p_dom = int(np.argmax(ms_sum[1]))
    avg_dom = ms_sum[1, p_dom] / max(1, cnt[1, p_dom])
    achieved = b32[p_dom] / (avg_dom * 1e-3) / 1e9

    # HBM bytes per launch of that kernel from the committed rocprofv3 counter
    # summary (FETCH_SIZE x2.0 and WRITE_SIZE x1.0, calibrated on a plain copy)
    traffic, traffic_src = None, None
    tj = os.path.join(ROOT, "profiles", "matvec_config2_hbm_traffic.json")
    if n == 100 and os.path.exists(tj):
        kj = json.load(open(tj))["kernels"]
        for name, v in kj.items():
            if name.startswith(kern[p_dom]):
                traffic = v["fetch_bytes"] + v["write_bytes"]
                traffic_src = f"profiles/matvec_config2_hbm_traffic.json ({name})"

    out = {
        "metric": METRIC,
        "value": args.steps / elapsed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 matvec + c128 refinement",
        "data": "synthetic",
        "config": {"workload": f"{n}^3 inhomogeneous cube (eps 2.25+0.01i, std 0.1, corr 100 nm), "
                               f"d=20 nm, lambda 632.8 nm, circular wave theta=30, FFT grid "
                               f"{sv.L[0]}x{sv.L[1]}x{sv.L[2]}, GPBICG x0=0 tol 1e-4",
                   "dipoles": nd, "method": "GPBICG1", "tol": 1e-4},
        "mixed": {"solves_per_s": args.steps / elapsed, "iterations": info32["nloop"],
                  "matvecs": info32["ncompte"], "outer": info32["outer"], "tol1": info32["tol1"],
                  "field_rel_err_vs_fp64": field_err,
                  "latency_ms": {"p50": float(np.percentile(lat, 50)) * 1e3,
                                 "p99": float(np.percentile(lat, 99)) * 1e3}},
        "fp64": {"solves_per_s": 1.0 / float(np.median(lat64)), "iterations": info64["nloop"],
                 "matvecs": info64["ncompte"], "tol1": info64["tol1"],
                 "latency_ms": {"p50": float(np.percentile(lat64, 50)) * 1e3,
                                "p99": float(np.percentile(lat64, 99)) * 1e3}},
        "mixed_speedup_vs_fp64": (args.steps / elapsed) * float(np.median(lat64)),
        "matvec": mv,
        "green_setup_s": green_s,
        "passes": passes,
        "roofline": {"bound": "hbm", "kernel": f"{names[p_dom]} ({kern[p_dom]})",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "alg_bytes_per_launch": b32[p_dom], "avg_us": avg_dom * 1e3},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sv, s.nx, s.ny, s.nz, s.tabdip, nd, info64["ncompte"])
    sv.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
