#!/usr/bin/env python
"""Benchmark of the curl-SVD + bi-anisotropic material matvec path on one MI355X.

Workload (the paper-size plain configuration): simple-cubic woodpile, 4 rod
layers, eps = 13 I, mu = I, xi = 0.2 i I, zeta = -0.2 i I in the rods, vacuum
outside, n = 60 Yee grid (4N = 864 000 complex128 unknowns per NFSEP vector),
40 k-points on Gamma-X-M-R-Gamma, FP64.

One step = one fncB (curl-SVD + material sandwich, 5 HIP kernels) applied to
all 40 k-points of the band path in one batched launch sequence, inputs
resident in HBM.  ``value`` = matvecs/s over the whole job (40 per step).
Also reported: the plain per-k rate (one k-point per launch sequence), the
batched/plain ratio, seconds per k-point of the full Lanczos solve (plain and
batched), per-kernel HIP-event timings, the roofline of the dominant kernel,
and the CPU oracle's rate on the host as a baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "curl+material matvecs/s and s per k-point at n=60 (batched vs plain), %HBM peak"
HBM_PEAK_GBS = 8000.0
# algorithmic HBM bytes per grid point and batch item for each fncB kernel (complex128 = 16 B):
#   zpass_mix_in  read 4 + write 6 words; xypass_* read 6 + write 6; material read 6 + write 6 (+1 B id);
#   zpass_mix_out read 6 + write 4
STAGE_BYTES_PER_POINT = {"zpass_mix_in": 160.0, "xypass_inv": 192.0, "material": 193.0, "xypass_fwd": 192.0,
                         "zpass_mix_out": 160.0}
FNCB_ALGO_BYTES_PER_POINT = 512.0  # one fncB: 32 complex words per grid point (3-stage fused minimum)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(n, seconds):
    """Oracle fncB (scipy.fft with all host threads + CSR SpMV, as the reference's
    flag_gpu=0 path) at the same n on one k-point, timed for about ``seconds``."""
    from oracle import fame_oracle as fo
    import __graft_entry__ as ge

    S = ge.load_package().structures
    dims = (n, n, n)
    theta = S.theta_of_k(np.pi * np.array([0.5, 0.3, 0.1]), S.SC_LATTICE)
    table = S.bi_isotropic_table(13.0, 0.2, S.SC_LATTICE)
    ids = S.woodpile_sc_ids(dims)
    T = table[ids]
    P = fo.Problem(dims, theta, T[:, 0], T[:, 1], T[:, 2])
    rng = np.random.default_rng(1234)
    x = rng.standard_normal(4 * P.N) + 1j * rng.standard_normal(4 * P.N)
    P.fncB(x)
    cnt, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        x = P.fncB(x)
        x /= np.linalg.norm(x)
        cnt += 1
    dt = time.perf_counter() - t0
    return {"value": cnt / dt, "unit": "matvecs/s", "cores": os.cpu_count(), "kind": "port",
            "sample": f"oracle fncB (CPU restatement of the reference algorithm, not MATLAB) at n={n}, one k-point, "
                      f"{cnt} calls in {dt:.1f} s; scipy.fft workers={os.cpu_count()}, CSR SpMV single-threaded"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=60)
    ap.add_argument("--nk", type=int, default=40, help="k-points of the band path (batch size)")
    ap.add_argument("--solve", choices=["none", "plain", "both"], default="both",
                    help="also time full Lanczos solves (s per k-point)")
    ap.add_argument("--solve-k", type=int, default=8,
                    help="k-points (spread over the path) solved one by one for the plain s/k-point")
    ap.add_argument("--nev", type=int, default=10)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--chunk", type=int, default=-1,
                    help="k-points per fncB chunk (-1: autotune on the plan, 0: whole batch)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; --gpus must be 1 "
                                   "(k-points shard across GPUs with no collectives, see DESIGN.md)"}))
        sys.exit(2)

    import torch
    import __graft_entry__ as ge

    pkg = ge.load_package()
    S = pkg.structures
    n, nk = args.n, args.nk
    dims = (n, n, n)
    N = n ** 3
    ks = S.sc_band_path()
    idx = np.linspace(0, len(ks) - 1, nk).round().astype(int) if nk != len(ks) else np.arange(len(ks))
    thetas = np.stack([S.theta_of_k(ks[i], S.SC_LATTICE) for i in idx])
    table = S.bi_isotropic_table(13.0, 0.2, S.SC_LATTICE)
    ids = S.woodpile_sc_ids(dims)
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)

    log(f"bench: n={n} nk={nk} building plans")
    pb = pkg.Plan(dims, thetas, table, ids)
    p1 = pkg.Plan(dims, thetas[:1], table, ids)
    rng = np.random.default_rng(1234)
    xb = torch.as_tensor(rng.standard_normal((nk, 4 * N)) + 1j * rng.standard_normal((nk, 4 * N))).to(dev)
    yb = torch.empty_like(xb)
    lib = pb.lib
    import ctypes

    def run_batched():
        rc = lib.fame_fncB(pb.h, ctypes.c_void_p(xb.data_ptr()), ctypes.c_void_p(yb.data_ptr()))
        if rc != 0:
            raise RuntimeError(lib.fame_last_error().decode())

    if args.chunk < 0:
        chunk, tune_ms = pb.autotune()
        log(f"bench: autotuned chunk = {chunk}, xy/z launch variants = {pb.variant()} "
            f"({tune_ms:.3f} ms per batched fncB)")
    else:
        chunk = args.chunk
        pb.set_chunk(chunk)
    x1, y1 = xb[0], yb[0]

    def run_plain():
        rc = lib.fame_fncB(p1.h, ctypes.c_void_p(x1.data_ptr()), ctypes.c_void_p(y1.data_ptr()))
        if rc != 0:
            raise RuntimeError(lib.fame_last_error().decode())

    # ---- batched (the metric) ----
    for _ in range(args.warmup):
        run_batched()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        run_batched()
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = e0.elapsed_time(e1) / args.steps
    value = nk * args.steps / (ms_step * args.steps / 1e3)
    log(f"bench: batched fncB {ms_step:.3f} ms/step ({nk} k) -> {value:.0f} matvecs/s (wall {wall:.3f} s)")

    # ---- plain per-k ----
    for _ in range(args.warmup):
        run_plain()
    torch.cuda.synchronize()
    reps = args.steps * nk
    e0.record()
    for _ in range(reps):
        run_plain()
    e1.record()
    torch.cuda.synchronize()
    ms_plain = e0.elapsed_time(e1) / reps
    plain_rate = 1e3 / ms_plain
    log(f"bench: plain fncB {ms_plain:.4f} ms/matvec -> {plain_rate:.0f} matvecs/s")

    # ---- per-kernel timings (HIP events between the five kernels) ----
    pb.profile(True)
    for _ in range(max(3, args.steps // 2)):
        run_batched()
    st, calls = pb.stage_times()
    pb.profile(False)
    stage_avg = {k: v / calls for k, v in st.items()}
    dom = max(stage_avg, key=stage_avg.get)
    dom_bytes = STAGE_BYTES_PER_POINT[dom] * N * nk
    achieved = dom_bytes / (stage_avg[dom] * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{dom}_n{n}.txt")
    if os.path.exists(tfile):
        for line in open(tfile):
            if line.startswith("traffic_bytes_per_launch:"):
                traffic = float(line.split(":", 1)[1])
    roofline = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "algo_bytes_per_launch": dom_bytes}
    log("bench: stage ms " + ", ".join(f"{k}={v:.4f}" for k, v in stage_avg.items()))
    fncb_gbs = FNCB_ALGO_BYTES_PER_POINT * N * value / 1e9

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "matvecs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"woodpile SC n={n} gamma=0.2, {nk}-k Gamma-X-M-R-Gamma band path, fncB batched over k",
                   "n": n, "nk": nk, "vector_len": 4 * N, "chunk": chunk,
                   "fft_variants": list(pb.variant())},
        "plain_matvecs_per_s": plain_rate,
        "batched_over_plain_matvec": value / plain_rate,
        "fncB_algo_GBs": fncb_gbs,
        "fncB_hbm_frac": fncb_gbs / HBM_PEAK_GBS,
        "stage_ms": stage_avg,
        "roofline": roofline,
    }

    # ---- s per k-point: full Lanczos solves (table setup + eigensolve, as the reference's time_eig window) ----
    if args.solve != "none":
        k = args.nev
        sample = np.unique(np.linspace(0, nk - 1, min(args.solve_k, nk)).round().astype(int))
        times, steps = [], []
        for j in sample:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            pj = pkg.Plan(dims, thetas[j], table, ids)
            ev, ew, it = pj.Lanczos(k, tol=1e-10, maxiter=10000, p=2 * k, cg_tol=1e-12, cg_maxit=10000)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
            steps.append(int(it[0]))
            pj.close()
            del ev
            log(f"bench: plain solve k#{j}: {times[-1]:.2f} s, {steps[-1]} Lanczos steps, "
                f"omega[:3]={np.sort(1.0 / ew[0])[:3]}")
        out["s_per_kpoint_plain"] = float(np.mean(times))
        out["plain_solve_sample"] = {"k_indices": sample.tolist(), "seconds": times, "lanczos_steps": steps,
                                     "note": "mean over this sample of the band path, one k-point per plan"}
        if args.solve == "both":
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            pl = pkg.Plan(dims, thetas, table, ids)
            pl.set_chunk(chunk)
            ev, ew, it = pl.Lanczos(k, tol=1e-10, maxiter=10000, p=2 * k, cg_tol=1e-12, cg_maxit=10000)
            torch.cuda.synchronize()
            tb = time.perf_counter() - t0
            pl.close()
            del ev
            out["s_per_kpoint_batched"] = tb / nk
            out["batched_solve_seconds"] = tb
            out["lanczos_steps_batched"] = it.tolist()
            out["batched_over_plain_solve"] = out["s_per_kpoint_plain"] / out["s_per_kpoint_batched"]
            log(f"bench: batched solve {nk} k: {tb:.2f} s -> {tb / nk:.3f} s/k "
                f"(x{out['batched_over_plain_solve']:.2f} vs plain)")

    if not args.no_cpu:
        log("bench: CPU baseline")
        out["cpu_baseline"] = cpu_baseline(n, args.cpu_seconds)
    pb.close()
    p1.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
