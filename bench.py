#!/usr/bin/env python3
"""BP5 benchmark on one MI355X.

Workload (24^3 hex box at p = 7, [-1,1]^3 with interior vertices perturbed
by U(-0.1h, 0.1h), lambda = 1 as in BP/BP5/bp5Main.cpp:47, seeded N(0,1)
right-hand side at the 167^3 = 4,657,463 masked DOFs):

  step = one solve to ||r|| / ||b|| < 1e-10 with the FP32 inner CG inside
         the FP64 defect-correction loop.

value = Ndofs * N_it(FP64 CG to 1e-10) / T_step, the FP64-equivalent CG
rate in GDOF*it/s; the plain FP64 CG (same tolerance) is timed beside it.
The roofline object covers the dominant kernel, the FP64 BP5 Ax, timed
with HIP events on the stream it is launched on; its algorithmic bytes are
the reference model (BK/BK5/src/bk5Run.cpp:116-122). The CPU baseline is the
oracle (C++/OpenMP restatement of the reference kernels) running a bounded
number of FP64 CG iterations of the same problem on the host cores.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "BP5 GDOF·it/s (24³ hex, p=7): FP64 vs FP32+refinement at 1e-10 rel resid"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def physical_cores():
    """lscpu Core(s) per socket x Socket(s) (libs/core/platformDeviceConfig.cpp:106-126)."""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, check=True).stdout
        kv = {}
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        return int(kv["Core(s) per socket"]) * int(kv["Socket(s)"])
    except Exception:
        return os.cpu_count() or 1


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_bp5_ax_fp64_p7_24cube_hbm_traffic.txt")


def measured_traffic():
    """HBM bytes per Ax launch from the committed counter summary (FETCH_SIZE and
    WRITE_SIZE scaled by a plain copy of known size), or None."""
    try:
        for line in open(TRAFFIC_FILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def bk5_bytes(Ndofs, Nel, Np):
    # BK/BK5/src/bk5Run.cpp:116-122 (non-affine hex, FP64, 4-byte dlong)
    return Ndofs * 8 + Nel * (Np * 7 * 8 + 4 + Np * 4 + Np * 8)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--N", type=int, default=7)
    ap.add_argument("--nx", type=int, default=24)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--interval", type=int, default=32,
                    help="FP32 iterations between FP64 defect corrections")
    ap.add_argument("--ax-launches", type=int, default=50)
    ap.add_argument("--cpu-its", type=int, default=1000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ax-only", action="store_true",
                    help="only the Ax launches and a 512 MiB device copy (for counter runs)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path (the BP5 solve is "
              "sized for one MI355X); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from benchparanumal_amd import capi

    torch.cuda.set_device(0)
    plat = capi.Platform(0)
    t0 = time.perf_counter()
    mesh = capi.Mesh(N=args.N, nx=args.nx, perturb=0.1, seed=1)
    bp = capi.BP(plat, mesh, 5, lam=1.0)
    Ndofs, Nel, Np = bp.Ngather, mesh.Nelements, mesh.Np
    log(f"setup: {time.perf_counter() - t0:.1f}s  Ndofs={Ndofs}  Nel={Nel}  p={args.N}")

    if args.ax_only:
        # counter calibration: one device-to-device copy of exactly 512 MiB
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        q = torch.from_numpy(capi.random_normal(2, Ndofs)).cuda()
        AqL = torch.empty(bp.NLocal, dtype=torch.float64, device="cuda")
        Aq = torch.empty_like(q)
        for _ in range(args.ax_launches):
            bp.ax_local(q, AqL)
        for _ in range(args.ax_launches):
            bp.operator(q, Aq)
        torch.cuda.synchronize()
        log(f"ax-only: {args.ax_launches} BK5 Ax launches, {args.ax_launches} operator (Ax + gather) "
            f"launches and one 512 MiB copy done")
        return

    b_host = capi.random_normal(7, Ndofs)
    b = torch.from_numpy(b_host).cuda()
    x = torch.zeros_like(b)
    r = torch.empty_like(b)

    # ---- FP64 CG to tol (reference iteration count and time)
    def fp64_solve():
        x.zero_()
        r.copy_(b)
        its, _ = bp.solve(x, r, tol=args.tol, maxit=100000, history=False)
        return its

    def mixed_solve():
        x.zero_()
        return bp.solve_mixed(x, b, tol=args.tol, interval=args.interval, maxit=100000)

    def timed(fn, k):
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = None
        for _ in range(k):
            out = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / k, out

    for _ in range(args.warmup):
        fp64_solve()
    t64, its64 = timed(fp64_solve, args.steps)
    Ax = torch.empty_like(b)
    bp.operator(x, Ax)
    rel64 = float(torch.linalg.norm(b - Ax) / torch.linalg.norm(b))
    log(f"fp64 CG: {its64} its to {args.tol:g}, {t64 * 1e3:.1f} ms/solve, true rel res {rel64:.2e}")

    for _ in range(args.warmup):
        mixed_solve()
    tmix, st = timed(mixed_solve, args.steps)
    bp.operator(x, Ax)
    relmix = float(torch.linalg.norm(b - Ax) / torch.linalg.norm(b))
    log(f"mixed: {st['iterations']} FP32 its, {st['updates']} FP64 corrections, {tmix * 1e3:.1f} ms/solve, "
        f"true rel res {relmix:.2e}")

    # ---- Ax kernel roofline (HIP events on the launch stream = torch's current stream)
    q = torch.from_numpy(capi.random_normal(2, Ndofs)).cuda()
    AqL = torch.empty(bp.NLocal, dtype=torch.float64, device="cuda")
    for _ in range(5):
        bp.ax_local(q, AqL)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(args.ax_launches):
        bp.ax_local(q, AqL)
    ev1.record()
    torch.cuda.synchronize()
    t_ax = ev0.elapsed_time(ev1) / 1e3 / args.ax_launches
    nbytes = bk5_bytes(Ndofs, Nel, Np)
    achieved = nbytes / t_ax / 1e9
    log(f"BK5 Ax: {t_ax * 1e6:.1f} us, {achieved:.0f} GB/s, {Ndofs / t_ax / 1e9:.1f} GDOF/s")

    # plain device copy of 512 MiB on the same stream: the attainable HBM rate
    src = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    ev0.record()
    for _ in range(10):
        dst.copy_(src)
    ev1.record()
    torch.cuda.synchronize()
    copy_gbs = 2 * src.numel() * 8 * 10 / (ev0.elapsed_time(ev1) / 1e3) / 1e9
    del src, dst
    log(f"512 MiB device copy: {copy_gbs:.0f} GB/s (read + write)")

    # ---- CPU baseline: the oracle's FP64 CG on the host cores, bounded sample
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib

        cores = min(physical_cores(), int(os.environ.get("OMP_NUM_THREADS", "1000")))
        threads = oracle_lib.set_threads(cores)
        P = oracle_lib.Problem(bp=5, N=args.N, nx=args.nx, lam=1.0, perturb=0.1, seed=1)
        assert P.Ngather == Ndofs
        tc = time.perf_counter()
        P.cg(b_host, args.cpu_its)
        tcpu = time.perf_counter() - tc
        cpu = {"value": Ndofs * args.cpu_its / tcpu / 1e9, "unit": "GDOF·it/s", "cores": threads,
               "kind": "port",
               "sample": f"oracle FP64 CG (restated OKL loop nests, OpenMP), same 24^3 p=7 BP5 problem, "
                         f"{args.cpu_its} iterations in {tcpu:.1f}s"}
        log(f"cpu baseline: {cpu['value']:.4f} GDOF*it/s on {threads} threads")

    value = Ndofs * its64 / tmix / 1e9
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "GDOF·it/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": tmix * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 outer + f32 inner",
        "data": "synthetic",
        "config": {
            "workload": f"BP5 hex p={args.N} {args.nx}^3 perturbed box, lambda=1, N(0,1) rhs, "
                        f"solve to {args.tol:g} rel resid (FP32 CG + FP64 defect correction every "
                        f"{args.interval} its)",
            "N": args.N, "nx": args.nx, "Ndofs": Ndofs, "tol": args.tol, "interval": args.interval,
        },
        "fp64_cg": {"value": Ndofs * its64 / t64 / 1e9, "unit": "GDOF·it/s", "iterations": its64,
                    "ms_per_solve": t64 * 1e3, "true_rel_res": rel64},
        "mixed": {"fp32_iterations": st["iterations"], "fp64_corrections": st["updates"],
                  "ms_per_solve": tmix * 1e3, "true_rel_res": relmix,
                  "own_rate": Ndofs * st["iterations"] / tmix / 1e9,
                  "speedup_vs_fp64": t64 / tmix},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "kernel": "k_bp5_ax (FP64, non-affine, Nq=8)", "us_per_launch": t_ax * 1e6,
                     "algorithmic_bytes": nbytes, "gdof_per_s": Ndofs / t_ax / 1e9,
                     "device_copy_gbs": copy_gbs},
        "tuning": bp.tuning(),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
