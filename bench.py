#!/usr/bin/env python3
"""BP5 benchmark on one MI355X.

Workload (24^3 hex box at p = 7, [-1,1]^3 with interior vertices perturbed
by U(-0.1h, 0.1h), lambda = 1 as in BP/BP5/bp5Main.cpp:47, seeded N(0,1)
right-hand side at the 167^3 = 4,657,463 masked DOFs):

  step = one solve to ||r|| / ||b|| < 1e-10 with the FP32 inner CG inside
         the FP64 defect-correction loop.

value = Ndofs * N_it(FP64 CG to 1e-10) / T_step, the FP64-equivalent CG
rate in GDOF*it/s; the plain FP64 CG (same tolerance) is timed beside it.
The roofline object covers the dominant kernel of the timed solve, the Ax
of the FP32 CG operator (the fused form when the setup's timing picked it),
timed inside one more solve identical to the timed ones, with HIP events
around each of its applications on the stream it is launched on (the same
kernel launched back to back is reported beside it); its algorithmic bytes
are the reference model (BK/BK5/src/bk5Run.cpp:116-122) with 4-byte values. The FP64 BK5 Ax (the reference's Ax benchmark) is
reported beside it with the FP64 model. The CPU baseline is the
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
TRAFFIC_FILE_F32 = os.path.join(ROOT, "profiles", "k_bp5_ax_fp32_cg_operator_p7_24cube_hbm_traffic.txt")


LIMITER_FILE_F32 = os.path.join(ROOT, "profiles", "k_bp5_ax_fp32_cg_operator_p7_24cube_sq_counters.txt")


def kernel_limiter(path):
    """What bounds the kernel, from the committed SQ counter summary
    (lines 'limiter_<key>: <value>'), or None."""
    out = {}
    try:
        for line in open(path):
            if line.startswith("limiter_"):
                k, v = line[len("limiter_"):].split(":", 1)
                v = v.strip()
                try:
                    out[k.strip()] = float(v)
                except ValueError:
                    out[k.strip()] = v
    except OSError:
        return None
    if out:
        out["source"] = "committed counter summary " + os.path.relpath(path, ROOT)
    return out or None


def measured_traffic(path=TRAFFIC_FILE):
    """HBM bytes per Ax launch from the committed counter summary (FETCH_SIZE and
    WRITE_SIZE scaled by a plain copy of known size), or None."""
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def bk5_bytes(Ndofs, Nel, Np, value_bytes=8):
    # BK/BK5/src/bk5Run.cpp:116-122 (non-affine hex, 4-byte dlong); value_bytes
    # = 4 for the FP32 operator (values float, indices unchanged)
    w = value_bytes
    return Ndofs * w + Nel * (Np * 7 * w + 4 + Np * 4 + Np * w)


def bk5_tri_bytes(Ndofs, Nel, Np):
    # the FP32 operator, which forms its geometric factors from each element's
    # trilinear map: bk5_bytes with 4-byte values and the 7 factors per node
    # replaced by the map's 24 coefficients per element
    return Ndofs * 4 + Nel * (24 * 4 + 4 + Np * 4 + Np * 4)


def bp5_ax_flops(Nq, Nel, affine=False, nfields=1):
    # BP/BP5/src/bp5Run.cpp:159-180 (hex): (12 Nq^4 + 18 Nq^3) per element,
    # (12 Nq^4 + 21 Nq^3) affine
    return (12 * Nq**4 + (21 if affine else 18) * Nq**3) * nfields * Nel


def gather_bytes(Ndofs, NLocal):
    # BP/BP5/src/bp5Run.cpp:136-139
    return (Ndofs + 1) * 4 + NLocal * 4 + NLocal * 8 + Ndofs * 8


def cg_solve_model(Ndofs, NLocal, ax_bytes, ax_flops, iterations):
    """Bytes and flops of a CG solve of `iterations` iterations, the
    reference's model (BP/BP5/src/bp5Run.cpp:141-142, 185-188): a first
    iteration of 4 vector passes and 5 Ndofs flops, then per iteration 11
    vector passes, the Ax and the gather."""
    gb = gather_bytes(Ndofs, NLocal)
    nbytes = (4 * Ndofs * 8 + ax_bytes + gb) + (11 * Ndofs * 8 + ax_bytes + gb) * iterations
    nflops = (5 * Ndofs + ax_flops + NLocal) + (11 * Ndofs + ax_flops + NLocal) * iterations
    return nbytes, nflops


def krhs_bytes(K, Ndofs, Nel, Np):
    # k right-hand sides: vectors K times, geometry and GlobalToLocal once
    return K * (Ndofs * 8 + Nel * Np * 8) + Nel * (Np * 7 * 8 + 4 + Np * 4)


def event_time(fn, reps, warm=2):
    """Average seconds per call of fn, HIP events on torch's current stream
    (the stream libbpx launches on)."""
    import torch

    for _ in range(warm):
        fn()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / 1e3 / reps


def bench_krhs(capi, plat):
    """k-RHS BP5 Ax at 20^3, p=7 on the problem's arrays: each form (chunk 0 =
    the default, the pencil form here; -1 the sheet workgroup form; 1, 2, 4
    the sweep form) against the single-vector Ax, both timed by HIP events
    their own dispatches stamp (kernel time, 10 launches after 2 warm-ups),
    on random right-hand sides."""
    import numpy as np

    m = capi.Mesh(N=7, nx=20, perturb=0.1, seed=1)
    bp = capi.BP(plat, m, 5, lam=1.0)
    Ndofs, Nel, Np = bp.Ngather, m.Nelements, m.Np
    out = {"workload": "BP5 Ax, 20^3 hex p=7 perturbed, K right-hand sides stored one after another",
           "Ndofs": Ndofs, "timing": "dispatch-stamped HIP events, mean of 10 launches"}
    import torch

    rows = []
    singles = []
    for K in (4, 8, 16, 32):
        q = torch.from_numpy(capi.random_normal(4, Ndofs * K)).cuda()
        per_chunk = {}
        for chunk in (0, -1, 1, 2, 4):
            if chunk > 0 and K % chunk:
                continue
            tk, t1 = bp.time_krhs(K, chunk, 10, q)
            per_chunk[chunk] = tk
            singles.append(t1)
        chunk = min(per_chunk, key=per_chunk.get)
        t = per_chunk[chunk] * 1e-6
        t1 = float(np.median(singles)) * 1e-6
        nb = krhs_bytes(K, Ndofs, Nel, Np)
        rows.append({"K": K, "chunk": chunk, "us": t * 1e6, "gdof_per_s": K * Ndofs / t / 1e9,
                     "gbs": nb / t / 1e9, "algorithmic_bytes": nb, "speedup_vs_K_single": K * t1 / t,
                     "us_by_chunk": per_chunk})
        del q
    out["single_us"] = float(np.median(singles))
    out["rows"] = rows
    return out


def bench_batch(capi, plat):
    """64 ragged BP5 problems (nx, ny, nz in [2, 8], p in 3..8) to 1e-8: one
    batched solve against the same problems solved one after another."""
    import numpy as np
    import time
    import torch

    rng = np.random.default_rng(2024)
    probs = []
    for i in range(64):
        nx, ny, nz = (int(v) for v in rng.integers(2, 9, size=3))
        probs.append({"N": int(rng.integers(3, 9)), "nx": nx, "ny": ny, "nz": nz, "perturb": 0.1, "seed": 100 + i})
    B = capi.Batch(plat, probs, lam=1.0)
    rhs = torch.from_numpy(capi.random_normal(77, B.Ntot)).cuda()
    x = torch.zeros(B.Ntot, dtype=torch.float64, device="cuda")
    B.solve(x, rhs, tol=1e-8, maxit=5000)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    its, _ = B.solve(x, rhs, tol=1e-8, maxit=5000)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    sizes = np.diff(B.offsets)
    work = float((sizes * its).sum())
    # the same problems one by one with the single-problem CG
    bps = []
    for pr in probs:
        m = capi.Mesh(N=pr["N"], nx=pr["nx"], ny=pr["ny"], nz=pr["nz"], perturb=0.1, seed=pr["seed"])
        bps.append((m, capi.BP(plat, m, 5, lam=1.0)))
    def seq():
        for i, (m, bp) in enumerate(bps):
            a, b = int(B.offsets[i]), int(B.offsets[i + 1])
            xs = torch.zeros(b - a, dtype=torch.float64, device="cuda")
            r = rhs[a:b].clone()
            bp.solve(xs, r, tol=1e-8, maxit=5000, history=False)
    seq()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    seq()
    torch.cuda.synchronize()
    ts = time.perf_counter() - t0
    # FP64 batch and FP32+refinement batch at 1e-10 (interval 32), and the
    # single-problem mixed solve one problem after another
    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, out
    t64, (its64, _) = timed(lambda: B.solve(x, rhs, tol=1e-10, maxit=20000))
    tmx, (itsm, upsm, relm) = timed(lambda: B.solve_mixed(x, rhs, tol=1e-10, interval=32, maxit=20000))
    def seq_mixed():
        for i, (m, bp) in enumerate(bps):
            a, b = int(B.offsets[i]), int(B.offsets[i + 1])
            xs = torch.zeros(b - a, dtype=torch.float64, device="cuda")
            bp.solve_mixed(xs, rhs[a:b].contiguous(), tol=1e-10, interval=32)
    tsm, _ = timed(seq_mixed)
    mixed = {"workload": "the same 64 problems to 1e-10 true relative residual",
             "fp64_batch_ms": t64 * 1e3, "fp64_iterations_max": int(its64.max()),
             "ms_per_solve": tmx * 1e3, "fp32_iterations_min": int(itsm.min()),
             "fp32_iterations_max": int(itsm.max()), "corrections_max": int(upsm.max()),
             "true_rel_res_max": float(relm.max()), "speedup_vs_fp64_batch": t64 / tmx,
             "sequential_mixed_ms": tsm * 1e3, "speedup_vs_sequential_mixed": tsm / tmx}
    return {"workload": "64 BP5 problems, nx/ny/nz in [2,8], p in 3..8, N(0,1) rhs, CG to 1e-8",
            "Ntot": B.Ntot, "iterations_min": int(its.min()), "iterations_max": int(its.max()),
            "ms_per_solve": tb * 1e3, "gdof_it_per_s": work / tb / 1e9,
            "sequential_ms": ts * 1e3, "speedup_vs_sequential": ts / tb, "tuning": B.tuning(),
            "graph_replays": B.info()["replays"], "fp32_refinement": mixed}


def bench_p15(capi, plat):
    """BK5 Ax at 24^3, p = 15 (Np 4096, ggeo 2.7 GB), and 100 iterations of
    the FP64 BP5 CG on the same mesh (N(0,1) right-hand side)."""
    import torch

    m = capi.Mesh(N=15, nx=24, perturb=0.1, seed=1)
    bp = capi.BP(plat, m, 5, lam=1.0)
    q = torch.from_numpy(capi.random_normal(2, bp.Ngather)).cuda()
    AqL = torch.empty(bp.NLocal, dtype=torch.float64, device="cuda")
    t = event_time(lambda: bp.ax_local(q, AqL), 20)
    nb = bk5_bytes(bp.Ngather, m.Nelements, m.Np)
    del AqL
    b = torch.from_numpy(capi.random_normal(7, bp.Ngather)).cuda()
    x = torch.zeros_like(b)
    r = torch.empty_like(b)

    def cg100():
        x.zero_()
        r.copy_(b)
        return bp.solve(x, r, tol=0.0, maxit=100, history=False)[0]

    t_cg = event_time(cg100, 1)
    return {"workload": "BK5 Ax, 24^3 hex p=15 perturbed", "Ndofs": bp.Ngather, "us": t * 1e6,
            "gdof_per_s": bp.Ngather / t / 1e9, "algorithmic_bytes": nb, "gbs": nb / t / 1e9,
            "cg100_ms": t_cg * 1e3, "cg_gdof_it_per_s": bp.Ngather * 100 / t_cg / 1e9,
            "cg_operator": "fused" if bp.tuning()["fused"] == 1 else "unfused"}


def bench_other_bps(capi, plat):
    """BP1, BP2, BP4 and BP6 Ax at 24^3, p = 7 (the perturbed box; three
    fields for BP2/4/6), against each one's own byte model: BP1/BP2
    BP/BP1/src/bp1Run.cpp:107-120 (cubature weights), BP4
    BP/BP4/src/bp4Run.cpp:106-119 (cubature factors), BP6
    BP/BP6/src/bp6Run.cpp:106-134 (node factors). Each is timed over 20
    launches after 200 untimed ones (8-48 ms): launches in the first ~20 ms
    after a problem's setup run up to 15% slower (BP4 268 against 235 us)."""
    import torch

    rows = []
    for bpn in (1, 2, 4, 6):
        m = capi.Mesh(N=7, nx=24, cubature=bpn <= 4, perturb=0.1, seed=1)
        bp = capi.BP(plat, m, bpn, lam=1.0)
        Nd, Nel, Np, nf = bp.Ngather, m.Nelements, m.Np, bp.Nfields
        q = torch.from_numpy(capi.random_normal(2, Nd)).cuda()
        AqL = torch.empty(bp.NLocal, dtype=torch.float64, device="cuda")
        t = event_time(lambda: bp.ax_local(q, AqL), 20, warm=200)
        geo = m.cubNp * 8 if bpn <= 2 else (m.cubNp if bpn == 4 else Np) * 7 * 8
        nb = Nd * 8 + Nel * (geo + 4 + Np * nf * (4 + 8))
        rows.append({"bp": bpn, "fields": nf, "Ndofs": Nd, "us": t * 1e6, "gdof_per_s": Nd / t / 1e9,
                     "algorithmic_bytes": nb, "gbs": nb / t / 1e9, "frac": nb / t / 1e9 / HBM_PEAK_GBS})
        del bp, m, q, AqL
        torch.cuda.empty_cache()
    return {"workload": "BP1/BP2/BP4/BP6 Ax, 24^3 hex p=7 perturbed, FP64", "rows": rows}


def bench_families(capi, plat):
    """BP5 Ax (FP64) on the quadrilateral, triangle and tetrahedron boxes at
    p = 7 (perturbed vertices), against the reference's byte model
    (BP/BP5/src/bp5Run.cpp:106-133: quads 4 factors per node, simplices 4 or
    7 per element as its affine meshes). The simplex operator is a GEMM with
    the dense reference matrices, so its rate is also given in FP64 TFLOP/s
    (2 (NG+1) Np^2 flops per element, NG = 3 or 6) against 78.6."""
    import torch

    rows = []
    for el, name, nx, ny, nz in ((capi.QUAD, "quadrilateral", 256, 256, 1), (capi.TRI, "triangle", 192, 192, 1),
                                 (capi.TET, "tetrahedron", 20, 20, 20)):
        m = capi.Mesh(N=7, nx=nx, ny=ny, nz=nz, element=el, perturb=0.1, seed=1)
        bp = capi.BP(plat, m, 5, lam=1.0)
        Nd, Nel, Np = bp.Ngather, m.Nelements, m.Np
        q = torch.from_numpy(capi.random_normal(2, Nd)).cuda()
        AqL = torch.empty(bp.NLocal, dtype=torch.float64, device="cuda")
        t = event_time(lambda: bp.ax_local(q, AqL), 20, warm=20)
        geo = Np * 4 * 8 if el == capi.QUAD else (7 if el == capi.TET else 4) * 8
        nb = Nd * 8 + Nel * (geo + 4 + Np * 12)
        row = {"element": name, "N": 7, "Nelements": Nel, "Np": Np, "Ndofs": Nd, "us": t * 1e6,
               "gdof_per_s": Nd / t / 1e9, "algorithmic_bytes": nb, "gbs": nb / t / 1e9,
               "frac": nb / t / 1e9 / HBM_PEAK_GBS}
        if el != capi.QUAD:
            ng = 6 if el == capi.TET else 3
            fl = 2.0 * (ng + 1) * Np * Np * Nel
            row.update({"gemm_flops": fl, "tflops": fl / t / 1e12, "tflops_frac_fp64": fl / t / 1e12 / 78.6})
        rows.append(row)
        del bp, m, q, AqL
        torch.cuda.empty_cache()
    return {"workload": "BP5 Ax FP64, p=7: 256^2 quads, 192^2 x 2 triangles, 20^3 x 6 tetrahedra", "rows": rows}


def bench_bp3(capi, plat, args):
    """BP3 (over-integrated stiffness, cubNq = p+2) at 24^3, p = 7: Ax, and
    the FP64 CG against the FP32 CG with FP64 defect correction, both to tol."""
    import time
    import torch

    m = capi.Mesh(N=7, nx=24, cubature=True, perturb=0.1, seed=1)
    bp = capi.BP(plat, m, 3, lam=1.0)
    Nd, Nel, Np, cubNp = bp.Ngather, m.Nelements, m.Np, m.cubNp
    q = torch.from_numpy(capi.random_normal(2, Nd)).cuda()
    AqL = torch.empty(bp.NLocal, dtype=torch.float64, device="cuda")
    t_ax = event_time(lambda: bp.ax_local(q, AqL), 20, warm=200)
    nb = Nd * 8 + Nel * (cubNp * 7 * 8 + 4 + Np * 12)  # BP/BP3/src/bp3Run.cpp:114-118
    b = torch.from_numpy(capi.random_normal(7, Nd)).cuda()
    x = torch.zeros_like(b)
    r = torch.empty_like(b)

    def fp64():
        x.zero_()
        r.copy_(b)
        return bp.solve(x, r, tol=args.tol, maxit=100000, history=False)[0]

    def mixed():
        x.zero_()
        return bp.solve_mixed(x, b, tol=args.tol, interval=args.interval, maxit=100000)

    fp64()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    its = fp64()
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t0
    mixed()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = mixed()
    torch.cuda.synchronize()
    tm = time.perf_counter() - t0
    return {"workload": f"BP3 hex p=7 24^3 perturbed, lambda=1, N(0,1) rhs, tol {args.tol:g}", "Ndofs": Nd,
            "ax_us": t_ax * 1e6, "ax_gbs": nb / t_ax / 1e9, "ax_algorithmic_bytes": nb,
            "fp64_iterations": its, "fp64_ms": t64 * 1e3, "mixed_fp32_iterations": st["iterations"],
            "mixed_corrections": st["updates"], "mixed_ms": tm * 1e3, "speedup_vs_fp64": t64 / tm,
            "fp64_equivalent_gdof_it_per_s": Nd * its / tm / 1e9}


def bench_adjoint(capi, plat):
    """dL/dkappa at 16^3, p=7: forward solve, adjoint solve and gradient to 1e-10."""
    import time
    import numpy as np
    import torch
    from benchparanumal_amd import adjoint

    m = capi.Mesh(N=7, nx=16, perturb=0.1, seed=1)
    bp = capi.BP(plat, m, 5, lam=0.0)
    o = capi.Ogs(mesh=m)
    W = torch.zeros(bp.Ngather, dtype=torch.float64, device="cuda")
    o.gather(plat, torch.from_numpy(np.ascontiguousarray(m.array("wJ"))).cuda(), W)
    kap = torch.from_numpy(np.exp(adjoint.smooth_lattice_field(adjoint.lattice_shape(m), 31).ravel())).cuda()
    u_obs = torch.from_numpy(capi.random_normal(33, bp.Ngather)).cuda()

    def step():
        kt = kap.clone().requires_grad_(True)
        u = adjoint.kappa_solve(bp, kt, W, tol=1e-10)
        L = 0.5 * ((u - u_obs) ** 2 * W).sum()
        L.backward()
        return kt.grad

    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    step()
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    u = torch.randn(bp.Ngather, dtype=torch.float64, device="cuda")
    g = torch.empty(bp.Nnodes, dtype=torch.float64, device="cuda")
    bp.set_kappa(kap)
    tg = event_time(lambda: bp.kappa_gradient(u, u, g), 10)
    return {"workload": "16^3 hex p=7, kappa = exp(smoothed N(0,1), std 0.5), L = 1/2 |u - u_obs|_M^2, tol 1e-10",
            "Ndofs": bp.Ngather, "ms_forward_adjoint_gradient": t * 1e3, "gradient_kernel_us": tg * 1e6}


def reference_protocol(bp, b, Ndofs, NLocal, Nel, Nq):
    """bp5_t::Run's timed protocol (BP/BP5/src/bp5Run.cpp:66-95): a
    1000-iteration warm-up solve with tol = 0, then a timed 100-iteration
    solve from x = 0, reported with the reference's byte and flop models
    (:106-188). The timing is host wall time around the solve with the
    device synchronised on both sides (GlobalPlatformTime)."""
    import torch

    x = torch.zeros_like(b)
    r = b.clone()
    bp.solve(x, r, tol=0.0, maxit=1000, history=False)
    x.zero_()
    r.copy_(b)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    its, _ = bp.solve(x, r, tol=0.0, maxit=100, history=False)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    Np = Nq**3
    nbytes, nflops = cg_solve_model(Ndofs, NLocal, bk5_bytes(Ndofs, Nel, Np), bp5_ax_flops(Nq, Nel), its)
    return {"protocol": "1000-iteration warm-up, then 100 timed iterations from x = 0 (tol 0)",
            "iterations": its, "elapsed_s": t, "time_per_dof_s": t / Ndofs, "gbs": nbytes / t / 1e9,
            "gflops": nflops / t / 1e9, "gdof_it_per_s": Ndofs * its / t / 1e9}


def cpu_baselines(args, b_host, Ndofs):
    """The oracle (the C++/OpenMP restatement of the reference's OKL loop
    nests, OCCA Serial/OpenMP mode: @outer -> omp parallel for) on the
    host: the benchmark's own problem (24^3, p=7) with a bounded number of
    FP64 CG iterations on all physical cores and on one, and the reference's
    sample problem (10^3, p=4) with the reference's protocol, both
    variants."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib

    phys = physical_cores()
    omp = min(phys, int(os.environ.get("OMP_NUM_THREADS", "1000")))
    out = {"physical_cores": phys, "openmp_threads": omp}

    P = oracle_lib.Problem(bp=5, N=args.N, nx=args.nx, lam=1.0, perturb=0.1, seed=1)
    assert P.Ngather == Ndofs
    runs = {}
    for name, threads, its in (("openmp", omp, args.cpu_its), ("serial", 1, args.cpu_serial_its)):
        got = oracle_lib.set_threads(threads)
        tc = time.perf_counter()
        P.cg(b_host, its)
        tcpu = time.perf_counter() - tc
        runs[name] = {"value": Ndofs * its / tcpu / 1e9, "unit": "GDOF·it/s", "threads": got,
                      "iterations": its, "seconds": tcpu}
        log(f"cpu baseline ({name}, {got} threads): {runs[name]['value']:.4f} GDOF*it/s, {its} its in {tcpu:.1f}s")
    del P
    out["problem_24cube_p7"] = runs

    # the reference's sample problem (README.md:109-115): BP5, 10^3, p=4
    rows = []
    for variant, kw in (("reference semantics: [-1,1]^3, lambda 1, forcing rhs", dict(lam=1.0)),
                        ("literal: [0,1]^3, lambda 0, b = M f",
                         dict(lam=0.0, lo=(0.0, 0.0, 0.0), dim=(1.0, 1.0, 1.0)))):
        Pc = oracle_lib.Problem(bp=5, N=4, nx=10, **kw)
        rhs = Pc.rhs_forcing()
        Nd, NL, Nel = Pc.Ngather, Pc.NLocal, Pc.Nelements
        for name, threads in (("openmp", omp), ("serial", 1)):
            got = oracle_lib.set_threads(threads)
            Pc.cg(rhs, 1000)
            tc = time.perf_counter()
            _, its, _, _ = Pc.cg(rhs, 100)
            tcpu = time.perf_counter() - tc
            nbytes, nflops = cg_solve_model(Nd, NL, bk5_bytes(Nd, Nel, 125), bp5_ax_flops(5, Nel), its)
            rows.append({"problem": "BP5 10^3 hex p=4 (" + variant + ")", "Ndofs": Nd, "mode": name,
                         "threads": got, "protocol": "1000-iteration warm-up, 100 timed iterations",
                         "elapsed_s": tcpu, "gdof_it_per_s": Nd * its / tcpu / 1e9,
                         "gbs": nbytes / tcpu / 1e9, "gflops": nflops / tcpu / 1e9})
            log(f"cpu config 1 ({variant}, {name}): {rows[-1]['elapsed_s']:.4f}s for 100 its")
        del Pc
    out["problem_10cube_p4"] = rows
    oracle_lib.set_threads(omp)
    return out


def roofline_object(ins, op32, nbytes32, achieved32, copy_gbs, Nel, args):
    """The roofline of the dominant kernel of the timed solve, the FP32 CG
    operator's Ax. Its bound is VALU issue (it forms the geometric factors
    from the trilinear maps instead of streaming them): the committed SQ
    counter summary gives its VALU instructions per launch, each wave64 FP32
    instruction takes 4 cycles of one of the 1,024 SIMDs, and the launch's
    own duration, timed in this run inside one more mixed solve, gives the
    issue rate it reaches. The HBM figure (the reference byte model with
    4-byte values and the map coefficients in place of the stored factors)
    is reported beside it."""
    t = ins["ax_us"] * 1e-6
    hbm = {"achieved": achieved32, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved32 / HBM_PEAK_GBS,
           "algorithmic_bytes": nbytes32, "traffic": measured_traffic(TRAFFIC_FILE_F32),
           "traffic_source": "committed counter summary " + os.path.relpath(TRAFFIC_FILE_F32, ROOT) +
                             " (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by a 512 MiB copy); not measured in "
                             "this run"}
    out = {"kernel": ("k_bp5_ax FP32 CG operator, " +
                      ("fused (interior rows + surface run)" if ins["fused"] else "plain") +
                      ", geometric factors from the trilinear map, non-affine, Nq=8"),
           "us_per_launch": ins["ax_us"], "launches_timed": ins["launches"],
           "timed_in": "one more mixed solve identical to the timed ones (dispatch-stamped HIP events)",
           "gather_us_per_launch": ins["gather_us"], "us_per_launch_back_to_back": op32["ax_us"],
           "device_copy_gbs": copy_gbs,
           "model_gflops": bp5_ax_flops(args.N + 1, Nel) / t / 1e9}
    lim = kernel_limiter(LIMITER_FILE_F32)
    if lim and "valu_instructions_per_launch" in lim:
        instr = lim["valu_instructions_per_launch"]
        lanes_per_clk = 64 / lim["valu_cycles_per_instruction"]
        peak = lim["simds"] * lanes_per_clk * lim["clock_ghz"] * 1e9  # VALU lane-operations per second
        achieved = instr * 64 / t
        out.update({"bound": "valu-issue", "achieved": achieved / 1e12, "peak": peak / 1e12,
                    "unit": "T VALU lane-ops/s", "frac": achieved / peak, "traffic": hbm["traffic"],
                    "valu_instructions_per_launch": instr,
                    "valu_issue_us": instr * lim["valu_cycles_per_instruction"] / (lim["simds"] * lim["clock_ghz"] * 1e9) * 1e6,
                    "counter_tflops": lim.get("counter_flops_per_launch", 0) / t / 1e12,
                    "limiter_source": lim["source"], "hbm": hbm})
    else:
        out.update({"bound": "hbm", "achieved": achieved32, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved32 / HBM_PEAK_GBS, "traffic": hbm["traffic"], "hbm": hbm})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--N", type=int, default=7)
    ap.add_argument("--nx", type=int, default=24)
    ap.add_argument("--tol", type=float, default=1e-10)
    # the solve stops at a correction, so it overshoots by about interval/2
    # FP32 iterations (~108 us each at 24^3, p=7) against N/interval
    # corrections (~215 us each): the sum is least near interval ~ 70 for
    # N ~ 1270; measured 140.2/138.0/136.7/133.7/138.4 ms at 24/32/48/64/96
    ap.add_argument("--interval", type=int, default=64,
                    help="FP32 iterations between FP64 defect corrections")
    ap.add_argument("--ax-launches", type=int, default=50)
    ap.add_argument("--cg-fold", type=int, default=1, choices=(0, 1, 2, 3),
                    help="where the fused CG operator also does the interior rows' direction update: "
                         "bit 0 = FP64 solves (default 1), bit 1 = FP32 iterations (same bits)")
    ap.add_argument("--setup", choices=("device", "host"), default="device",
                    help="where the mesh, ogs maps and solve-space maps are built (same arrays either way)")
    ap.add_argument("--no-setup-compare", action="store_true",
                    help="skip timing the other setup of the same problem")
    ap.add_argument("--setup-only", action="store_true",
                    help="only time the device and the host setup of the problem (mesh, then BP5 problem)")
    ap.add_argument("--cpu-its", type=int, default=1000)
    ap.add_argument("--cpu-serial-its", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extensions", action="store_true",
                    help="skip the k-RHS, batch and adjoint measurements")
    ap.add_argument("--ax-only", action="store_true",
                    help="only the Ax launches and a 512 MiB device copy (for counter runs)")
    ap.add_argument("--counter-kernels", action="store_true",
                    help="only 3 launches each of the FP64 Ax of BK5 p=15 (vector-ALU and matrix-core "
                         "kernels), BP3 p=7 and BK5 p=7 on the 24^3 perturbed box (for counter runs)")
    ap.add_argument("--krhs-only", action="store_true",
                    help="only 3 single-vector and 3 k-RHS (K = 32, --krhs-chunk) FP64 BP5 Ax launches "
                         "on the 20^3 p=7 perturbed box, random right-hand sides (for counter runs)")
    ap.add_argument("--krhs-chunk", type=int, default=0)
    ap.add_argument("--cub-only", action="store_true",
                    help="only 10 launches each of the FP64 BP1-BP4 Ax (default kernel form) on the 24^3 p=7 "
                         "perturbed box and one 512 MiB device copy (for counter runs)")
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
    if args.krhs_only:
        K = 32
        mk = capi.Mesh(N=7, nx=20, perturb=0.1, seed=1)
        bk = capi.BP(plat, mk, 5, lam=1.0)
        q1 = torch.from_numpy(capi.random_normal(3, bk.Ngather)).cuda()
        A1 = torch.empty(bk.NLocal, dtype=torch.float64, device="cuda")
        qk = torch.from_numpy(capi.random_normal(4, bk.Ngather * K)).cuda()
        Ak = torch.empty(bk.Ngather * K, dtype=torch.float64, device="cuda")
        for _ in range(3):
            bk.ax_local(q1, A1)
        for _ in range(3):
            bk.operator_krhs(qk, Ak, K, args.krhs_chunk)  # the k-RHS Ax, then a gather per right-hand side
        torch.cuda.synchronize()
        log(f"krhs-only: 3 single-vector Ax and 3 k-RHS operator launches (K = {K}, chunk "
            f"{args.krhs_chunk}), 20^3 p=7, Ndofs {bk.Ngather}")
        return
    if args.cub_only:
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        del src, dst
        for bpn in (1, 2, 3, 4):
            mk = capi.Mesh(N=7, nx=args.nx, cubature=True, perturb=0.1, seed=1)
            bk = capi.BP(plat, mk, bpn, lam=1.0)
            qk = torch.from_numpy(capi.random_normal(2, bk.Ngather)).cuda()
            AqLk = torch.empty(bk.NLocal, dtype=torch.float64, device="cuda")
            for _ in range(10):
                bk.ax_local(qk, AqLk)
            torch.cuda.synchronize()
            log(f"cub-only: BP{bpn} p=7 {args.nx}^3, 10 Ax launches")
            del bk, mk, qk, AqLk
            torch.cuda.empty_cache()
        return
    if args.counter_kernels:
        for bp, N, cub, engines in ((5, 15, False, (1, 2)), (3, 7, True, (0,)), (5, 7, False, (0,))):
            mk = capi.Mesh(N=N, nx=args.nx, cubature=cub, perturb=0.1, seed=1)
            bk = capi.BP(plat, mk, bp, lam=1.0)
            qk = torch.from_numpy(capi.random_normal(2, bk.Ngather)).cuda()
            AqLk = torch.empty(bk.NLocal, dtype=torch.float64, device="cuda")
            for eng in engines:
                bk.set_ax_engine(eng)
                for _ in range(3):
                    bk.ax_local(qk, AqLk)
                torch.cuda.synchronize()
                log(f"counter kernels: BP{bp} p={N} engine {eng}, 3 Ax launches")
            del bk, mk, qk, AqLk
            torch.cuda.empty_cache()
        return

    def build(where):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        m = capi.Mesh(N=args.N, nx=args.nx, perturb=0.1, seed=1, platform=plat if where == "device" else None)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        b = capi.BP(plat, m, 5, lam=1.0)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        log(f"{where} setup: mesh {t1 - t0:.2f}s, BP5 problem {t2 - t1:.2f}s (of which the launch-choice "
            f"timing is the same either way)")
        return m, b, {"mesh_s": t1 - t0, "problem_s": t2 - t1, "total_s": t2 - t0}

    setup = {"workload": f"BP5 hex p={args.N} {args.nx}^3 perturbed box: mesh (nodes, ids, geometric "
                         f"factors), then the problem (masked ids, ogs T maps, GlobalToLocal, fused "
                         f"operator's solve-space maps, operator data, launch-choice timing)"}
    mesh, bp, setup[args.setup] = build(args.setup)
    other = "host" if args.setup == "device" else "device"
    if args.setup_only or not args.no_setup_compare:
        mo, bo, setup[other] = build(other)
        del mo, bo
        torch.cuda.empty_cache()
        setup["host_over_device"] = setup["host"]["total_s"] / setup["device"]["total_s"]
        setup["host_over_device_without_problem"] = setup["host"]["mesh_s"] / setup["device"]["mesh_s"]
    if args.setup_only:
        print(json.dumps({"setup": setup}), flush=True)
        return
    bp.set_cg_fold(args.cg_fold)
    Ndofs, Nel, Np = bp.Ngather, mesh.Nelements, mesh.Np
    log(f"setup ({args.setup}): {setup[args.setup]['total_s']:.1f}s  Ndofs={Ndofs}  Nel={Nel}  p={args.N}")

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
        op32 = bp.time_operator(32, args.ax_launches)
        torch.cuda.synchronize()
        log(f"ax-only: {args.ax_launches} BK5 Ax launches, {args.ax_launches} operator (Ax + gather) "
            f"launches, {args.ax_launches + 3} FP32 CG operator launches (fused: {op32['fused']}) "
            f"and one 512 MiB copy done")
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
    x_mixed = x.clone()

    # ---- the same mixed solve with the FP32 operator streaming the stored
    # geometric factors rounded to FP32 (the reference kernel with dfloat =
    # float), timed the same way; the default forms them from the elements'
    # trilinear maps
    bp.set_fp32_geometry(1)
    for _ in range(args.warmup):
        mixed_solve()
    tmix_s, st_s = timed(mixed_solve, args.steps)
    bp.operator(x, Ax)
    relmix_s = float(torch.linalg.norm(b - Ax) / torch.linalg.norm(b))
    x.zero_()
    ins_s = bp.time_operator_in_solve(x, b, tol=args.tol, interval=args.interval)
    stored32 = bk5_bytes(Ndofs, Nel, Np, 4)
    mixed_stored = {"fp32_geometry": "stored factors rounded to FP32", "fp32_iterations": st_s["iterations"],
                    "fp64_corrections": st_s["updates"], "ms_per_solve": tmix_s * 1e3, "true_rel_res": relmix_s,
                    "fp64_equivalent_gdof_it_per_s": Ndofs * its64 / tmix_s / 1e9,
                    "speedup_vs_fp64": t64 / tmix_s,
                    "ax_us_in_solve": ins_s["ax_us"], "ax_algorithmic_bytes": stored32,
                    "ax_gbs": stored32 / (ins_s["ax_us"] * 1e-6) / 1e9,
                    "ax_hbm_frac": stored32 / (ins_s["ax_us"] * 1e-6) / 1e9 / HBM_PEAK_GBS}
    log(f"mixed, stored FP32 factors: {st_s['iterations']} FP32 its, {tmix_s * 1e3:.1f} ms/solve, "
        f"Ax {ins_s['ax_us']:.1f} us in the solve")
    bp.set_fp32_geometry(0)
    mixed_solve()  # rebuild the default FP32 operator before it is timed below

    # ---- the reference's own timed protocol on the FP64 CG (100 iterations)
    refp = reference_protocol(bp, b, Ndofs, bp.NLocal, Nel, args.N + 1)
    log(f"reference protocol (FP64 CG, 100 its): {refp['elapsed_s'] * 1e3:.1f} ms, {refp['gbs']:.0f} GB/s, "
        f"{refp['gflops']:.0f} GFLOP/s, {refp['gdof_it_per_s']:.2f} GDOF*it/s")

    # ---- roofline: the Ax of the FP32 CG operator, the dominant kernel of the
    # timed solve, timed inside one more solve identical to the timed ones
    # (events around each application on the library's stream); back to
    # back for comparison (its inputs then stay partly in the 256 MB
    # Infinity Cache)
    x.zero_()
    ins = bp.time_operator_in_solve(x, b, tol=args.tol, interval=args.interval)
    op32 = bp.time_operator(32, args.ax_launches)
    nbytes32 = bk5_tri_bytes(Ndofs, Nel, Np)
    achieved32 = nbytes32 / (ins["ax_us"] * 1e-6) / 1e9
    log(f"FP32 CG operator in the solve ({ins['launches']} applications): Ax {ins['ax_us']:.1f} us "
        f"({achieved32:.0f} GB/s), gather {ins['gather_us']:.1f} us, fused {ins['fused']}; back to back: "
        f"Ax {op32['ax_us']:.1f} us, gather {op32['gather_us']:.1f} us")

    # ---- FP64 BK5 Ax (HIP events on the launch stream = torch's current stream)
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
        cb = cpu_baselines(args, b_host, Ndofs)
        omp_run = cb["problem_24cube_p7"]["openmp"]
        cpu = {"value": omp_run["value"], "unit": "GDOF·it/s", "cores": omp_run["threads"], "kind": "port",
               "sample": f"oracle FP64 CG (restated OKL loop nests, OpenMP over elements), same {args.nx}^3 "
                         f"p={args.N} BP5 problem and right-hand side, {omp_run['iterations']} iterations in "
                         f"{omp_run['seconds']:.1f}s",
               "physical_cores": cb["physical_cores"], "serial": cb["problem_24cube_p7"]["serial"],
               "config1_10cube_p4": cb["problem_10cube_p4"]}

    ext = None
    if not args.no_extensions:
        # the BP Ax rows first: run after the others (the p = 15 problem takes
        # several GB), the same kernels read up to 18% slower (BP2 112.8 against
        # 95.3 us at 24^3, p = 7)
        ext = {"other_bps": bench_other_bps(capi, plat), "bp3": bench_bp3(capi, plat, args),
               "families": bench_families(capi, plat),
               "krhs_ax": bench_krhs(capi, plat), "ragged_batch_cg": bench_batch(capi, plat),
               "adjoint": bench_adjoint(capi, plat), "bk5_p15": bench_p15(capi, plat)}
        log("extensions: " + json.dumps({k: {kk: vv for kk, vv in v.items() if kk in
                                                 ("ms_per_solve", "speedup_vs_sequential",
                                                  "ms_forward_adjoint_gradient", "single_us",
                                                  "speedup_vs_fp64", "gbs", "ax_gbs")}
                                             for k, v in ext.items()}))

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
                    "ms_per_solve": t64 * 1e3, "true_rel_res": rel64,
                    "gflops": cg_solve_model(Ndofs, bp.NLocal, nbytes, bp5_ax_flops(args.N + 1, Nel), its64)[1]
                    / t64 / 1e9,
                    "gbs": cg_solve_model(Ndofs, bp.NLocal, nbytes, bp5_ax_flops(args.N + 1, Nel), its64)[0]
                    / t64 / 1e9,
                    "reference_protocol": refp},
        "mixed_stored_factors": mixed_stored,
        "mixed": {"fp32_iterations": st["iterations"], "fp64_corrections": st["updates"],
                  "ms_per_solve": tmix * 1e3, "true_rel_res": relmix,
                  "own_rate": Ndofs * st["iterations"] / tmix / 1e9,
                  "speedup_vs_fp64": t64 / tmix},
        "roofline": roofline_object(ins, op32, nbytes32, achieved32, copy_gbs, Nel, args),
        "bk5_fp64": {"achieved": achieved, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(), "kernel": "k_bp5_ax (FP64, non-affine, Nq=8)",
                     "us_per_launch": t_ax * 1e6, "algorithmic_bytes": nbytes,
                     "gdof_per_s": Ndofs / t_ax / 1e9,
                     "gflops": bp5_ax_flops(args.N + 1, Nel) / t_ax / 1e9},
        "tuning": bp.tuning(),
        "setup": setup,
        "cpu_baseline": cpu,
        "extensions": ext,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
