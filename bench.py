"""Benchmark: fp64 CSR SpMV HBM bandwidth and JPCG time-to-1e-10 (fp64 vs mixed-refined) on one MI355X.

Workload (default): 3D 27-point SPD operator on a 160^3 grid (n = 4,096,000,
nnz = 109,215,352, diag 26, off-diagonals -1, CSR fp64 values, int32
indices); b = A*1 + U[-1e-3, 1e-3] (seed 1); tolerance 1e-10.  Inputs are
generated on the host, uploaded, and resident in HBM before any timed region.

  value       fp64 SpMV algorithmic HBM GB/s (B = 12 nnz + 4(n+1) + 16 n per launch),
              100 launches timed with HIP events on the launch stream
  steps       timed JPCG solves per precision (x0 = 0 to ||r||/||b|| < 1e-10)
  ms_per_step fp64 JPCG time-to-1e-10 (median of the timed solves)
  jpcg        fp64 and mixed-refined times, iterations, true residuals

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--grid 160] [--no-cpu]
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

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spmv_bytes(n, nnz, vbytes=8):
    return (vbytes + 4) * nnz + 4 * (n + 1) + 16 * n


def iter_bytes(n, nnz):
    return 12 * nnz + 4 * (n + 1) + 112 * n


def cpu_baseline(A, budget_s=15.0):
    """Oracle SpMV (the project's C restatement) on host cores: 1 thread and all threads."""
    from oracle import oracle
    x = np.ones(A.n)
    nb = spmv_bytes(A.n, A.nnz)
    out = {}
    threads_all = min(os.cpu_count() or 1, 16)
    for th in (1, threads_all):
        oracle.spmv(A.rowptr, A.col, A.val, x, threads=th)  # warm
        reps, t0 = 0, time.perf_counter()
        while True:
            oracle.spmv(A.rowptr, A.col, A.val, x, threads=th)
            reps += 1
            el = time.perf_counter() - t0
            if el > budget_s / 2 or reps >= 20:
                break
        out[th] = (nb * reps / el / 1e9, reps)
    return out, threads_all


LAYOUT_NAMES = {1: "tiles", 2: "panels", 3: "patterns", 4: "diagonals"}
LAYOUT_KERNELS = {1: "spmv_tiles<double, %s columns>", 2: "spmv_panels<double>", 3: "spmv_pattern<double>",
                  4: "spmv_dia<double>"}
TRAFFIC_PROFILES = {"tiles": "spmv_tiles_stencil27_160.md", "patterns": "spmv_pattern_stencil27_160.md",
                    "diagonals": "spmv_dia_stencil27_160.md"}


def measured_traffic(grid, layout_name, compact=True):
    """HBM bytes per SpMV launch from the committed counter summary of that kernel (only for the workload it
    measured: the 27-point 160^3 operator; the tile kernel on its int16 column stream)."""
    import re
    fname = TRAFFIC_PROFILES.get(layout_name)
    if grid != 160 or fname is None or (layout_name == "tiles" and not compact):
        return None, None
    path = os.path.join(ROOT, "profiles", fname)
    if not os.path.exists(path):
        return None, None
    m = re.search(r"^traffic_bytes_per_launch:\s*(\d+)", open(path).read(), re.M)
    return (int(m.group(1)), "profiles/" + fname) if m else (None, None)


def achievable_copy(dev, stream, nbytes=2 << 30, reps=10):
    """Achievable HBM rate: the library's stream kernel copying nbytes, (read + write bytes) / time, median of
    reps."""
    import torch
    from hpc_amd import device
    src = torch.empty(nbytes // 8, dtype=torch.float64, device=dev).fill_(1.0)
    dst = torch.empty_like(src)
    device.stream_copy(src, dst)
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        device.stream_copy(src, dst)
        e1.record(stream)
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1))
    assert bool((dst == 1.0).all())
    del src, dst
    return 2 * nbytes / (float(np.median(ts)) * 1e-3) / 1e9


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def spmv_profile(args):
    """Short run for hardware-counter passes: the library's stream kernel copying 256 MiB (the calibration
    kernel of known size), then 10 launches of the fp64 SpMV in the layout the operator chooses, on the
    27-point 160^3 operator or (--profile-workload graph) the graph Laplacian."""
    import torch
    from hpc_amd import device, workloads
    dev = torch.device("cuda:0")
    if args.profile_workload == "graph":
        A = workloads.c3_graph(args.graph_n).A
    else:
        A = workloads.c2_stencil27(args.grid).A
    op = device.CsrOperator(torch.from_numpy(A.rowptr).to(dev), torch.from_numpy(A.col).to(dev),
                            torch.from_numpy(A.val).to(dev), rowptr_host=A.rowptr)
    src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    device.stream_copy(src, dst)
    x = torch.ones(A.n, dtype=torch.float64, device=dev)
    y = torch.empty_like(x)
    for _ in range(10):
        op.spmv(x, y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_each_way": src.numel() * 8, "spmv_algorithmic_bytes": spmv_bytes(A.n, A.nnz),
                      "layout": LAYOUT_NAMES[op.layout], "compact_columns": op.compact, "n": A.n, "nnz": A.nnz}),
          flush=True)


def _events(stream):
    import torch
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    return e0, e1


def _timed(fn, steps, warmup, stream):
    """Runs fn warmup times untimed, then steps times between HIP events on `stream`; (ms list, last result)."""
    import torch
    out = None
    for _ in range(warmup):
        out = fn()
    torch.cuda.synchronize()
    ms = []
    for _ in range(steps):
        e0, e1 = _events(stream)
        out = fn()
        e1.record(stream)
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    return ms, out


def _spmv_rate(op, n, stream, repeats, nbytes):
    import torch
    x = torch.ones(n, dtype=torch.float64, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(3):
        op.spmv(x, y)
    e0, e1 = _events(stream)
    for _ in range(repeats):
        op.spmv(x, y)
    e1.record(stream)
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / repeats
    return {"us": round(us, 2), "GBs": round(nbytes / (us * 1e-6) / 1e9, 1), "bytes_per_launch": nbytes}


def bench_graph(args):
    """Graph Laplacian + 1e-2 I (lognormal degrees mu 2.5 sigma 1 capped at 2000, weights U[0.1,1] in fp32,
    fp64 diagonal): fp64 JPCG vs the fp32-valued refined solve to 1e-10, both checked against the oracle
    (the same JPCG on host cores)."""
    import torch
    from hpc_amd import device, workloads
    from oracle import oracle
    dev = torch.device("cuda:0")
    t0 = time.time()
    sysm = workloads.c3_graph(args.graph_n)
    A = sysm.A
    deg = np.diff(A.rowptr)
    log(f"graph n={A.n} nnz={A.nnz} max row {deg.max()} generated in {time.time() - t0:.1f}s")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    op = device.CsrOperator(t(A.rowptr), t(A.col), t(A.val), rowptr_host=A.rowptr)
    srp, scol, sv32, sdiag, exact = device.split_fp32(A.rowptr, A.col, A.val)
    inner = device.CsrOperator(t(srp), t(scol), t(sv32), dsplit=t(sdiag), rowptr_host=srp)
    # residual operator: the split one when it equals A (fewer bytes), else the fp64 one; a separate
    # object from the inner operator so their work buffers stay apart
    outer = device.CsrOperator(t(srp), t(scol), t(sv32), dsplit=t(sdiag), rowptr_host=srp) if exact else op
    diag, b = t(A.diag), t(sysm.b)
    stream = torch.cuda.current_stream()
    # SpMV in both layouts (row tiles, panel-swept copy); the solves use each operator's creation-time choice
    chosen = {"fp64": op.layout, "fp32_values": inner.layout}
    sp64, sp32 = {}, {}
    for lay, name in ((1, "tiles"), (2, "panels")):
        op.set_layout(lay)
        inner.set_layout(lay)
        sp64[name] = _spmv_rate(op, A.n, stream, args.spmv_repeats, spmv_bytes(A.n, A.nnz))
        sp32[name] = _spmv_rate(inner, A.n, stream, args.spmv_repeats, 8 * inner.nnz + 4 * (A.n + 1) + 24 * A.n)
    for o in (op, inner, outer):
        o.set_layout(0)
    log(f"SpMV fp64 {sp64}  fp32-values {sp32}  layouts at creation {chosen} op={op.info()} compact={op.compact}")
    x = torch.empty_like(b)
    ms64, (x, info) = _timed(lambda: op.solve(diag, b, sysm.max_iter, sysm.tol, x=x), args.steps, args.warmup, stream)
    xm = torch.empty_like(b)
    msm, (xm, minfo) = _timed(lambda: device.solve_mixed(outer, inner, diag, b, sysm.tol, inner_tol=args.inner_tol,
                                                         x=xm), args.steps, args.warmup, stream)
    bb = float(torch.dot(b, b))
    true64 = (op.residual(b, x)[1] / bb) ** 0.5
    truem = (op.residual(b, xm)[1] / bb) ** 0.5
    log(f"fp64 {np.mean(ms64):.2f} ms ({info.executed} it, true res {true64:.2e}); mixed {np.mean(msm):.2f} ms "
        f"(outer {minfo.outer}, inner {minfo.inner_total}, true res {truem:.2e})")
    th = min(os.cpu_count() or 1, 16)
    log(f"oracle JPCG on {th} host threads ...")
    t0 = time.perf_counter()
    ref = oracle.jpcg(A.rowptr, A.col, A.val, A.diag, sysm.b, sysm.max_iter, sysm.tol, threads=th)
    cpu_s = time.perf_counter() - t0
    xr = ref["x"]
    nr = np.linalg.norm(xr)
    e64 = float(np.linalg.norm(x.cpu().numpy() - xr) / nr)
    em = float(np.linalg.norm(xm.cpu().numpy() - xr) / nr)
    out = {"metric": "graph-Laplacian JPCG time-to-1e-10: mixed-refined vs fp64, 1 GPU",
           "workload": f"graph Laplacian + 1e-2 I, n={A.n}, nnz={A.nnz}, degrees lognormal(2.5, 1.0) cap 2000, "
                       "weights U[0.1,1] (fp32-exact), b U[-1,1] seed 2, tol 1e-10",
           "steps": args.steps, "warmup": args.warmup, "split_exact": exact,
           "spmv": {"fp64": sp64, "fp32_values": sp32, "compact_columns": op.compact,
                    "layout_used": {k: ("panels" if v == 2 else "tiles") for k, v in chosen.items()}},
           "fp64": {"ms": round(float(np.mean(ms64)), 3), "iterations": info.executed, "true_res": true64,
                    "rel_err_vs_oracle": e64},
           "mixed": {"ms": round(float(np.mean(msm)), 3), "outer": minfo.outer, "inner_total": minfo.inner_total,
                     "true_res": truem, "rel_err_vs_oracle": em},
           "speedup_mixed_vs_fp64": round(float(np.mean(ms64) / np.mean(msm)), 3),
           "oracle": {"iterations": ref["executed"], "res": ref["res"], "seconds": round(cpu_s, 2), "threads": th}}
    print(json.dumps(out), flush=True)


def bench_batched(args):
    """512 anisotropic 2D/3D diffusion systems (n log-uniform in [5k, 400k]) x 16 right-hand sides, tol 1e-8:
    batched fp64 JPCG vs the fp32-valued refined batch, with per-pair checks against the oracle on a sample."""
    import torch
    from hpc_amd import batched, workloads
    from oracle import oracle
    dev = torch.device("cuda:0")
    t0 = time.time()
    mats, rhs = workloads.c4_systems(args.batch_count, 5000, 400_000, args.nrhs, seed=3)
    asm64 = batched.assemble(mats)
    asm32 = batched.assemble(mats, split_fp32=True)
    N, R = int(asm64["sys_row"][-1]), args.nrhs
    log(f"{len(mats)} systems, N={N}, nnz={asm64['col'].size}, R={R}, generated in {time.time() - t0:.1f}s")
    bh = np.concatenate(rhs)
    del rhs
    b = torch.from_numpy(bh).to(dev)
    diag = torch.from_numpy(asm64["diag"]).to(dev)
    B64 = batched.Batch.from_host(asm64, R, dev)
    outer = batched.Batch.from_host(asm32, R, dev)
    inner = batched.Batch.from_host(asm32, R, dev)
    stream = torch.cuda.current_stream()
    t0 = time.time()
    B64.configure(tune=True)
    inner.configure(tune=True)
    log(f"autotuned in {time.time() - t0:.1f}s: fp64 {B64.info()} fp32 {inner.info()}")
    tol, maxit = 1e-8, 100000
    x = torch.empty_like(b)
    ms64, (x, info) = _timed(lambda: B64.solve(diag, b, maxit, tol, x=x), args.steps, args.warmup, stream)
    its = info.iterations.cpu().numpy()
    log(f"fp64 batch: {np.mean(ms64):.1f} ms, {info.launched} launched iterations, pair iterations "
        f"{its.min()}..{its.max()}, max res {float(info.residual.max()):.2e}")
    xm = torch.empty_like(b)
    msm, (xm, minfo, steps) = _timed(lambda: batched.solve_mixed(outer, inner, diag, b, tol, inner_tol=args.inner_tol,
                                                                 x=xm), args.steps, args.warmup, stream)
    log(f"mixed batch: {np.mean(msm):.1f} ms, outer {steps}, launched {minfo.launched}, max true res "
        f"{float(minfo.residual.max()):.2e}")
    sizes = np.diff(asm64["sys_row"])
    order = np.argsort(sizes)
    picks = [int(order[0]), int(order[len(order) // 4]), int(order[len(order) // 2])]
    th = min(os.cpu_count() or 1, 16)
    xh, xmh = x.cpu().numpy(), xm.cpu().numpy()
    checks = []
    for s in picks:
        r0, r1 = int(asm64["sys_row"][s]), int(asm64["sys_row"][s + 1])
        A = mats[s]
        for j in (0, R - 1):
            ref = oracle.jpcg(A.rowptr, A.col, A.val, A.diag, bh[r0:r1, j].copy(), maxit, tol, threads=th)
            nr = np.linalg.norm(ref["x"])
            checks.append({"system": s, "n": A.n, "rhs": j, "oracle_iterations": ref["executed"],
                           "iterations": int(its[s, j]),
                           "fp64_rel_err": float(np.linalg.norm(xh[r0:r1, j] - ref["x"]) / nr),
                           "mixed_rel_err": float(np.linalg.norm(xmh[r0:r1, j] - ref["x"]) / nr)})
            log(f"check {checks[-1]}")
    out = {"metric": "batched JPCG time-to-1e-8 (all pairs): mixed-refined vs fp64, 1 GPU",
           "workload": f"{len(mats)} anisotropic 2D5/3D7 diffusion systems, n log-uniform [5k,400k] (N={N}, "
                       f"nnz={asm64['col'].size}), {R} rhs U[-1,1] seed 3, tol 1e-8",
           "steps": args.steps, "warmup": args.warmup,
           "fp64": {"ms": round(float(np.mean(ms64)), 2), "launched_iterations": info.launched,
                    "pair_iterations_max": int(its.max()), "pair_iterations_sum": int(its.sum()),
                    "max_res": float(info.residual.max()), "config": B64.info()},
           "mixed": {"ms": round(float(np.mean(msm)), 2), "outer": steps, "launched_iterations": minfo.launched,
                     "inner_tol": args.inner_tol, "max_true_res": float(minfo.residual.max()),
                     "config": inner.info()},
           "speedup_mixed_vs_fp64": round(float(np.mean(ms64) / np.mean(msm)), 3),
           "checks": checks}
    print(json.dumps(out), flush=True)


def bench_ingest(args):
    """One-shot pcg.h solve from a shuffled COO of the 27-point operator (--grid, default 160^3): the
    reference's metric phases (matrix processing, vectors, solver) with the CSR built on the GPU, beside the
    oracle's host COO-to-CSR on the same arrays."""
    import torch
    from hpc_amd import workloads
    from hpc_amd.pcg import JPCG
    from oracle import oracle
    sysm = workloads.c2_stencil27(args.grid)
    A = sysm.A
    rng = np.random.default_rng(5)
    p = rng.permutation(A.nnz)
    rows = np.repeat(np.arange(A.n, dtype=np.uint32), np.diff(A.rowptr))[p]
    cols = A.col.astype(np.uint32)[p]
    vals = A.val[p]
    del p
    log(f"COO of n={A.n}, nnz={A.nnz} shuffled")
    out = {}
    with JPCG(0) as h:
        for rep in range(2):
            t0 = time.perf_counter()
            res = h.coo_solve(A.n, rows, cols, vals, A.diag, sysm.b, sysm.max_iter, sysm.tol)
            wall = time.perf_counter() - t0
            m = h.metrics()
            out[f"call{rep}"] = {"wall_s": round(wall, 3), "matProc_s": round(m.m_matProc, 3),
                                 "vecProc_s": round(m.m_vecProc, 3), "solver_s": round(m.m_solver, 3),
                                 "iterations": res.iterations, "status": int(res.status)}
            log(f"pcg.h call {rep}: {out[f'call{rep}']}")
    t0 = time.perf_counter()
    oracle.coo_to_csr(A.n, rows, cols, vals)
    cpu = time.perf_counter() - t0
    log(f"oracle host COO->CSR (1 thread): {cpu:.2f} s")
    print(json.dumps({"metric": "one-shot xJPCG_cooSolver phases, CSR built on the GPU",
                      "workload": f"27-point stencil {args.grid}^3 as a shuffled COO (n={A.n}, nnz={A.nnz})",
                      **out, "oracle_coo_to_csr_s": round(cpu, 3)}), flush=True)


def bench_signature(args):
    """Signature files (the reference's partitioned-matrix format) of the 27-point operator (--grid): encode on
    the host, store / load, then to a device CSR with the streams decoded on the GPU, beside the host decode;
    the CSR is checked entry for entry against the generator's.  The reference's own partitioner (oracle/_ref,
    where built) and this encoder are timed on a bounded 40^3 sample."""
    import shutil
    import tempfile

    import torch
    from hpc_amd import workloads
    from hpc_amd.signature import Signature
    from oracle import oracle
    A = workloads.c2_stencil27(args.grid).A
    rows = np.repeat(np.arange(A.n, dtype=np.uint32), np.diff(A.rowptr))
    cols = A.col.astype(np.uint32)
    out = {}
    t0 = time.perf_counter()
    sig = Signature.from_coo(A.n, A.n, rows, cols, A.val)
    out["encode_s"] = round(time.perf_counter() - t0, 3)
    info = sig.info()
    sbytes = sum(len(sig.buffer(c)) for c in range(16))
    log(f"encoded n={A.n} nnz={A.nnz}: {info}, streams {sbytes / 1e9:.2f} GB in {out['encode_s']} s")
    root = "/dev/shm" if os.path.isdir("/dev/shm") else None
    d = tempfile.mkdtemp(prefix="sig_", dir=root)
    try:
        t0 = time.perf_counter()
        sig.store(d)
        out["store_s"] = round(time.perf_counter() - t0, 3)
        del sig
        t0 = time.perf_counter()
        sig = Signature.load(d)
        out["load_s"] = round(time.perf_counter() - t0, 3)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    dev = torch.device("cuda:0")
    times = []
    for rep in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rp, col, val = sig.to_csr(dev)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        if rep == 0:
            same = (np.array_equal(rp.cpu().numpy(), A.rowptr) and np.array_equal(col.cpu().numpy(), A.col)
                    and np.array_equal(val.cpu().numpy(), A.val))
            log(f"signature -> device CSR equals the generator's CSR: {same}")
            out["csr_equal"] = bool(same)
        del rp, col, val
    out["to_csr_s"] = [round(t, 3) for t in times]
    t0 = time.perf_counter()
    r, c, v = sig.to_coo()
    out["host_decode_s"] = round(time.perf_counter() - t0, 3)
    del r, c, v
    log(f"timings: {out}")
    sample = {}
    S = workloads.c2_stencil27(40).A
    srow = np.repeat(np.arange(S.n, dtype=np.uint32), np.diff(S.rowptr))
    t0 = time.perf_counter()
    Signature.from_coo(S.n, S.n, srow, S.col.astype(np.uint32), S.val)
    sample["encode_s"] = round(time.perf_counter() - t0, 3)
    if oracle.ref_lib() is not None:
        d = tempfile.mkdtemp(prefix="sigref_", dir=root)
        try:
            t0 = time.perf_counter()
            oracle.ref_sig_coo(d, S.n, S.n, srow, S.col, S.val)
            sample["reference_partitioner_s"] = round(time.perf_counter() - t0, 3)
        finally:
            shutil.rmtree(d, ignore_errors=True)
    log(f"40^3 sample (n={S.n}, nnz={S.nnz}): {sample}")
    print(json.dumps({"metric": "signature files to device CSR",
                      "workload": f"27-point stencil {args.grid}^3 (n={A.n}, nnz={A.nnz})",
                      "nnz_pad": info["nnz_pad"], "row_blocks": info["row_blocks"],
                      "partitions": info["partitions"], "stream_bytes": sbytes, **out,
                      "sample_40cubed": sample}), flush=True)


def batched_profile(args):
    """Short run for hardware-counter passes on the batched kernels: the batch of --batch-count systems
    x --nrhs right-hand sides, 10 fp64 iterations (set-up plus 10 graph-free iteration launches)."""
    import torch
    from hpc_amd import batched, workloads
    dev = torch.device("cuda:0")
    mats, rhs = workloads.c4_systems(args.batch_count, 5000, 400_000, args.nrhs, seed=3)
    asm = batched.assemble(mats)
    b = torch.from_numpy(np.concatenate(rhs)).to(dev)
    diag = torch.from_numpy(asm["diag"]).to(dev)
    B = batched.Batch.from_host(asm, args.nrhs, dev)
    x, info = B.solve(diag, b, 9, 1e-30)
    torch.cuda.synchronize()
    print(json.dumps({"N": B.N, "nnz": int(asm["col"].size), "R": args.nrhs, "launched": info.launched,
                      "config": B.info()}), flush=True)


def bench_autograd(args):
    """Differentiable solve on 2D variable-coefficient diffusion 512^2: forward JPCG, then backward
    (adjoint JPCG + SDDMM) for L = 0.5 ||x - x_target||^2, fp64."""
    import torch
    from hpc_amd import autograd, workloads
    dev = torch.device("cuda:0")
    A, bh, xt = workloads.c5_diffusion()
    pat = autograd.SparsePattern(A.rowptr, A.col, dev)
    base = torch.from_numpy(A.val.copy()).to(dev)
    b0 = torch.from_numpy(bh).to(dev)
    xt_d = torch.from_numpy(xt).to(dev)
    stream = torch.cuda.current_stream()
    tol = 1e-12

    def fwd_bwd():
        v = base.clone().requires_grad_(True)
        bt = b0.clone().requires_grad_(True)
        e0, e1 = _events(stream)
        x = autograd.jpcg_solve(pat, v, bt, tol=tol, max_iter=200000)
        e1.record(stream)
        loss = 0.5 * ((x - xt_d) ** 2).sum()
        e2, e3 = _events(stream)
        loss.backward()
        e3.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1), e2.elapsed_time(e3), v.grad, bt.grad

    for _ in range(args.warmup):
        fwd_bwd()
    f, bw = [], []
    for _ in range(args.steps):
        a, c, gv, gb = fwd_bwd()
        f.append(a)
        bw.append(c)
    out = {"metric": "differentiable JPCG forward + backward (adjoint JPCG + SDDMM), 1 GPU",
           "workload": f"2D variable-coefficient diffusion 512^2 (n={A.n}, nnz={A.nnz}), exp(GRF) conductivities "
                       f"(length 16, variance 1, seed 4), tol {tol}",
           "steps": args.steps, "warmup": args.warmup,
           "forward_ms": round(float(np.mean(f)), 3), "backward_ms": round(float(np.mean(bw)), 3),
           "grad_values_norm": float(gv.norm()), "grad_b_norm": float(gb.norm())}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=160)
    ap.add_argument("--spmv-repeats", type=int, default=100)
    ap.add_argument("--inner-tol", type=float, default=0.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mode", choices=["full", "spmv-profile", "graph", "batched", "autograd", "batched-profile",
                                       "ingest", "signature"],
                    default="full",
                    help="spmv-profile: one 256 MiB device copy (counter calibration) then 10 SpMV launches, "
                         "for rocprofv3 --pmc runs; graph / batched / autograd: the graph-Laplacian, batched "
                         "and differentiable workloads, one JSON line each")
    ap.add_argument("--graph-n", type=int, default=4_000_000)
    ap.add_argument("--profile-workload", choices=["stencil27", "graph"], default="stencil27")
    ap.add_argument("--batch-count", type=int, default=512)
    ap.add_argument("--nrhs", type=int, default=16)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "one GPU only: the row-block split over several GPUs is designed, not "
                                   "built; run with --gpus 1"}))
        sys.exit(2)
    if args.mode == "spmv-profile":
        spmv_profile(args)
        return
    if args.mode in ("graph", "batched", "autograd", "batched-profile", "ingest", "signature"):
        import torch
        assert torch.cuda.is_available(), "bench.py needs a GPU"
        torch.cuda.set_device(0)
        {"graph": bench_graph, "batched": bench_batched, "autograd": bench_autograd,
         "batched-profile": batched_profile, "ingest": bench_ingest, "signature": bench_signature}[args.mode](args)
        return

    import torch
    from hpc_amd import device, workloads

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    torch.cuda.set_device(0)

    t0 = time.time()
    sysm = workloads.c2_stencil27(args.grid)
    A = sysm.A
    log(f"generated 27-pt {args.grid}^3: n={A.n} nnz={A.nnz} in {time.time() - t0:.1f}s")

    rp = torch.from_numpy(A.rowptr).to(dev)
    col = torch.from_numpy(A.col).to(dev)
    val = torch.from_numpy(A.val).to(dev)
    diag = torch.from_numpy(A.diag).to(dev)
    b = torch.from_numpy(sysm.b).to(dev)
    op = device.CsrOperator(rp, col, val, rowptr_host=A.rowptr)
    # mixed: fp32 off-diagonal values + fp64 split diagonal (exact for this operator)
    srp, scol, sv32, sdiag, exact = device.split_fp32(A.rowptr, A.col, A.val)
    srp_d, scol_d = torch.from_numpy(srp).to(dev), torch.from_numpy(scol).to(dev)
    sv32_d, sdiag_d = torch.from_numpy(sv32).to(dev), torch.from_numpy(sdiag).to(dev)
    inner = device.CsrOperator(srp_d, scol_d, sv32_d, dsplit=sdiag_d, rowptr_host=srp)
    outer = op if not exact else device.CsrOperator(srp_d, scol_d, sv32_d, dsplit=sdiag_d, rowptr_host=srp)
    torch.cuda.synchronize()
    log(f"uploaded; op={op.info()} split_exact={exact}")

    stream = torch.cuda.current_stream()

    # ---- SpMV: per-launch HIP events on the launch stream, every layout that applies; the reported one is
    # the layout the operator chose for its standalone SpMV at creation (its solves use their own choice)
    xs = torch.ones(A.n, dtype=torch.float64, device=dev)
    ys = torch.empty_like(xs)
    nb = spmv_bytes(A.n, A.nnz)
    chosen = op.layout

    def spmv_series(repeats):
        for _ in range(max(3, args.warmup)):
            op.spmv(xs, ys)
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(repeats)]
        for e0, e1 in evs:
            e0.record(stream)
            op.spmv(xs, ys)
            e1.record(stream)
        torch.cuda.synchronize()
        return np.array([e0.elapsed_time(e1) for e0, e1 in evs])

    layouts = {}
    for lay in (1, 3, 4):
        try:
            op.set_layout(lay)
        except Exception:
            continue
        ms = spmv_series(args.spmv_repeats)
        layouts[LAYOUT_NAMES[lay]] = {"median_us": round(float(np.median(ms)) * 1e3, 2),
                                      "GBs": round(nb / (float(np.median(ms)) * 1e-3) / 1e9, 1)}
    op.set_layout(0)
    spmv_ms = spmv_series(args.spmv_repeats)
    spmv_med = float(np.median(spmv_ms))
    spmv_gbs = nb / (spmv_med * 1e-3) / 1e9
    log(f"SpMV fp64 ({LAYOUT_NAMES[chosen]}): median {spmv_med * 1e3:.1f} us  ->  {spmv_gbs:.0f} GB/s; "
        f"per layout {layouts}")
    # fp32-value SpMV of the inner operator, every layout that applies, then its own choice
    def spmv32_time():
        for _ in range(3):
            inner.spmv(xs, ys)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.spmv_repeats):
            inner.spmv(xs, ys)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / args.spmv_repeats

    inner_layouts = {}
    for lay in (1, 3, 4):
        try:
            inner.set_layout(lay)
        except Exception:
            continue
        inner_layouts[LAYOUT_NAMES[lay]] = round(spmv32_time() * 1e3, 2)
    inner.set_layout(0)
    spmv32_ms = spmv32_time()
    log(f"fp32-value SpMV per layout (us): {inner_layouts}, chosen {LAYOUT_NAMES[inner.layout]}")
    nb32 = 8 * inner.nnz + 4 * (A.n + 1) + 24 * A.n
    log(f"SpMV fp32-values+fp64 diag: {spmv32_ms * 1e3:.1f} us  ->  {nb32 / (spmv32_ms * 1e-3) / 1e9:.0f} GB/s")

    copy_gbs = achievable_copy(dev, stream)
    log(f"achievable copy rate: {copy_gbs:.0f} GB/s")

    # ---- JPCG fp64
    x = torch.empty_like(b)
    for _ in range(args.warmup):
        op.solve(diag, b, sysm.max_iter, sysm.tol, x=x)
    torch.cuda.synchronize()
    times, info = [], None
    t_all0 = time.perf_counter()
    for _ in range(args.steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        x, info = op.solve(diag, b, sysm.max_iter, sysm.tol, x=x)
        e1.record(stream)
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    t_fp64_wall = (time.perf_counter() - t_all0) / args.steps
    _, rr = op.residual(b, x)
    bb = float(torch.dot(b, b))
    true_res64 = (rr / bb) ** 0.5
    fp64_ms = float(np.median(times))
    it_bytes = iter_bytes(A.n, A.nnz) * info.executed
    log(f"JPCG fp64 ({LAYOUT_NAMES[op.solve_layout]}): {fp64_ms:.1f} ms, {info.executed} iterations, recursive "
        f"res {info.residual:.2e}, true res {true_res64:.2e}, {it_bytes / (fp64_ms * 1e-3) / 1e9:.0f} GB/s "
        f"(B_iter formula)")

    # ---- JPCG mixed-refined
    xm = torch.empty_like(b)
    for _ in range(args.warmup):
        device.solve_mixed(outer, inner, diag, b, sysm.tol, inner_tol=args.inner_tol, x=xm)
    torch.cuda.synchronize()
    mtimes, minfo = [], None
    for _ in range(args.steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        xm, minfo = device.solve_mixed(outer, inner, diag, b, sysm.tol, inner_tol=args.inner_tol, x=xm)
        e1.record(stream)
        torch.cuda.synchronize()
        mtimes.append(e0.elapsed_time(e1))
    mixed_ms = float(np.median(mtimes))
    _, rrm = op.residual(b, xm)
    true_res_m = (rrm / bb) ** 0.5
    err_vs64 = float(torch.linalg.norm(xm - x) / torch.linalg.norm(x))
    log(f"JPCG mixed: {mixed_ms:.1f} ms, outer {minfo.outer}, inner total {minfo.inner_total}, "
        f"true res {true_res_m:.2e}, |x_mixed - x_fp64|/|x_fp64| = {err_vs64:.2e}")

    cpu = None
    if not args.no_cpu:
        log("CPU baseline (oracle SpMV on host cores) ...")
        res, th = cpu_baseline(A)
        cpu = {"value": round(res[th][0], 2), "unit": "GB/s", "cores": th, "kind": "port",
               "sample": f"oracle CSR SpMV on the same 27-pt 160^3 matrix, {res[th][1]} repeats on {th} threads "
                         f"(1 thread: {res[1][0]:.2f} GB/s, {res[1][1]} repeats); same byte formula; host "
                         f"{cpu_model()}, {os.cpu_count()} logical CPUs visible"}
        log(f"CPU: {cpu}")

    traffic, traffic_src = measured_traffic(args.grid, LAYOUT_NAMES[chosen], op.compact)
    kernel = LAYOUT_KERNELS[chosen] % ("int16-delta" if op.compact else "int32") if chosen == 1 \
        else LAYOUT_KERNELS[chosen]
    out = {
        "metric": "JPCG time-to-1e-10 residual & SpMV HBM GB/s: mixed-refined vs fp64, 1 GPU",
        "value": round(spmv_gbs, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(fp64_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3D 27-point SPD stencil {args.grid}^3 (n={A.n}, nnz={A.nnz}), CSR fp64 values "
                               "int32 indices, b=A*1+U[-1e-3,1e-3] seed 1, tol 1e-10",
                   "parallelism": "single GPU"},
        "spmv": {"repeats": args.spmv_repeats, "median_us": round(spmv_med * 1e3, 2),
                 "mean_us": round(float(spmv_ms.mean()) * 1e3, 2), "bytes_per_launch": nb,
                 "layout": LAYOUT_NAMES[chosen], "per_layout": layouts,
                 "fp32_values_us": round(spmv32_ms * 1e3, 2), "fp32_values_per_layout_us": inner_layouts,
                 "fp32_values_GBs": round(nb32 / (spmv32_ms * 1e-3) / 1e9, 1)},
        "jpcg": {
            "fp64": {"time_to_tol_ms": round(fp64_ms, 3), "wall_ms": round(t_fp64_wall * 1e3, 3),
                     "layout": LAYOUT_NAMES[op.solve_layout],
                     "iterations": info.executed, "recursive_res": info.residual, "true_res": true_res64,
                     "iter_GBs_formula": round(it_bytes / (fp64_ms * 1e-3) / 1e9, 1)},
            "mixed": {"time_to_tol_ms": round(mixed_ms, 3), "outer": minfo.outer, "inner_total": minfo.inner_total,
                      "inner_layout": LAYOUT_NAMES[inner.solve_layout],
                      "inner_tol": args.inner_tol, "true_res": true_res_m, "rel_diff_vs_fp64": err_vs64},
            "speedup_mixed_vs_fp64": round(fp64_ms / mixed_ms, 3),
        },
        "roofline": {"bound": "hbm", "achieved": round(spmv_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(spmv_gbs / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": kernel,
                     "algorithmic_bytes_per_launch": nb,
                     "achievable_copy_GBs": round(copy_gbs, 1),
                     "physical_GBs": round(traffic / (spmv_med * 1e-3) / 1e9, 1) if traffic else None,
                     "physical_frac": round(traffic / (spmv_med * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
                     "bytes_basis": "algorithmic = CSR bytes 12*nnz + 4*(n+1) + 16*n (int32 columns); the "
                                    "diagonal and pattern layouts move fewer bytes than that (no column stream), "
                                    "so frac can reach 1; physical_* use the counter-measured L2 fill bytes",
                     "traffic_source": (traffic_src + " (rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated)")
                     if traffic_src else None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
