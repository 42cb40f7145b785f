"""Benchmark: fp64 SpMV HBM bandwidth and JPCG time-to-tolerance (fp64 vs mixed-refined) on one MI355X.

Headline workload: 3D 27-point SPD operator on a 160^3 grid (n = 4,096,000, nnz = 109,215,352, diag 26,
off-diagonals -1), given to the library as a CSR with fp64 values and int32 indices; b = A*1 +
U[-1e-3, 1e-3] (seed 1); tolerance 1e-10.  Inputs are generated on the host, uploaded, and resident in HBM
before any timed region.

  value       fp64 SpMV rate in GB/s on the bytes the timed layout moves per launch (its matrix streams with
              padding, its per-row tables, x once, y once), over the mean launch duration of 100 launches
              timed with HIP events on the launch stream; the layout is the one the operator chose
  ms_per_step fp64 JPCG time-to-1e-10 (median of the timed solves)
  spmv        every layout's rate on its own bytes and on the CSR formula 12 nnz + 4(n+1) + 16 n (the int32
              row tiles are the input format as given)
  jpcg        fp64 and mixed-refined times, iterations, true residuals
  cpu_baseline the oracle's JPCG (the C restatement of the reference loop) to the same tolerance on the host
  workloads   the other measured workloads: 2D Laplacian 256^2 (GPU vs CPU to 1e-12), the 4M-row graph
              Laplacian (fp64 vs mixed, oracle-checked), 512 systems x 16 right-hand sides (fp64 vs mixed),
              the differentiable solve on 512^2 exp(GRF) diffusion, and refinement with inexact fp32 values

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--grid 160] [--no-cpu] [--no-extra] [--mode ...]
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


def csr_bytes(n, nnz, vbytes=8):
    """SpMV bytes of a CSR with int32 indices: values and columns, row pointers, x once, y once."""
    return (vbytes + 4) * nnz + 4 * (n + 1) + 16 * n


def iter_bytes(n, nnz):
    """JPCG iteration bytes after the reference's stage split (SpMV 12 nnz + 4(n+1) + 24 n, x update 24 n,
    r/z update 40 n, p update 24 n)."""
    return 12 * nnz + 4 * (n + 1) + 112 * n


def cpu_threads():
    """Host threads of this job: OMP_NUM_THREADS when set (the CPU share given to one GPU), else every CPU in
    the affinity mask; and the logical CPUs visible."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    th = min(int(env), aff) if env.isdigit() and int(env) > 0 else aff
    return max(1, th), os.cpu_count() or aff


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


LAYOUT_NAMES = {1: "tiles", 2: "panels", 3: "patterns", 4: "diagonals"}
LAYOUT_KERNELS = {1: "spmv_tiles<%s, %s>", 2: "spmv_panels<%s>", 3: "spmv_pattern<%s>", 4: "spmv_dia<%s, %d, %d>"}
DIA_WIDE_ROWS = 5 << 16        # from this many rows the fp64 diagonal kernel takes two rows per lane
DIA_WIDE_BLOCK_ROWS = 3 << 18  # and from this many, 512-lane workgroups
TRAFFIC_PROFILES = {("stencil27_160", "tiles"): "spmv_tiles_stencil27_160.md",
                    ("stencil27_160", "patterns"): "spmv_pattern_stencil27_160.md",
                    ("stencil27_160", "diagonals"): "spmv_dia_stencil27_160.md",
                    ("graph4m", "panels"): "spmv_panels_graph4m.md"}


def kernel_name(layout, f32, compact, n):
    vt = "float" if f32 else "double"
    if layout == 1:
        return LAYOUT_KERNELS[1] % (vt, "int16_t" if compact else "int")
    if layout == 4:
        two = not f32 and n >= DIA_WIDE_ROWS
        return LAYOUT_KERNELS[4] % (vt, 2 if two else 1, 512 if two and n >= DIA_WIDE_BLOCK_ROWS else 256)
    return LAYOUT_KERNELS[layout] % vt


def measured_traffic(workload, layout_name):
    """HBM bytes per SpMV launch from the committed counter summary of that kernel on that workload."""
    import re
    fname = TRAFFIC_PROFILES.get((workload, layout_name))
    if fname is None:
        return None, None
    path = os.path.join(ROOT, "profiles", fname)
    if not os.path.exists(path):
        return None, None
    m = re.search(r"^traffic_bytes_per_launch:\s*(\d+)", open(path).read(), re.M)
    return (int(m.group(1)), "profiles/" + fname) if m else (None, None)


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


def spmv_launches(op, n, stream, repeats, warmup=3):
    """Per-launch durations (ms) of `repeats` SpMV launches on the operator's current layout, each between
    its own pair of HIP events on the launch stream; x = 1."""
    import torch
    x = torch.ones(n, dtype=torch.float64, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(warmup):
        op.spmv(x, y)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(repeats)]
    for e0, e1 in evs:
        e0.record(stream)
        op.spmv(x, y)
        e1.record(stream)
    torch.cuda.synchronize()
    return np.array([e0.elapsed_time(e1) for e0, e1 in evs])


def layout_rates(op, layouts, stream, repeats, csr_nb):
    """Every applicable layout of `op`: mean / median launch time, its own bytes and rate, the CSR-formula
    rate.  The int32 row tiles (set_compact(False)) are the input CSR as given."""
    out = {}
    for lay in layouts:
        variants = [(lay, None)]
        if lay == 1 and op.compact:
            variants = [(1, True), (1, False)]
        for L, compact in variants:
            try:
                op.set_layout(L)
            except Exception:
                continue
            if compact is not None:
                op.set_compact(compact)
            ms = spmv_launches(op, op.n, stream, repeats)
            nb = op.spmv_bytes(L)
            mean_us, med_us = float(ms.mean()) * 1e3, float(np.median(ms)) * 1e3
            name = LAYOUT_NAMES[L] + ("" if compact is None else ("_int16" if compact else "_int32"))
            out[name] = {"mean_us": round(mean_us, 2), "median_us": round(med_us, 2), "bytes_per_launch": nb,
                         "GBs": round(nb / (mean_us * 1e-6) / 1e9, 1),
                         "frac_of_peak": round(nb / (mean_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
                         "csr_equivalent_GBs": round(csr_nb / (mean_us * 1e-6) / 1e9, 1),
                         "kernel": kernel_name(L, op.f32, bool(compact), op.n)}
            if compact is not None:
                op.set_compact(True)
    op.set_layout(0)
    return out


def achievable_copy(dev, stream, nbytes=2 << 30, reps=10):
    """Achievable HBM rate: the library's stream kernel copying nbytes, (read + write bytes) / time, median."""
    import torch
    from hpc_amd import device
    src = torch.empty(nbytes // 8, dtype=torch.float64, device=dev).fill_(1.0)
    dst = torch.empty_like(src)
    device.stream_copy(src, dst)
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        e0, e1 = _events(stream)
        device.stream_copy(src, dst)
        e1.record(stream)
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1))
    assert bool((dst == 1.0).all())
    del src, dst
    return 2 * nbytes / (float(np.median(ts)) * 1e-3) / 1e9


def cpu_jpcg(A, b, max_iter, tol, threads, reps):
    """The oracle's JPCG on `threads` host threads: (median seconds, last result)."""
    from oracle import oracle
    ts, ref = [], None
    for _ in range(reps):
        t0 = time.perf_counter()
        ref = oracle.jpcg(A.rowptr, A.col, A.val, A.diag, b, max_iter, tol, threads=threads)
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ref


def _split_ops(A, dev):
    """(fp64 operator, fp32-valued inner operator, outer residual operator, exact): the inner operator holds
    the off-diagonal values in fp32 and the diagonal in fp64; the outer one is the split operator when those
    fp32 values equal A's, else the fp64 operator."""
    import torch
    from hpc_amd import device
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    op = device.CsrOperator(t(A.rowptr), t(A.col), t(A.val), rowptr_host=A.rowptr)
    srp, scol, sv32, sdiag, exact = device.split_fp32(A.rowptr, A.col, A.val)
    inner = device.CsrOperator(t(srp), t(scol), t(sv32), dsplit=t(sdiag), rowptr_host=srp)
    outer = device.CsrOperator(t(srp), t(scol), t(sv32), dsplit=t(sdiag), rowptr_host=srp) if exact else op
    return op, inner, outer, exact


# ---------------------------------------------------------------------------------------------------------
# workloads besides the headline


def bench_laplace2d(args):
    """2D 5-point Laplacian 256^2 (x_true U[-1,1] seed 0, b = A x_true, tol 1e-12, at most 5000 iterations):
    GPU JPCG (multi-kernel loop, persistent loop, allocation-free entry point) against the oracle's JPCG on one
    host thread and on all of the job's threads."""
    import torch
    from hpc_amd import device, workloads
    dev = torch.device("cuda:0")
    sysm = workloads.c1_laplace2d(256)
    A = sysm.A
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    rp, col, val = t(A.rowptr), t(A.col), t(A.val)
    op = device.CsrOperator(rp, col, val, rowptr_host=A.rowptr)
    diag, b = t(A.diag), t(sysm.b)
    stream = torch.cuda.current_stream()
    steps, warm = max(3, min(args.steps, 10)), max(1, min(args.warmup, 3))
    out = {"workload": f"2D 5-point Laplacian 256^2 (n={A.n}, nnz={A.nnz}), b = A x_true (x_true U[-1,1] seed 0), "
                       f"tol 1e-12, at most 5000 iterations"}
    x = torch.empty_like(b)
    for path, name in ((device.CsrOperator.PATH_MULTI, "multi_kernel"), (device.CsrOperator.PATH_PERSISTENT,
                                                                         "persistent")):
        op.set_path(path)
        ms, (x, info) = _timed(lambda: op.solve(diag, b, sysm.max_iter, sysm.tol, x=x), steps, warm, stream)
        out[f"gpu_{name}"] = {"ms": round(float(np.median(ms)), 3), "iterations": info.executed,
                              "us_per_iteration": round(float(np.median(ms)) * 1e3 / max(1, info.executed), 2),
                              "loop_used": {1: "persistent", 2: "multi_kernel"}.get(op.last_path, op.last_path)}
    op.set_path(device.CsrOperator.PATH_AUTO)
    xs = torch.empty_like(b)
    ms, (xs, sinfo) = _timed(lambda: device.csr_solve_static(rp, col, val, diag, b, sysm.max_iter, sysm.tol, x=xs),
                             steps, warm, stream)
    out["gpu_allocation_free"] = {"ms": round(float(np.median(ms)), 3), "iterations": sinfo.executed}
    x, info = op.solve(diag, b, sysm.max_iter, sysm.tol, x=x)
    xh = x.cpu().numpy()
    th, visible = cpu_threads()
    s1, ref = cpu_jpcg(A, sysm.b, sysm.max_iter, sysm.tol, 1, 3)
    sa, _ = cpu_jpcg(A, sysm.b, sysm.max_iter, sysm.tol, th, 3)
    out["rel_err_vs_oracle"] = float(np.linalg.norm(xh - ref["x"]) / np.linalg.norm(ref["x"]))
    out["oracle_iterations"] = ref["executed"]
    out["cpu_oracle_jpcg"] = {"ms_1_thread": round(s1 * 1e3, 2), "ms_all": round(sa * 1e3, 2), "threads": th,
                              "logical_cpus_visible": visible, "runs": "median of 3"}
    best = min(out["gpu_multi_kernel"]["ms"], out["gpu_persistent"]["ms"])
    out["gpu_speedup_vs_cpu_all"] = round(sa * 1e3 / best, 2)
    log(f"laplace2d 256^2: {out}")
    return out


def bench_graph(args):
    """Graph Laplacian + 1e-2 I (lognormal degrees mu 2.5 sigma 1 capped at 2000, weights U[0.1,1] in fp32,
    fp64 diagonal), n = 4M: fp64 JPCG vs the fp32-valued refined solve to 1e-10, both checked against the
    oracle's JPCG on the host cores; SpMV per layout on each layout's own bytes."""
    import torch
    from hpc_amd import device, workloads
    dev = torch.device("cuda:0")
    t0 = time.time()
    sysm = workloads.c3_graph(args.graph_n)
    A = sysm.A
    deg = np.diff(A.rowptr)
    log(f"graph n={A.n} nnz={A.nnz} max row {deg.max()} generated in {time.time() - t0:.1f}s")
    op, inner, outer, exact = _split_ops(A, dev)
    diag, b = torch.from_numpy(A.diag).to(dev), torch.from_numpy(sysm.b).to(dev)
    stream = torch.cuda.current_stream()
    chosen = {"fp64": LAYOUT_NAMES[op.layout], "fp32_values": LAYOUT_NAMES[inner.layout],
              "fp64_solve": LAYOUT_NAMES[op.solve_layout], "fp32_values_solve": LAYOUT_NAMES[inner.solve_layout]}
    sp64 = layout_rates(op, (1, 2), stream, args.spmv_repeats, csr_bytes(A.n, A.nnz))
    sp32 = layout_rates(inner, (1, 2), stream, args.spmv_repeats, csr_bytes(A.n, inner.nnz, 4) + 8 * A.n)
    log(f"graph SpMV fp64 {sp64} fp32-values {sp32} layouts {chosen}")
    steps, warm = max(1, min(args.steps, 5)), max(1, min(args.warmup, 2))
    x = torch.empty_like(b)
    ms64, (x, info) = _timed(lambda: op.solve(diag, b, sysm.max_iter, sysm.tol, x=x), steps, warm, stream)
    xm = torch.empty_like(b)
    msm, (xm, minfo) = _timed(lambda: device.solve_mixed(outer, inner, diag, b, sysm.tol, inner_tol=args.inner_tol,
                                                         x=xm), steps, warm, stream)
    bb = float(torch.dot(b, b))
    true64 = (op.residual(b, x)[1] / bb) ** 0.5
    truem = (op.residual(b, xm)[1] / bb) ** 0.5
    xmh = xm.cpu().numpy()
    # fp32 work vectors in the inner operator (panel layout): p gathered in fp32
    mixed32 = None
    if inner.solve_layout in (2, 4):
        inner.set_vectors(32)
        msw, (xm, winfo) = _timed(lambda: device.solve_mixed(outer, inner, diag, b, sysm.tol,
                                                             inner_tol=args.inner_tol, x=xm), steps, warm, stream)
        inner.set_vectors(64)
        mixed32 = {"ms": round(float(np.median(msw)), 3), "outer": winfo.outer, "inner_total": winfo.inner_total,
                   "true_res": (op.residual(b, xm)[1] / bb) ** 0.5,
                   "speedup_vs_fp64": round(float(np.median(ms64) / np.median(msw)), 3)}
    th, visible = cpu_threads()
    log(f"graph: fp64 {np.median(ms64):.2f} ms ({info.executed} it), mixed {np.median(msm):.2f} ms, with fp32 "
        f"work vectors {mixed32['ms'] if mixed32 else None} ms; oracle JPCG on {th} host threads ...")
    cpu_s, ref = cpu_jpcg(A, sysm.b, sysm.max_iter, sysm.tol, th, 1)
    xr = ref["x"]
    nr = np.linalg.norm(xr)
    out = {"workload": f"graph Laplacian + 1e-2 I, n={A.n}, nnz={A.nnz}, degrees lognormal(2.5, 1.0) cap 2000 "
                       f"(max row {int(deg.max())}), weights U[0.1,1] (fp32-exact), b U[-1,1] seed 2, tol 1e-10",
           "steps": steps, "split_exact": exact, "layouts": chosen,
           "spmv": {"fp64": sp64, "fp32_values": sp32},
           "fp64": {"ms": round(float(np.median(ms64)), 3), "iterations": info.executed, "true_res": true64,
                    "rel_err_vs_oracle": float(np.linalg.norm(x.cpu().numpy() - xr) / nr)},
           "mixed": {"ms": round(float(np.median(msm)), 3), "outer": minfo.outer, "inner_total": minfo.inner_total,
                     "true_res": truem, "rel_err_vs_oracle": float(np.linalg.norm(xmh - xr) / nr)},
           "speedup_mixed_vs_fp64": round(float(np.median(ms64) / np.median(msm)), 3),
           "mixed_fp32_vectors": dict(mixed32, rel_err_vs_oracle=float(np.linalg.norm(xm.cpu().numpy() - xr) / nr))
           if mixed32 else None,
           "cpu_oracle_jpcg": {"iterations": ref["executed"], "ms_all": round(cpu_s * 1e3, 1), "threads": th,
                               "logical_cpus_visible": visible}}
    out["gpu_speedup_vs_cpu_all"] = round(cpu_s * 1e3 / out["fp64"]["ms"], 1)
    log(f"graph: {out['fp64']} {out['mixed']} cpu {out['cpu_oracle_jpcg']}")
    return out


def batch_bytes(mats, its, launched_sys):
    """Algorithmic bytes of a batched solve after the reference's stage split: each system's matrix
    (12 nnz + 4(n+1)) once per batched iteration in which it ran, and 112 n vector bytes per (system, rhs)
    iteration."""
    nb = 0
    for s, A in enumerate(mats):
        nb += int(launched_sys[s]) * (12 * A.nnz + 4 * (A.n + 1)) + int(its[s].sum()) * 112 * A.n
    return nb


def bench_batched(args):
    """512 anisotropic 2D/3D diffusion systems (n log-uniform in [5k, 400k]) x 16 right-hand sides, tol 1e-8:
    batched fp64 JPCG vs the fp32-valued refined batch, with per-pair checks against the oracle on a sample
    that includes the largest system."""
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
    if args.tune:
        t0 = time.time()
        B64.configure(tune=True)
        inner.configure(tune=True)
        log(f"autotuned in {time.time() - t0:.1f}s")
    tol, maxit = 1e-8, 100000
    steps, warm = max(1, min(args.steps, args.batch_steps)), 0
    x = torch.empty_like(b)
    ms64, (x, info) = _timed(lambda: B64.solve(diag, b, maxit, tol, x=x), steps, warm, stream)
    its = info.iterations.cpu().numpy()
    log(f"fp64 batch: {np.median(ms64):.1f} ms, {info.launched} launched iterations, pair iterations "
        f"{its.min()}..{its.max()}, max res {float(info.residual.max()):.2e}")
    # the same solve as one persistent cooperative launch (same bits)
    B64.set_path(1)
    xp = torch.empty_like(b)
    msp, (xp, pinfo) = _timed(lambda: B64.solve(diag, b, maxit, tol, x=xp), 1, 0, stream)
    persistent = {"ms": round(float(np.median(msp)), 2), "loop": B64.last_path[0], "grid": B64.last_path[1],
                  "iterations": pinfo.launched,
                  "same_bits_as_launches": bool((xp == x).all()) and bool((pinfo.iterations == info.iterations).all())}
    B64.set_path(0)
    del xp
    log(f"fp64 batch, one persistent launch: {persistent}")
    xm = torch.empty_like(b)
    msm, (xm, minfo, steps_m) = _timed(lambda: batched.solve_mixed(outer, inner, diag, b, tol,
                                                                   inner_tol=args.inner_tol, x=xm),
                                       steps, warm, stream)
    log(f"mixed batch: {np.median(msm):.1f} ms, outer {steps_m}, launched {minfo.launched}, max true res "
        f"{float(minfo.residual.max()):.2e}")
    # the inner batch with fp32 work vectors too (fp64 arithmetic, dots and outer loop)
    inner32 = batched.Batch.from_host(asm32, R, dev)
    inner32.set_vectors(32)
    if args.tune:
        inner32.configure(tune=True)
    xw = torch.empty_like(b)
    msw, (xw, winfo, steps_w) = _timed(lambda: batched.solve_mixed(outer, inner32, diag, b, tol,
                                                                   inner_tol=args.inner_tol, x=xw),
                                       steps, warm, stream)
    log(f"mixed batch, fp32 work vectors: {np.median(msw):.1f} ms, outer {steps_w}, launched {winfo.launched}, max "
        f"true res {float(winfo.residual.max()):.2e}")
    del inner32
    sizes = np.diff(asm64["sys_row"])
    order = np.argsort(sizes, kind="stable")
    picks = [int(order[-1]), int(order[0]), int(order[len(order) // 2])]
    th, _ = cpu_threads()
    xh, xmh, xwh = x.cpu().numpy(), xm.cpu().numpy(), xw.cpu().numpy()
    checks = []
    for s in picks:
        r0, r1 = int(asm64["sys_row"][s]), int(asm64["sys_row"][s + 1])
        A = mats[s]
        ref = oracle.jpcg(A.rowptr, A.col, A.val, A.diag, bh[r0:r1, 0].copy(), maxit, tol, threads=th)
        nr = np.linalg.norm(ref["x"])
        checks.append({"system": s, "n": A.n, "rhs": 0, "oracle_iterations": ref["executed"],
                       "iterations": int(its[s, 0]),
                       "fp64_rel_err": float(np.linalg.norm(xh[r0:r1, 0] - ref["x"]) / nr),
                       "mixed_rel_err": float(np.linalg.norm(xmh[r0:r1, 0] - ref["x"]) / nr),
                       "mixed_fp32_vectors_rel_err": float(np.linalg.norm(xwh[r0:r1, 0] - ref["x"]) / nr)})
    # the batch runs a system while any of its pairs runs: its matrix moves max_j(its) times
    nb = batch_bytes(mats, its, its.max(axis=1))
    t64 = float(np.median(ms64))
    # the fp32-vector refined batch, inner solves only: the split operator (fp32 off-diagonal values, int32
    # columns, row pointers, fp64 diagonal) per launched iteration, and the stage split at the stored widths
    # per (system, rhs) iteration -- SpMV 12 n, x update 20 n (x fp64), r/z 20 n, p 12 n = 64 n
    itw = winfo.iterations.cpu().numpy().astype(np.int64)
    nbw = sum(int(itw[s_].max()) * (8 * (A.nnz - A.n) + 4 * (A.n + 1) + 8 * A.n) + int(itw[s_].sum()) * 64 * A.n
              for s_, A in enumerate(mats))
    tw = float(np.median(msw))
    out = {"workload": f"{len(mats)} anisotropic 2D5/3D7 diffusion systems, n log-uniform [5k,400k] (N={N}, "
                       f"nnz={asm64['col'].size}), {R} rhs U[-1,1] seed 3, tol 1e-8",
           "steps": steps, "autotuned": bool(args.tune),
           "fp64": {"ms": round(t64, 2), "launched_iterations": info.launched,
                    "pair_iterations_max": int(its.max()), "pair_iterations_sum": int(its.sum()),
                    "max_res": float(info.residual.max()), "config": B64.info(),
                    "algorithmic_bytes": nb, "GBs": round(nb / (t64 * 1e-3) / 1e9, 1),
                    "frac_of_peak": round(nb / (t64 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
           "mixed": {"ms": round(float(np.median(msm)), 2), "outer": steps_m, "launched_iterations": minfo.launched,
                     "inner_tol": args.inner_tol, "max_true_res": float(minfo.residual.max()),
                     "config": inner.info()},
           "fp64_one_persistent_launch": persistent,
           "speedup_mixed_vs_fp64": round(t64 / float(np.median(msm)), 3),
           "mixed_fp32_vectors": {"ms": round(tw, 2), "outer": steps_w,
                                  "launched_iterations": winfo.launched,
                                  "max_true_res": float(winfo.residual.max()),
                                  "speedup_vs_fp64": round(t64 / tw, 3),
                                  "inner_pair_iterations_sum": int(itw.sum()),
                                  "algorithmic_bytes": nbw, "GBs": round(nbw / (tw * 1e-3) / 1e9, 1),
                                  "frac_of_peak": round(nbw / (tw * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                                  "bytes_basis": "inner solves only, a lower bound: per system and launched "
                                                 "iteration (its busiest pair's total) 8 (nnz - n) + 4 (n+1) + "
                                                 "8 n; per (system, rhs) inner iteration 64 n (SpMV 12 n, "
                                                 "x update 20 n with x in fp64, r/z 20 n, p 12 n); the outer "
                                                 "fp64 residual steps are not counted"},
           "bytes_basis": "per system and batched iteration it runs: 12 nnz + 4(n+1); per (system, rhs) "
                          "iteration: 112 n (the reference's SpMV / x / r-z / p stage split)",
           "checks": checks}
    log(f"batched: {out['fp64']['ms']} ms fp64, {out['mixed']['ms']} ms mixed, checks {checks}")
    return out


def bench_shard(args):
    """One GPU's share of the 160^3 27-point stencil split by contiguous row blocks over P = 2, 4, 8 GPUs: a
    160 x 160 x (160/P) slab (Dirichlet faces where the neighbours would be), the SpMV on the layout the operator
    picks, timed on its own bytes, and the JPCG iteration (graph-replayed multi-kernel loop, 200 iterations);
    beside them the p halo a rank would exchange per iteration (two 160^2 planes, fp64) and its time on one
    xGMI link at 153 GB/s.  The transport and the two all-reduces per iteration are not measured."""
    import torch
    from hpc_amd import device, workloads
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    m = args.grid
    out = {"workload": f"row-block share of the 27-point stencil {m}^3 for P GPUs: {m} x {m} x {m}/P slab",
           "shares": {}}
    for P in (2, 4, 8):
        nz = m // P
        A = workloads.stencil(workloads.STENCIL_3D27, m, m, nz, [26.0, -1.0])
        op = device.CsrOperator(t(A.rowptr), t(A.col), t(A.val), rowptr_host=A.rowptr)
        ms = spmv_launches(op, A.n, stream, args.spmv_repeats)
        nb = op.spmv_bytes()
        mean_us = float(ms.mean()) * 1e3
        diag = t(A.diag)
        b = t(np.ones(A.n))
        x = torch.empty_like(b)
        it_ms, _ = _timed(lambda: op.solve(diag, b, 199, 1e-300, x=x), 3, 1, stream)
        halo = 2 * m * m * 8
        out["shares"][f"P{P}"] = {
            "rows": A.n, "nnz": A.nnz, "layout": LAYOUT_NAMES[op.layout],
            "spmv_us": round(mean_us, 2), "spmv_bytes": nb, "spmv_GBs": round(nb / (mean_us * 1e-6) / 1e9, 1),
            "spmv_frac_of_peak": round(nb / (mean_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
            "jpcg_us_per_iteration": round(float(np.median(it_ms)) * 1e3 / 200, 2),
            "halo_bytes_per_iteration": halo, "halo_us_one_xgmi_link": round(halo / 153e9 * 1e6, 2)}
        log(f"shard P={P}: {out['shares'][f'P{P}']}")
        del op
        torch.cuda.empty_cache()
    return out


def bench_autograd(args):
    """Differentiable solve on 2D variable-coefficient diffusion 512^2: forward JPCG, then backward
    (adjoint JPCG + SDDMM) for L = 0.5 ||x - x_target||^2, fp64, tol 1e-12."""
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

    steps, warm = max(1, min(args.steps, 5)), max(1, min(args.warmup, 2))
    for _ in range(warm):
        fwd_bwd()
    f, bw = [], []
    for _ in range(steps):
        a, c, gv, gb = fwd_bwd()
        f.append(a)
        bw.append(c)
    # the gradient kernel on its own: g_k = -lambda_row(k) * x_col(k) over the pattern, 100 launches
    from hpc_amd import device
    op = device.CsrOperator(torch.from_numpy(A.rowptr).to(dev), torch.from_numpy(A.col).to(dev), base,
                            rowptr_host=A.rowptr)
    lam = torch.from_numpy(workloads.uniform(4, 3, A.n, -1, 1)).to(dev)
    xs = torch.from_numpy(workloads.uniform(4, 4, A.n, -1, 1)).to(dev)
    g = torch.empty(A.nnz, dtype=torch.float64, device=dev)
    sd, _ = _timed(lambda: op.sddmm(lam, xs, g), 100, 5, stream)
    sd_us = float(np.median(sd)) * 1e3
    sd_bytes = 12 * A.nnz + 4 * (A.n + 1) + 16 * A.n
    del op
    out = {"workload": f"2D variable-coefficient diffusion 512^2 (n={A.n}, nnz={A.nnz}), exp(GRF) conductivities "
                       f"(length 16, variance 1, seed 4), tol {tol}",
           "steps": steps, "forward_ms": round(float(np.median(f)), 3), "backward_ms": round(float(np.median(bw)), 3),
           "sddmm": {"median_us": round(sd_us, 2), "bytes": sd_bytes,
                     "GBs": round(sd_bytes / (sd_us * 1e-6) / 1e9, 1),
                     "bytes_basis": "12 nnz + 4(n+1) + 16 n (row pointers, columns, gradient out, lambda and x); "
                                    "21 MB stays in the 256 MB MALL, so the rate is not an HBM figure"},
           "grad_values_norm": float(gv.norm()), "grad_b_norm": float(gb.norm())}
    log(f"autograd: {out}")
    return out


def bench_inexact_mixed(args):
    """Refinement where the fp32 values are not A: the 512^2 exp(GRF) diffusion operator (conductivities not
    representable in fp32), tol 1e-10: fp64 JPCG vs the refined solve (fp32 off-diagonal values, fp64 split
    diagonal, fp64 residual through the fp64 operator), errors against the oracle's fp64 JPCG."""
    import torch
    from hpc_amd import device, workloads
    dev = torch.device("cuda:0")
    A, bh, _ = workloads.c5_diffusion()
    op, inner, outer, exact = _split_ops(A, dev)
    diag, b = torch.from_numpy(A.diag).to(dev), torch.from_numpy(bh).to(dev)
    stream = torch.cuda.current_stream()
    tol, maxit = 1e-10, 200000
    steps, warm = max(1, min(args.steps, 5)), 1
    x = torch.empty_like(b)
    ms64, (x, info) = _timed(lambda: op.solve(diag, b, maxit, tol, x=x), steps, warm, stream)
    res = {}
    for itol in (0.0, 1e-3):
        xm = torch.empty_like(b)
        msm, (xm, minfo) = _timed(lambda: device.solve_mixed(outer, inner, diag, b, tol, inner_tol=itol, x=xm,
                                                             inner_max_iter=maxit), steps, warm, stream)
        res[itol] = (float(np.median(msm)), minfo, xm.cpu().numpy())
    th, _ = cpu_threads()
    _, ref = cpu_jpcg(A, bh, maxit, tol, th, 1)
    nr = np.linalg.norm(ref["x"])
    out = {"workload": f"2D exp(GRF) diffusion 512^2 (n={A.n}, nnz={A.nnz}), b U[-1,1] seed 4, tol 1e-10; fp32 "
                       f"values exact: {exact}",
           "fp64": {"ms": round(float(np.median(ms64)), 3), "iterations": info.executed,
                    "rel_err_vs_oracle": float(np.linalg.norm(x.cpu().numpy() - ref["x"]) / nr),
                    "oracle_iterations": ref["executed"]}}
    for itol, (ms, minfo, xmh) in res.items():
        out[f"mixed_inner_tol_{itol:g}"] = {"ms": round(ms, 3), "outer": minfo.outer, "inner_total": minfo.inner_total,
                                           "true_res": minfo.residual,
                                           "rel_err_vs_oracle": float(np.linalg.norm(xmh - ref["x"]) / nr),
                                           "speedup_vs_fp64": round(float(np.median(ms64)) / ms, 3)}
    log(f"inexact mixed: {out}")
    return out


# ---------------------------------------------------------------------------------------------------------
# short runs for profilers


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
    if args.profile_layout:
        op.set_layout(args.profile_layout)
    src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    device.stream_copy(src, dst)
    x = torch.ones(A.n, dtype=torch.float64, device=dev)
    y = torch.empty_like(x)
    for _ in range(10):
        op.spmv(x, y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_each_way": src.numel() * 8, "layout_bytes": op.spmv_bytes(),
                      "csr_bytes": csr_bytes(A.n, A.nnz), "layout": LAYOUT_NAMES[op.layout],
                      "compact_columns": op.compact, "n": A.n, "nnz": A.nnz}), flush=True)


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
    B.set_vectors(args.batch_vectors)
    x, info = B.solve(diag, b, 9, 1e-30)
    torch.cuda.synchronize()
    print(json.dumps({"N": B.N, "nnz": int(asm["col"].size), "R": args.nrhs, "launched": info.launched,
                      "vector_bits": B.vector_bits, "config": B.info()}), flush=True)


def bench_ingest(args):
    """One-shot pcg.h solve from a shuffled COO of the 27-point operator (--grid, default 160^3): the
    reference's metric phases (matrix processing, vectors, solver) with the CSR built on the GPU, beside the
    oracle's host COO-to-CSR on the same arrays."""
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
    return {"metric": "one-shot xJPCG_cooSolver phases, CSR built on the GPU",
            "workload": f"27-point stencil {args.grid}^3 as a shuffled COO (n={A.n}, nnz={A.nnz})",
            **out, "oracle_coo_to_csr_s": round(cpu, 3)}


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
    return {"metric": "signature files to device CSR",
            "workload": f"27-point stencil {args.grid}^3 (n={A.n}, nnz={A.nnz})",
            "nnz_pad": info["nnz_pad"], "row_blocks": info["row_blocks"],
            "partitions": info["partitions"], "stream_bytes": sbytes, **out,
            "sample_40cubed": sample}


# ---------------------------------------------------------------------------------------------------------
# headline


def bench_headline(args):
    import torch
    from hpc_amd import device, workloads
    dev = torch.device("cuda:0")
    t0 = time.time()
    sysm = workloads.c2_stencil27(args.grid)
    A = sysm.A
    log(f"generated 27-pt {args.grid}^3: n={A.n} nnz={A.nnz} in {time.time() - t0:.1f}s")
    op, inner, outer, exact = _split_ops(A, dev)
    diag = torch.from_numpy(A.diag).to(dev)
    b = torch.from_numpy(sysm.b).to(dev)
    torch.cuda.synchronize()
    log(f"uploaded; op={op.info()} split_exact={exact}")
    stream = torch.cuda.current_stream()
    csr_nb = csr_bytes(A.n, A.nnz)

    # ---- SpMV: every layout on its own bytes; the reported one is the layout the operator chose at creation
    chosen = op.layout
    layouts = layout_rates(op, (1, 3, 4), stream, args.spmv_repeats, csr_nb)
    ms = spmv_launches(op, A.n, stream, args.spmv_repeats)
    nb = op.spmv_bytes()
    mean_us, med_us = float(ms.mean()) * 1e3, float(np.median(ms)) * 1e3
    spmv_gbs = nb / (mean_us * 1e-6) / 1e9
    log(f"SpMV fp64 ({LAYOUT_NAMES[chosen]}): mean {mean_us:.1f} us on {nb} bytes -> {spmv_gbs:.0f} GB/s; "
        f"per layout {layouts}")
    inner_layouts = layout_rates(inner, (1, 3, 4), stream, args.spmv_repeats, csr_bytes(A.n, inner.nnz, 4) + 8 * A.n)
    log(f"fp32-value SpMV per layout: {inner_layouts}, chosen {LAYOUT_NAMES[inner.layout]}")
    copy_gbs = achievable_copy(dev, stream)
    log(f"achievable copy rate: {copy_gbs:.0f} GB/s")

    # ---- JPCG fp64
    x = torch.empty_like(b)
    times, (x, info) = _timed(lambda: op.solve(diag, b, sysm.max_iter, sysm.tol, x=x), args.steps, args.warmup,
                              stream)
    bb = float(torch.dot(b, b))
    true_res64 = (op.residual(b, x)[1] / bb) ** 0.5
    fp64_ms = float(np.median(times))
    it_bytes = iter_bytes(A.n, A.nnz) * info.executed
    log(f"JPCG fp64 ({LAYOUT_NAMES[op.solve_layout]}): {fp64_ms:.1f} ms, {info.executed} iterations, recursive "
        f"res {info.residual:.2e}, true res {true_res64:.2e}")
    # ---- JPCG mixed-refined
    xm = torch.empty_like(b)
    mtimes, (xm, minfo) = _timed(lambda: device.solve_mixed(outer, inner, diag, b, sysm.tol, inner_tol=args.inner_tol,
                                                            x=xm), args.steps, args.warmup, stream)
    mixed_ms = float(np.median(mtimes))
    true_res_m = (op.residual(b, xm)[1] / bb) ** 0.5
    err_vs64 = float(torch.linalg.norm(xm - x) / torch.linalg.norm(x))
    log(f"JPCG mixed: {mixed_ms:.1f} ms, outer {minfo.outer}, inner total {minfo.inner_total}, "
        f"true res {true_res_m:.2e}, |x_mixed - x_fp64|/|x_fp64| = {err_vs64:.2e}")
    # ---- JPCG mixed-refined with fp32 work vectors in the inner operator (diagonal solve layout only)
    mixed32 = None
    if inner.solve_layout == 4:
        inner.set_vectors(32)
        wtimes, (xm, winfo) = _timed(lambda: device.solve_mixed(outer, inner, diag, b, sysm.tol,
                                                                inner_tol=args.inner_tol, x=xm),
                                     args.steps, args.warmup, stream)
        inner.set_vectors(64)
        w_ms = float(np.median(wtimes))
        true_res_w = (op.residual(b, xm)[1] / bb) ** 0.5
        err_w = float(torch.linalg.norm(xm - x) / torch.linalg.norm(x))
        log(f"JPCG mixed, fp32 work vectors: {w_ms:.1f} ms, outer {winfo.outer}, inner total {winfo.inner_total}, "
            f"true res {true_res_w:.2e}, |x - x_fp64|/|x_fp64| = {err_w:.2e}")
        mixed32 = {"time_to_tol_ms": round(w_ms, 3), "outer": winfo.outer, "inner_total": winfo.inner_total,
                   "inner_layout": LAYOUT_NAMES[inner.solve_layout], "inner_vectors": "r, p, Ap, J fp32; x, dots fp64",
                   "true_res": true_res_w, "rel_diff_vs_fp64": err_w,
                   "speedup_vs_fp64": round(fp64_ms / w_ms, 3)}
    xh = x.cpu().numpy()

    cpu = None
    if not args.no_cpu:
        th, visible = cpu_threads()
        log(f"CPU baseline: oracle JPCG to {sysm.tol:g} on {th} threads (median of 3), then 1 thread ...")
        s_all, ref = cpu_jpcg(A, sysm.b, sysm.max_iter, sysm.tol, th, 3)
        s_1, _ = cpu_jpcg(A, sysm.b, sysm.max_iter, sysm.tol, 1, 1)
        e = float(np.linalg.norm(xh - ref["x"]) / np.linalg.norm(ref["x"]))
        cpu = {"value": round(s_all * 1e3, 1), "unit": "ms (JPCG time to 1e-10)", "cores": th, "kind": "port",
               "sample": f"the oracle's JPCG (C restatement of the reference loop, CSR fp64/int32, OpenMP) on the "
                         f"same 27-pt {args.grid}^3 system to 1e-10: {ref['executed']} iterations, median of 3 runs "
                         f"on {th} threads (the job's CPU share; {visible} logical CPUs visible; host "
                         f"{cpu_model()}); 1 thread, one run: {s_1 * 1e3:.0f} ms",
               "ms_1_thread": round(s_1 * 1e3, 1), "iterations": ref["executed"],
               "gpu_fp64_speedup": round(s_all * 1e3 / fp64_ms, 1),
               "gpu_mixed_speedup": round(s_all * 1e3 / mixed_ms, 1),
               "gpu_mixed_fp32_vectors_speedup": round(s_all / mixed32["time_to_tol_ms"] * 1e3, 1) if mixed32 else None,
               "gpu_rel_err_vs_oracle": e}
        log(f"CPU: {cpu}")

    wl = f"stencil27_{args.grid}"
    lname = LAYOUT_NAMES[chosen]
    traffic, traffic_src = measured_traffic(wl, lname)
    kname = kernel_name(chosen, False, op.compact, op.n)
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
        "config": {"workload": f"3D 27-point SPD stencil {args.grid}^3 (n={A.n}, nnz={A.nnz}) given as CSR fp64 "
                               f"values + int32 indices; SpMV timed on the layout the operator chose: {lname} "
                               f"({kname}); b=A*1+U[-1e-3,1e-3] seed 1, tol 1e-10",
                   "parallelism": "single GPU"},
        "spmv": {"repeats": args.spmv_repeats, "layout": lname, "kernel": kname, "mean_us": round(mean_us, 2),
                 "median_us": round(med_us, 2), "bytes_per_launch": nb, "csr_bytes_per_launch": csr_nb,
                 "csr_equivalent_GBs": round(csr_nb / (mean_us * 1e-6) / 1e9, 1), "per_layout": layouts,
                 "fp32_values_layout": LAYOUT_NAMES[inner.layout], "fp32_values_per_layout": inner_layouts},
        "jpcg": {
            "fp64": {"time_to_tol_ms": round(fp64_ms, 3), "layout": LAYOUT_NAMES[op.solve_layout],
                     "iterations": info.executed, "recursive_res": info.residual, "true_res": true_res64,
                     "iter_GBs_reference_stage_bytes": round(it_bytes / (fp64_ms * 1e-3) / 1e9, 1)},
            "mixed": {"time_to_tol_ms": round(mixed_ms, 3), "outer": minfo.outer, "inner_total": minfo.inner_total,
                      "inner_layout": LAYOUT_NAMES[inner.solve_layout], "inner_tol": args.inner_tol,
                      "true_res": true_res_m, "rel_diff_vs_fp64": err_vs64},
            "speedup_mixed_vs_fp64": round(fp64_ms / mixed_ms, 3),
            "mixed_fp32_vectors": mixed32,
        },
        "roofline": {"bound": "hbm", "achieved": round(spmv_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(spmv_gbs / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kname,
                     "algorithmic_bytes_per_launch": nb, "avg_launch_us": round(mean_us, 2),
                     "bytes_basis": "the bytes the timed layout must move per launch: its matrix streams with "
                                    "their zero padding, its per-row tables, x once and y once "
                                    "(xjpcg_op_spmv_bytes)",
                     "counter_GBs": round(traffic / (mean_us * 1e-6) / 1e9, 1) if traffic else None,
                     "counter_frac": round(traffic / (mean_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
                     "csr_equivalent_GBs": round(csr_nb / (mean_us * 1e-6) / 1e9, 1),
                     "achievable_copy_GBs": round(copy_gbs, 1),
                     "traffic_source": (traffic_src + " (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)")
                     if traffic_src else None},
        "cpu_baseline": cpu,
    }
    del op, inner, outer, x, xm, b, diag
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=160)
    ap.add_argument("--spmv-repeats", type=int, default=100)
    ap.add_argument("--inner-tol", type=float, default=0.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extra", action="store_true", help="headline workload only")
    ap.add_argument("--tune", action="store_true", help="autotune the batched solves first")
    ap.add_argument("--batch-steps", type=int, default=1, help="timed batched solves (each ~10 s)")
    ap.add_argument("--mode", choices=["full", "spmv-profile", "laplace2d", "graph", "batched", "autograd",
                                       "inexact-mixed", "batched-profile", "ingest", "signature", "shard"],
                    default="full",
                    help="spmv-profile: one 256 MiB device copy (counter calibration) then 10 SpMV launches, "
                         "for rocprofv3 --pmc runs; laplace2d / graph / batched / autograd / inexact-mixed: one "
                         "workload, one JSON line")
    ap.add_argument("--graph-n", type=int, default=4_000_000)
    ap.add_argument("--profile-workload", choices=["stencil27", "graph"], default="stencil27")
    ap.add_argument("--profile-layout", type=int, default=0)
    ap.add_argument("--batch-count", type=int, default=512)
    ap.add_argument("--nrhs", type=int, default=16)
    ap.add_argument("--batch-vectors", type=int, default=64, choices=[32, 64],
                    help="batched-profile: work vectors r, p, Ap in fp64 or fp32")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "one GPU only: the row-block split over several GPUs is designed, not "
                                   "built; run with --gpus 1"}))
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    if args.mode == "spmv-profile":
        spmv_profile(args)
        return
    if args.mode == "batched-profile":
        batched_profile(args)
        return
    single = {"laplace2d": bench_laplace2d, "graph": bench_graph, "batched": bench_batched,
              "autograd": bench_autograd, "inexact-mixed": bench_inexact_mixed, "ingest": bench_ingest,
              "signature": bench_signature, "shard": bench_shard}
    if args.mode in single:
        print(json.dumps(single[args.mode](args)), flush=True)
        return
    out = bench_headline(args)
    if not args.no_extra:
        extra = {}
        for name, fn in (("laplace2d_256", bench_laplace2d), ("graph_4m", bench_graph),
                         ("batch_512x16", bench_batched), ("autograd_512", bench_autograd),
                         ("inexact_fp32_512", bench_inexact_mixed), ("shard_160", bench_shard),
                         ("ingest_coo_160", bench_ingest), ("signature_160", bench_signature)):
            t0 = time.time()
            try:
                extra[name] = fn(args)
            except Exception as e:  # a failing extra is reported, the headline line still prints
                extra[name] = {"error": f"{type(e).__name__}: {e}"}
                log(f"{name} failed: {e}")
            extra[name]["bench_seconds"] = round(time.time() - t0, 1)
            torch.cuda.empty_cache()
        out["workloads"] = extra
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
