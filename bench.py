"""Benchmark of the HQRRP hot path on one MI355X.

Workload: the plain all-FP64 HQRRP of a 16384 x 16384 matrix
A = U diag(sigma) V' (U, V Haar-orthogonal, sigma_i = 10^(-12 i/n)),
nb = 64, pp = 10, panel pivoting on, Omega = the reference's stream.
A step is one complete factorisation (NoFLA_HQRRP_WY_blk_var4 semantics,
device-pointer entry) of a fresh copy of A that is already resident in HBM;
the device-to-device copy that restores A is inside the timed step.

metric: QRCP GFLOP/s with the flop count 2mn^2 - 2/3 n^3 (m >= n).

--workload batched: 8192 independent matrices (m ~ U[64,512], n ~ U[32,m],
singular values 10^(-8 i/n) with Haar singular vectors) factored in one
persistent launch; a step restores the packed inputs (device copy) and
factors all of them; metric: matrices/s, with the summed flops of the
realised shapes beside it.

Also reported:
  phase_ms      -- per-phase kernel time of one extra, untimed step in which
                   every phase is bracketed by HIP events;
  roofline      -- the bandwidth- or compute-bound phase with the largest
                   time in that step (the dominant kernel): during the timed
                   steps only its launches are bracketed by HIP events
                   (recorded on the stream they run on), and its algorithmic
                   flops (or bytes) per launch are divided by its average
                   launch duration;
  chain         -- the pivot chain's latency-bound kernels (sketch and panel
                   QRCP, one persistent launch per block): time per launch
                   and per dependent pivot step;
  cpu_baseline  -- the reference's dgeqp4 compiled from its source against
                   OpenBLAS (oracle/_ref; the oracle's restatement where it
                   was not built) on the whole of the same matrix, timed on
                   the host cores; cpu_dgeqp3 beside it: LAPACK dgeqp3 from
                   the same OpenBLAS on a matrix of the same family at
                   --cpu-dgeqp3-n (its 16384 run takes minutes);
  quality       -- the result's backward error ||AP - QR||_F / ||A||_F,
                   orthogonality max|Q'Q - I| (Q from the library's dorgqr,
                   products in FP64) and the rank-revealing ratio
                   sigma_k(A) / |R_kk| (sigma_k known exactly for the
                   generated matrices).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_MFMA = 78.6e12  # MI355X dense FP64 matrix, FLOP/s
PEAK_HBM = 8.0e12         # MI355X HBM3E, B/s

# HBM bytes per launch of a phase's kernel, measured with rocprofv3 --pmc
# FETCH_SIZE and --pmc WRITE_SIZE on this workload (summary and commands in
# profiles/hqrrp_16384_hbm_counters.md), each counter scaled by what it reads
# for a plain copy kernel of known size (FETCH_SIZE x2.0, WRITE_SIZE x1.0):
# the KiB summed over the phase's kernels of all 255 blocks, per block.
# Keyed by (phase, m, n, nb); other shapes report null.
MEASURED_TRAFFIC = {("trail_update", 16384, 16384, 64):
                    round((106359050.4375 * 2.0 + 179231455.125) * 1024 / 255)}

# Phase -> (bound, kernel symbol prefix in rocprof)
PHASE_BOUND = {
    "sketch_gemm": "mfma",
    "sketch_qrcp": "latency",
    "gather": "hbm",
    "panel": "latency",
    "panel_post": "hbm",
    "trail_vtc": "mfma",
    "trail_tw": "mfma",
    "trail_update": "hbm",
    "downdate": "hbm",
    "select_update": "hbm",
}


def qr_flops(m, n):
    if m >= n:
        return 2.0 * m * n * n - 2.0 / 3.0 * n ** 3
    return 2.0 * n * m * m - 2.0 / 3.0 * m ** 3


def _median_time(fn, reps):
    """Median wall time of `reps` calls after one warm-up (reps > 1), or of
    a single call."""
    import numpy as np

    if reps > 1:
        fn()
    ts = []
    for _ in range(max(1, reps)):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), len(ts)


def cpu_baseline(A_host, nb, pp, threads, reps):
    """The reference's dgeqp4 (compiled from its source, OpenBLAS) on the
    whole matrix; the oracle's restatement of its driver where the compiled
    reference is absent."""
    import oracle

    oracle.set_threads(threads)
    m, n = A_host.shape
    use_ref = oracle.ref() is not None and (nb, pp) == (64, 10)
    if use_ref:
        fn = lambda: oracle.dgeqp4(A_host, use_ref=True)  # noqa: E731
        what = ("the reference's dgeqp4 (NoFLA_HQRRP_WY_blk_var4.c compiled "
                "from source, OpenBLAS")
    else:
        fn = lambda: oracle.hqrrp(A_host, nb, pp, 1)  # noqa: E731
        what = ("the oracle's restatement of NoFLA_HQRRP_WY_blk_var4 "
                "(OpenBLAS")
    dt, runs = _median_time(fn, reps)
    fl = qr_flops(m, n)
    how = (f"median of {runs} after a warm-up" if runs > 1 else "one run")
    return {"value": round(fl / dt / 1e9, 2), "unit": "GFLOP/s",
            "cores": threads, "kind": "reference" if use_ref else "port",
            "sample": (f"{what}, {threads} threads) on the whole {m}x{n} "
                       f"matrix: {dt:.2f} s ({how})")}


def cpu_dgeqp3(A_host, threads, reps):
    import oracle

    oracle.set_threads(threads)
    m, n = A_host.shape
    dt, runs = _median_time(lambda: oracle.dgeqp3(A_host), reps)
    return {"value": round(qr_flops(m, n) / dt / 1e9, 2), "unit": "GFLOP/s",
            "cores": threads, "kind": "lapack",
            "sample": (f"LAPACK dgeqp3 (OpenBLAS, {threads} threads) on a "
                       f"{m}x{n} matrix of the same family: {dt:.2f} s "
                       f"({'median of %d after a warm-up' % runs if runs > 1 else 'one run'})")}


def factor_quality(h, A0, Af, jp, tau, m, n, sigma=None, kq=None):
    """Backward error, orthogonality and sigma_k / |R_kk| of a factorisation
    on the device (FP64 products)."""
    import torch

    import hqrrp_amd

    k = min(m, n)
    Qt = hqrrp_amd.orgqr_device(h, Af, tau, m, k, k)
    h.synchronize()
    torch.cuda.synchronize()
    R = torch.triu(Af.t()[:k, :])
    AP = A0[(jp.long() - 1)].t()
    res = (torch.linalg.matrix_norm(AP - Qt.t() @ R)
           / torch.linalg.matrix_norm(AP)).item()
    del AP
    orth = (Qt @ Qt.t() - torch.eye(k, dtype=torch.float64, device=Af.device)
            ).abs().max().item()
    eps = 2.0 ** -52
    out = {"backward_error": res, "orthogonality": orth,
           "backward_error_over_n_eps": round(res / (n * eps), 4),
           "orthogonality_over_n_eps": round(orth / (n * eps), 4),
           "bound_over_n_eps": 50}
    if sigma is not None:
        d = torch.diagonal(R).abs()
        ratio = sigma[:k] / d
        out["rank_ratio"] = {
            "sigma_k_over_abs_Rkk_min": round(ratio.min().item(), 4),
            "sigma_k_over_abs_Rkk_max": round(ratio.max().item(), 4),
            "sigma_k_over_abs_Rkk_median": round(ratio.median().item(), 4)}
        if kq is not None:
            out["rank_ratio"]["k"] = kq
            out["rank_ratio"]["at_k"] = round(ratio[kq - 1].item(), 4)
    return out


def _oracle_slice(job):
    import oracle

    oracle.set_threads(1)
    mats, nbs, pp = job
    for A, nb in zip(mats, nbs):
        oracle.hqrrp(A, int(nb), pp, 1)
    return len(mats)


def cpu_baseline_batched(mats, nbs, pp, workers, budget_s=15.0):
    """The oracle on the host cores (one OpenBLAS thread per worker process)
    over a sample of the batch sized to about budget_s seconds."""
    import multiprocessing as mp

    import oracle

    oracle.set_threads(1)
    t0 = time.perf_counter()
    probe = 0
    while probe < len(mats) and time.perf_counter() - t0 < 1.0:
        oracle.hqrrp(mats[probe], int(nbs[probe]), pp, 1)
        probe += 1
    per = (time.perf_counter() - t0) / max(1, probe)
    count = int(min(len(mats), max(workers, budget_s * workers / per)))
    chunks = [(mats[i:count:workers], nbs[i:count:workers], pp)
              for i in range(workers)]
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(workers) as pool:
        done = sum(pool.map(_oracle_slice, chunks))
    dt = time.perf_counter() - t0
    return {"value": round(done / dt, 1), "unit": "matrices/s",
            "cores": workers, "kind": "port",
            "sample": (f"oracle restatement of NoFLA_HQRRP_WY_blk_var4 "
                       f"(OpenBLAS, 1 thread per process, {workers} "
                       f"processes) on the first {done} matrices of the "
                       f"batch with the same nb and pp: {dt:.2f} s")}


def bench_batched(args):
    import numpy as np
    import torch

    import hqrrp_amd
    import oracle
    from hqrrp_amd import matgen

    count, pp = args.count, 8
    workers = min(16, os.cpu_count() or 1)
    print(f"# generating {count} matrices ({workers} processes)",
          file=sys.stderr, flush=True)
    m, n = matgen.batch_shapes(count, seed=0)
    mats = matgen.batch_matrices(m, n, seed=0, workers=workers)
    h = hqrrp_amd.Handle(0)
    h.use_stream(torch.cuda.current_stream())
    tune_s, table = None, None
    if args.batch_nb is None:
        t0 = time.perf_counter()
        table = h.batch_tune(pp=pp, panel_pivoting=1)
        tune_s = time.perf_counter() - t0
        print(f"# autotuned nb per 64x64 bucket in {tune_s:.1f} s:\n{table}",
              file=sys.stderr, flush=True)
    plan = hqrrp_amd.BatchPlan(h, m, n, nb=args.batch_nb, pp=pp)
    nbs = plan.nb
    src = hqrrp_amd.PackedBatch(m, n)
    src.load(mats)
    work = hqrrp_amd.PackedBatch(m, n)
    ident = src.jpvt.clone()

    def step():
        work.A.copy_(src.A)
        work.jpvt.copy_(ident)
        plan.factor(h, work.A_ptrs, work.jpvt_ptrs, work.tau_ptrs)

    for i in range(args.warmup):
        step()
        h.synchronize()
        print(f"# warmup {i} done", file=sys.stderr, flush=True)
    # one untimed step with the kernel's per-phase cycle counts
    phase_share = {}
    if not args.no_profile:
        plan.profile(True)
        step()
        h.synchronize()
        cyc = plan.profile_read()
        plan.profile(False)
        tot = sum(cyc.values()) or 1.0
        phase_share = {k: round(v / tot, 4) for k, v in cyc.items()}
        print(f"# phase shares {phase_share}", file=sys.stderr, flush=True)
    # The timed step is a HIP graph replay: the input copies and the
    # factorisation (one stream-ordered launch per block-size class plus a
    # queue reset kernel) captured once.
    gs = torch.cuda.Stream()
    graph = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(graph, stream=gs):
        h.use_stream(gs)
        step()
    h.use_stream(torch.cuda.current_stream())
    graph.replay()
    torch.cuda.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        graph.replay()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    # The batched kernel's own duration (the roofline's denominator): HIP
    # events around the launch alone, in eager steps after the timed ones.
    ev = [(torch.cuda.Event(enable_timing=True),
           torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    for i in range(args.steps):
        work.A.copy_(src.A)
        work.jpvt.copy_(ident)
        ev[i][0].record()
        plan.factor(h, work.A_ptrs, work.jpvt_ptrs, work.tau_ptrs)
        ev[i][1].record()
    h.synchronize()
    torch.cuda.synchronize()
    kern_ms = sum(a.elapsed_time(b) for a, b in ev) / args.steps
    print(f"# timed {args.steps} steps: {ms:.2f} ms/step, kernel "
          f"{kern_ms:.2f} ms", file=sys.stderr, flush=True)
    flops = float(sum(qr_flops(int(a), int(b)) for a, b in zip(m, n)))
    # backward error of a sample of the factorisations
    out = work.unpack()
    worst = 0.0
    for i in range(0, count, max(1, count // 64)):
        res, orth = oracle.residuals(mats[i], *out[i])
        worst = max(worst, res / (max(m[i], n[i]) * np.finfo(float).eps),
                    orth / (max(m[i], n[i]) * np.finfo(float).eps))
    ach = flops / (kern_ms / 1e3)
    roof = {"bound": "mfma", "kernel": "batch_hqrrp_kernel",
            "achieved": round(ach / 1e12, 3), "peak": PEAK_FP64_MFMA / 1e12,
            "unit": "TFLOP/s", "frac": round(ach / PEAK_FP64_MFMA, 4),
            "traffic": None, "launches": args.steps,
            "avg_launch_us": round(kern_ms * 1e3, 1),
            "algorithmic_flops_per_launch": flops}
    cpu = None
    if not args.no_cpu:
        print("# cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline_batched(mats, nbs, pp, workers)
    plan.close()
    h.close()
    line = {
        "metric": "batched QRCP/s (matrices per second)",
        "value": round(count / (ms / 1e3), 1), "unit": "matrices/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": (f"batched HQRRP {count} matrices, m~U[64,512], "
                                f"n~U[32,m], sigma_i=10^(-8i/n)"),
                   "count": count, "pp": pp, "panel_pivoting": 1,
                   "nb_counts": {int(k): int(v) for k, v in
                                 zip(*np.unique(nbs, return_counts=True))},
                   "nb_choice": ("autotuned per 64x64 shape bucket "
                                 f"({tune_s:.1f} s, untimed)"
                                 if tune_s is not None else
                                 f"fixed {args.batch_nb}"),
                   "omega": "reference stream seed 12",
                   "step": "HIP graph replay of the input copies and the "
                           "batched launch"},
        "gflops_realised": round(flops / (ms / 1e3) / 1e9, 1),
        "kernel_ms": round(kern_ms, 3),
        "max_backward_error_over_n_eps": round(worst, 3),
        "phase_cycle_share": phase_share,
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def bench_adjoint(args):
    """QR at fixed pivots of a 4096 x 1024 Gaussian truncated to k = 256
    (Q, R) and its adjoint with Gaussian cotangents: ms per forward and per
    backward, each timed with HIP events over the steps."""
    import torch

    import hqrrp_amd

    m, n, k = 4096, 1024, 256
    g = torch.Generator().manual_seed(0)
    A = torch.randn(m, n, dtype=torch.float64, generator=g).cuda()
    p = (torch.randperm(n, generator=g).to(torch.int32) + 1).cuda()
    gQ = torch.randn(m, k, dtype=torch.float64, generator=g).cuda()
    gR = torch.randn(k, n, dtype=torch.float64, generator=g).cuda()
    h = hqrrp_amd.Handle(0)
    h.use_stream(torch.cuda.current_stream())
    for _ in range(args.warmup):
        Q, R = hqrrp_amd.qr_fixed(h, A, p, k)
        hqrrp_amd.qr_fixed_adjoint(h, A, p, k, Q, R, gQ, gR)
    torch.cuda.synchronize()
    f_ms = b_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        e[0].record()
        Q, R = hqrrp_amd.qr_fixed(h, A, p, k)
        e[1].record()
        gA = hqrrp_amd.qr_fixed_adjoint(h, A, p, k, Q, R, gQ, gR)
        e[2].record()
        torch.cuda.synchronize()
        f_ms += e[0].elapsed_time(e[1])
        b_ms += e[1].elapsed_time(e[2])
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    f_ms /= args.steps
    b_ms /= args.steps
    assert torch.isfinite(gA).all()
    h.close()
    # forward: QR of the k pivoted columns, Q' applied to the other n-k, Q
    # formed; backward: the products of the adjoint formula
    fl_f = (2.0 * m * k * k - 2.0 / 3.0 * k ** 3 + 4.0 * m * k * (n - k)
            + 4.0 * m * k * k)
    fl_b = 2.0 * m * k * (n - k) * 2 + 2.0 * k * k * n + 2.0 * m * k * k * 3
    line = {
        "metric": "fixed-pivot QR + adjoint ms (4096x1024, k=256)",
        "value": round(ms, 3), "unit": "ms", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "QR at fixed pivots 4096x1024 Gaussian, k=256, "
                               "Gaussian cotangents",
                   "m": m, "n": n, "k": k},
        "forward_ms": round(f_ms, 3), "backward_ms": round(b_ms, 3),
        "forward_gflops": round(fl_f / (f_ms / 1e3) / 1e9, 1),
        "backward_gflops": round(fl_b / (b_ms / 1e3) / 1e9, 1),
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def bench_classic(args):
    """dgeqp4's small-size path: the classic greedy QRCP (dgeqp3 semantics)
    of a 999 x 999 Gaussian, the largest size the reference hands to LAPACK
    dgeqp3; GFLOP/s of 2mn^2 - 2/3 n^3 from HIP events around the device
    entry (the input copy stays outside the events)."""
    import numpy as np
    import torch

    import hqrrp_amd
    from hqrrp_amd import matgen

    m = n = 999
    A_host = matgen.gaussian(m, n, 0)
    A0 = hqrrp_amd.to_colmajor(A_host)
    W = torch.empty_like(A0)
    ids = torch.arange(1, n + 1, dtype=torch.int32, device="cuda")
    jp = torch.empty_like(ids)
    tau = torch.empty(n, dtype=torch.float64, device="cuda")
    h = hqrrp_amd.Handle(0)
    h.use_stream(torch.cuda.current_stream())

    def step():
        W.copy_(A0)
        jp.copy_(ids)
        e = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        e[0].record()
        hqrrp_amd.geqp3_device(h, W, jp, tau, m, n)
        e[1].record()
        return e

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ms = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        e = step()
        torch.cuda.synchronize()
        ms += e[0].elapsed_time(e[1])
    wall = (time.perf_counter() - t0) * 1e3 / args.steps
    ms /= args.steps
    h.close()
    fl = qr_flops(m, n)
    # Algorithmic HBM bytes: step i reads and writes the trailing block
    # A(i:m, i:n) once, 16 bytes per element (the kernels move whole columns,
    # rows above i included, so they move more than this).
    byts = sum(16.0 * (m - i) * (n - i) for i in range(min(m, n)))
    ach = byts / (ms / 1e3) / 1e9
    roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": 8000.0,
            "unit": "GB/s", "frac": round(ach / 8000.0, 4), "traffic": None}
    cpu = None
    if not args.no_cpu:
        import oracle

        threads = min(16, os.cpu_count() or 1)
        oracle.set_threads(threads)
        use_ref = oracle.ref() is not None
        ts = []
        for _ in range(4):
            t1 = time.perf_counter()
            oracle.dgeqp4(A_host, use_ref=use_ref)
            ts.append(time.perf_counter() - t1)
        dt = float(np.median(ts[1:]))
        cpu = {"value": round(fl / dt / 1e9, 2), "unit": "GFLOP/s",
               "cores": threads, "kind": "reference" if use_ref else "port",
               "sample": (f"dgeqp4 on the same {m}x{n} matrix (its LAPACK "
                          f"dgeqp3 branch, OpenBLAS, {threads} threads): "
                          f"median of 3 after a warm-up, {dt * 1e3:.1f} ms")}
    line = {
        "metric": "classic QRCP GFLOP/s (2mn^2-2/3n^3), dgeqp4 small-size "
                  "branch",
        "value": round(fl / (ms / 1e3) / 1e9, 2), "unit": "GFLOP/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3), "wall_ms_per_step": round(wall, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "classic QRCP (dgeqp3 semantics) of a "
                               "999x999 Gaussian", "m": m, "n": n},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--m", type=int, default=16384)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--nb", type=int, default=64)
    ap.add_argument("--pp", type=int, default=10)
    ap.add_argument("--cpu-reps", type=int, default=1,
                    help="CPU baseline runs (> 1: a warm-up, then the median)")
    ap.add_argument("--cpu-dgeqp3-n", type=int, default=8192,
                    help="size of LAPACK dgeqp3's sample (0: skip; the "
                         "benchmark size takes minutes)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-quality", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--lookahead", type=int, default=1,
                    help="overlap the next panel with the trailing update")
    ap.add_argument("--panel", choices=("default", "cqr", "step"),
                    default="default",
                    help="panel factorisation: the library's default, the "
                         "two-Gram-pass kernel (cqr) or the per-step kernel")
    ap.add_argument("--input", choices=("graded", "gaussian", "lowrank"),
                    default="graded",
                    help="graded: the benchmark matrix; gaussian: same shape, "
                         "no QR in its generation (for hardware-counter runs, "
                         "whose bytes per launch depend only on the shapes); "
                         "lowrank: rank 512 (Gaussian factors, sigma_i = 1/i) "
                         "+ 1e-8 N(0,1), the mixed-precision workload")
    ap.add_argument("--sketch", choices=("fp64", "fp32", "bf16"),
                    default="fp64",
                    help="sketch precision (panel, T and trailing update are "
                         "FP64 in every case)")
    ap.add_argument("--workload",
                    choices=("single", "batched", "adjoint", "classic"),
                    default="single",
                    help="single: one large HQRRP (the headline); batched: "
                         "many small matrices in one launch; adjoint: QR at "
                         "fixed pivots and its adjoint; classic: dgeqp4's "
                         "small-size classic QRCP")
    ap.add_argument("--count", type=int, default=8192,
                    help="matrices of the batched workload")
    ap.add_argument("--batch-nb", type=int, default=None,
                    help="block size of every batched matrix (default: "
                         "chosen per shape)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain copy kernel (read MIB MiB, write "
                         "MIB MiB) three times: the HBM-counter calibration "
                         "for the roofline traffic")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            torch.add(x, 0.0, out=y)
        torch.cuda.synchronize()
        print(f"# copy kernel: {n * 8} bytes read and written per launch",
              flush=True)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: HQRRP is single-device in this "
              "build (multi-GPU sharding is a design note only); run with "
              "--gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batched":
        bench_batched(args)
        return
    if args.workload == "adjoint":
        bench_adjoint(args)
        return
    if args.workload == "classic":
        bench_classic(args)
        return

    import numpy as np
    import torch

    import hqrrp_amd
    from hqrrp_amd import matgen

    m, n, nb, pp = args.m, args.n, args.nb, args.pp
    dev = "cuda:0"
    print(f"# generating {args.input} {m}x{n} input on the GPU",
          file=sys.stderr, flush=True)
    sigma, kq = None, None
    if args.input == "graded":
        A0 = matgen.graded_device(m, n, seed=0, device=dev)
        sigma = 10.0 ** (-12.0 * torch.arange(min(m, n), device=dev,
                                              dtype=torch.float64) / min(m, n))
        kq = min(m, n) // 2
    elif args.input == "lowrank":
        A0, sigma = matgen.lowrank_device(m, n, rank=512, seed=0, device=dev,
                                          core_sigma=True)
        sigma = torch.cat([sigma, torch.zeros(min(m, n) - sigma.numel(),
                                              device=dev,
                                              dtype=torch.float64)])
        kq = 512
    else:
        A0 = matgen.gaussian_device(m, n, seed=0, device=dev)
    sketch_prec = {"fp64": 0, "fp32": 1, "bf16": 2}[args.sketch]
    A = torch.empty_like(A0)
    jp = torch.empty(n, dtype=torch.int32, device=dev)
    ident = torch.arange(1, n + 1, dtype=torch.int32, device=dev)
    tau = torch.empty(min(m, n), dtype=torch.float64, device=dev)
    h = hqrrp_amd.Handle(0)
    h.use_stream(torch.cuda.current_stream())
    if args.panel != "default":
        h.set_panel_alg(0 if args.panel == "cqr" else 1)

    def step():
        A.copy_(A0)
        jp.copy_(ident)
        hqrrp_amd.hqrrp_device(h, A, jp, tau, m, n, nb=nb, pp=pp,
                               panel_pivoting=1, lookahead=args.lookahead,
                               sketch_prec=sketch_prec)

    for i in range(args.warmup):
        step()
        h.synchronize()
        print(f"# warmup {i} done", file=sys.stderr, flush=True)

    # One untimed step with every phase timed: the breakdown, and the choice
    # of the dominant phase (timing every phase costs ~10% of the step).
    prof_all, dominant = {}, None
    if not args.no_profile:
        h.profile(True)
        step()
        h.synchronize()
        prof_all = h.profile_read()
        # The roofline is the bandwidth- or compute-bound phase with the most
        # time; the pivot chain's persistent kernels (sketch and panel QRCP)
        # are bound by the latency of their dependent pivot steps and are
        # reported per step in "chain".
        dominant = max((k for k in prof_all if PHASE_BOUND[k] != "latency"),
                       key=lambda k: prof_all[k][0])
        h.profile(True, phases=[dominant])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
    h.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    print(f"# timed {args.steps} steps: {ms:.1f} ms/step", file=sys.stderr,
          flush=True)
    prof = h.profile_read() if not args.no_profile else {}
    h.profile(False)

    flops = qr_flops(m, n)
    value = flops / (ms / 1e3) / 1e9

    roof = None
    phase_ms = {}
    gemm_roof = {}
    chain = {}
    if prof_all:
        # The trailing-update kernels carry ~97% of the flops; report their
        # rooflines (from the breakdown step) beside the dominant phase's.
        for name in ("trail_vtc", "trail_update"):
            t_ms, launches, fl, by = prof_all[name]
            if t_ms > 0:
                gemm_roof[name] = {
                    "tflops": round(fl / (t_ms / 1e3) / 1e12, 2),
                    "frac_mfma": round(fl / (t_ms / 1e3) / PEAK_FP64_MFMA, 4),
                    "hbm_gbs": round(by / (t_ms / 1e3) / 1e9, 1),
                    "frac_hbm": round(by / (t_ms / 1e3) / PEAK_HBM, 4)}
        phase_ms = {k: round(v[0], 3) for k, v in prof_all.items()}
        # Pivot chain: one launch per block, nb dependent pivot steps each.
        for name in ("sketch_qrcp", "panel"):
            t_ms, launches, fl, by = prof_all[name]
            if launches:
                chain[name] = {
                    "ms": round(t_ms, 3), "launches": launches,
                    "us_per_launch": round(1e3 * t_ms / launches, 2),
                    "us_per_pivot_step": round(1e3 * t_ms / launches / nb, 3),
                    "share_of_step": round(t_ms / ms, 4)}
    if prof and dominant:
        name = dominant
        t_ms, launches, fl, by = prof[name]
        bound = PHASE_BOUND[name]
        avg_s = t_ms / 1e3 / max(1, launches)
        traffic = MEASURED_TRAFFIC.get((name, m, n, nb))
        if bound == "mfma":
            ach = fl / max(1, launches) / avg_s / 1e12
            roof = {"bound": "mfma", "kernel": name, "achieved": round(ach, 3),
                    "peak": PEAK_FP64_MFMA / 1e12, "unit": "TFLOP/s",
                    "frac": round(ach * 1e12 / PEAK_FP64_MFMA, 4),
                    "traffic": traffic, "launches": launches,
                    "avg_launch_us": round(avg_s * 1e6, 2)}
        else:
            ach = by / max(1, launches) / avg_s / 1e9
            roof = {"bound": "hbm", "kernel": name, "achieved": round(ach, 1),
                    "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                    "frac": round(ach * 1e9 / PEAK_HBM, 4),
                    "traffic": traffic,
                    "algorithmic_bytes_per_launch": round(by / max(1, launches)),
                    "launches": launches,
                    "avg_launch_us": round(avg_s * 1e6, 2)}

    quality = None
    if not args.no_quality:
        quality = factor_quality(h, A0, A, jp, tau, m, n, sigma, kq)
        if args.input == "lowrank" and quality.get("rank_ratio"):
            # sigma_k of A differs from the noise-free core's by at most
            # ||noise||_2 <= 1e-8 (sqrt(m) + sqrt(n)) (Weyl); only k <= 512
            # is meaningful, so report at k = 512.
            rr = quality["rank_ratio"]
            quality["rank_ratio"] = {"k": rr["k"], "at_k": rr["at_k"]}
        print(f"# quality {quality}", file=sys.stderr, flush=True)
    cpu, cpu3 = None, None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        print(f"# cpu baseline on {threads} threads", file=sys.stderr,
              flush=True)
        A_host = np.asfortranarray(A0.cpu().numpy().T)
        cpu = cpu_baseline(A_host, nb, pp, threads, args.cpu_reps)
        print(f"# cpu baseline {cpu}", file=sys.stderr, flush=True)
        del A_host
        if args.cpu_dgeqp3_n > 0:
            q = min(args.cpu_dgeqp3_n, m, n)
            Aq = (matgen.graded_device(q, q, seed=0, device=dev)
                  if args.input == "graded" else
                  matgen.gaussian_device(q, q, seed=0, device=dev))
            Aq_host = np.asfortranarray(Aq.cpu().numpy().T)
            del Aq
            cpu3 = cpu_dgeqp3(Aq_host, threads, args.cpu_reps)
            print(f"# cpu dgeqp3 {cpu3}", file=sys.stderr, flush=True)

    h.close()
    line = {
        "metric": ("QRCP GFLOP/s (2mn^2-2/3n^3), plain FP64 sketch"
                   if args.sketch == "fp64" else
                   f"QRCP GFLOP/s (2mn^2-2/3n^3), mixed {args.sketch} sketch"),
        "value": round(value, 2), "unit": "GFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": (
                       f"HQRRP plain FP64 {m}x{n} graded "
                       f"(Haar U,V, sigma_i=10^(-12i/n))"
                       if args.input == "graded" and args.sketch == "fp64" else
                       f"HQRRP {args.sketch} sketch {m}x{n} {args.input}"),
                   "m": m, "n": n, "nb": nb, "pp": pp, "panel_pivoting": 1,
                   "sketch": args.sketch, "omega": "reference stream seed 12",
                   "lookahead": args.lookahead, "panel": args.panel},
        "roofline": roof, "cpu_baseline": cpu, "cpu_dgeqp3": cpu3,
        "vs_cpu_baseline": (round(value / cpu["value"], 2) if cpu else None),
        "quality": quality, "phase_ms": phase_ms,
        "gemm_roofline": gemm_roof, "chain": chain,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
