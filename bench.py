"""Benchmark: pair-histogram throughput on one MI355X.

Workload (default): the spherical joint-2D configuration -- N = 262144 points on the unit sphere
(half uniform, half on 512 great-circle tracks), Haversine(R=1), operator du_L
(ProjectedStructureFunctionType{1,0}), joint histogram of 64 log distance bins [1e-5, pi] x 96
linear du_L bins over [-4s, 4s].  One step = one sf_hist_device call over all N(N-1)/2 pairs
with the inputs already resident in HBM.  The default precision is "mixed_fixed": binary32 pair
arithmetic, binary64 re-binning of pairs near any bin edge (counts bit-exact), joint cells kept
in LDS as 64-bit fixed-point words and flushed to binary64.  The strict "mixed" and the binary64
kernels are timed on the same inputs, and the per-bin error against binary64 is reported.

Prints one JSON line (metric Gpairs/s) with a live roofline entry for the dominant kernel (HIP
events around its launch, recorded by the library on the call's stream) and a CPU baseline (the
C++ restatement of the reference's SerialBackend/ThreadedBackend, timed on a bounded sample of the
same workload).  The HBM traffic of that kernel comes from two rocprofv3 counter runs (FETCH_SIZE,
WRITE_SIZE) of this script with --copy-probe, scaled by the counters' reading for a plain copy of
known size; the figures are kept in TRAFFIC_FILE and read back here.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "structurefunctions.jl_amd")
for p in (PKG, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

# FLOP-equivalents per pair (FMA = 2, add/mul/cmp/cvt = 1, binary32 transcendental = 4; integer
# and LDS work 0) of the algorithmic pair computation, per workload.  Spherical joint du_L: the
# common spherical part 43 (differences, chord^2, chord^2-space bin, sin^2 sigma, ok test, rsq,
# tangent projections) + du_L 8 + linear value bin 8 + one accumulation 1.
FLOPS_PER_PAIR = {"sphere_joint_L1": 60, "flat2_L2": 28}
# algorithmic FLOP-equivalents per pair of every benchmark configuration's kernels (FMA 2, other
# arithmetic 1, binary32 transcendental 4; integer and LDS work 0): flat 2D one operator per pass,
# flat 3D joint du_L, sphere 1D operators and the sphere joint du_L
TABLE_FLOPS = {("cfg1", "L2"): 28,
               ("cfg2", "L1"): 27, ("cfg2", "T1"): 30, ("cfg2", "L2"): 28, ("cfg2", "T2"): 32,
               ("cfg2", "L3"): 29, ("cfg2", "T3"): 32,
               ("cfg3", "joint L1"): 41,
               ("cfg4", "L2"): 53, ("cfg4", "T2"): 53, ("cfg4", "L3"): 54, ("cfg4", "T3"): 54,
               ("cfg4", "joint L1"): 60,
               # the six single-pass invariants from one pair evaluation: flat 2D 40; sphere: the
               # common spherical part 43, du_L 8, du_T 8, |du|^2 and the four moments 10
               ("cfg2", "single pass"): 40, ("cfg4", "single pass"): 69,
               # fused outputs (one pair evaluation, sf_hist_multi_device), counted the same way:
               # flat 2D all six projected operators 40; sphere four 1D + joint du_L 76
               ("cfg2", "fused L1 T1 L2 T2 L3 T3"): 40, ("cfg4", "fused L2 T2 L3 T3 + joint L1"): 76}
# (configuration, operator) rows also timed with the binary64 drop-in kernel (against the FP64 peak)
FP64_ROWS = {("cfg1", "L2"), ("cfg2", "L2"), ("cfg2", "T3"), ("cfg3", "joint L1"), ("cfg4", "L2"),
             ("cfg4", "joint L1")}
FUSED = {"fused L1 T1 L2 T2 L3 T3": ["L1", "T1", "L2", "T2", "L3", "T3"],
         "fused L2 T2 L3 T3 + joint L1": ["L2", "T2", "L3", "T3", "joint L1"]}
OPS = {"L1": (0, 1, 0), "T1": (0, 0, 1), "L2": (0, 2, 0), "T2": (0, 0, 2), "L3": (0, 3, 0),
       "T3": (0, 0, 3), "joint L1": (0, 1, 0), "single pass": (6, 0, 0)}
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
METRIC = "Gpairs/s + % FP32 VALU roofline; mixed-prec vs FP64 GPU path at fixed error"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hist32_sphere_joint_fixed_N262144_traffic.txt")
PREC_CODE = {"fp64": 0, "mixed": 1, "mixed_fixed": 2}
DTYPE = {"fp64": "f64",
         "mixed": "f32 pair arithmetic, f64 re-binning and accumulation",
         "mixed_fixed": "f32 pair arithmetic, f64 re-binning, 64-bit fixed-point LDS cells, f64 accumulation"}
COPY_PROBE_BYTES = 1 << 30


def read_traffic(path):
    """'traffic_bytes_per_launch: <int>' from the counter summary, or None."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(float(line.split(":", 1)[1]))
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", choices=["mixed_fixed", "mixed", "fp64"], default="mixed_fixed")
    ap.add_argument("--n", type=int, default=262144)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the binary64 comparison step")
    ap.add_argument("--no-fused", action="store_true",
                    help="skip the timing of every output of the configuration in one fused call")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one device copy of 1 GiB (counter calibration for HBM traffic)")
    ap.add_argument("--workload", choices=["sphere", "batch", "local", "single_pass_2d", "aux", "table"],
                    default="sphere",
                    help="sphere: the headline (N=262144 spherical joint); batch: 256 ragged 2D time "
                         "slices in one plan (hipGraph) against sequential binary64 launches; local: "
                         "the sphere field binned only up to 0.05 rad (spatial sort and tile culling "
                         "in one-shot calls, against a plan, which evaluates every pair); "
                         "single_pass_2d: the six invariants' joint histograms of the sphere field; "
                         "aux: eight velocity fields on the sphere points (auxiliary axis, B-strips)")
    ap.add_argument("--report-errors", action="store_true",
                    help="print the per-order error of the binary32 kernels against binary64 instead")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding "
                                   "is a design note only)"}))
        sys.exit(2)
    if args.report_errors:
        return report_errors(args)
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "local":
        return bench_local(args)
    if args.workload == "single_pass_2d":
        return bench_single_pass_2d(args)
    if args.workload == "aux":
        return bench_aux(args)
    if args.workload == "table":
        return bench_table(args)

    import numpy as np
    import torch

    from sfamd import _native, synthetic

    x, u, dedges, vedges = synthetic.config4(N=args.n)
    N = x.shape[1]
    pairs = N * (N - 1) // 2
    dev = torch.device("cuda:0")
    # column-major (W, N) == row-major (N, W)
    dx = torch.from_numpy(np.ascontiguousarray(x.T)).to(dev)
    du = torch.from_numpy(np.ascontiguousarray(u.T)).to(dev)
    nd, nv = len(dedges) - 1, len(vedges) - 1
    sums = torch.zeros(nd * nv, dtype=torch.float64, device=dev)
    counts = torch.zeros(nd * nv, dtype=torch.int64, device=dev)
    ctx = _native.Context(0)
    stream = torch.cuda.current_stream().cuda_stream

    def problem(prec):
        p = _native.SfProblem()
        p.D, p.W, p.N = 2, 2, N
        p.x, p.u = dx.data_ptr(), du.data_ptr()
        p.geometry, p.radius = 1, 1.0
        p.op, p.nl, p.nt = 0, 1, 0
        p.n_dist, p.dist_edges = nd, dedges.ctypes.data
        p.n_val, p.val_edges = nv, vedges.ctypes.data
        p.precision = PREC_CODE[prec]
        p.count_bits = 64
        return p

    def step(prec):
        ctx.hist_device([problem(prec)], sums.data_ptr(), counts.data_ptr(), stream)

    for _ in range(args.warmup):
        step(args.precision)
    torch.cuda.synchronize()
    kern_ns = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(args.precision)
        st = ctx.stats()
        kern_ns.append(st["hist_ns"])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms = dt * 1e3 / args.steps
    gps = pairs * args.steps / dt / 1e9
    kern_s = float(np.mean(kern_ns)) * 1e-9
    flops = FLOPS_PER_PAIR["sphere_joint_L1"]
    achieved = pairs * flops / kern_s / 1e12
    peak = PEAK_FP64_TFLOPS if args.precision == "fp64" else PEAK_FP32_TFLOPS
    rebinned = st["rebinned"]

    def rate(prec, reps=2):
        best = None
        for _ in range(reps):
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            step(prec)
            torch.cuda.synchronize()
            dt1 = time.perf_counter() - t1
            best = dt1 if best is None else min(best, dt1)
        return pairs / best / 1e9

    def result(prec):
        sums.zero_()
        counts.zero_()
        step(prec)
        torch.cuda.synchronize()
        return sums.cpu().numpy().copy(), counts.cpu().numpy().copy()

    extra = {}
    if not args.no_fp64 and args.precision != "fp64":
        # binary64 kernel on the same inputs: throughput, and each binary32 result's error against it
        s64, c64 = result("fp64")
        g64 = rate("fp64")
        extra["fp64_gpairs_s"] = round(g64, 3)
        extra["speedup_vs_fp64"] = round(gps / g64, 3)
        # per-cell relative error |s32 - s64| / |s64| over every non-empty cell.  A joint
        # (distance, du_L) cell holds values of one sign (the value edge at 0 splits the signs),
        # so |sum v| = sum |v| there and the ratio is the error relative to the cell's magnitude.
        # The binary64 kernel's counts equal the oracle's and its sums agree to 1e-12 (the
        # scale parity tests), so it stands in for the oracle at this size.
        nz = c64 > 0
        precs = [args.precision] + (["mixed"] if args.precision == "mixed_fixed" else [])
        for prec in precs:
            s32, c32 = result(prec)
            rel = np.abs(s32 - s64)[nz] / np.maximum(np.abs(s64[nz]), 1e-300)
            cnt = c64[nz]
            bad = rel > 1e-6
            extra[f"{prec}_vs_fp64"] = {
                "counts_bit_exact": bool(np.array_equal(c32, c64)),
                "cells_nonempty": int(nz.sum()),
                "sums_max_rel_err_all_cells": float(rel.max()) if rel.size else None,
                "sums_median_rel_err_all_cells": float(np.median(rel)) if rel.size else None,
                "cells_over_1e-6": int(bad.sum()),
                "pairs_in_cells_over_1e-6": int(cnt[bad].sum()),
                "largest_count_of_a_cell_over_1e-6": int(cnt[bad].max()) if bad.any() else 0,
                "sums_max_rel_err_cells_ge_1000_pairs": float(rel[cnt >= 1000].max()) if (cnt >= 1000).any() else None}
        if args.precision == "mixed_fixed":
            extra["mixed_gpairs_s"] = round(rate("mixed"), 3)

    if not args.no_fused and args.precision != "fp64":
        extra["all_outputs"] = fused_vs_separate(ctx, stream, dx, du, N, dedges, vedges, args.precision)

    if args.copy_probe:
        a = torch.empty(COPY_PROBE_BYTES // 4, dtype=torch.float32, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        extra["copy_probe_bytes"] = COPY_PROBE_BYTES
        del a, b

    cpu = None
    if not args.no_cpu:
        host = host_info()
        threads = host["threads"]

        def oracle_rows(rows, nthreads):
            import _oracle
            _oracle.hist(x, u, dedges, op=(0, 1, 0), geometry="deg", vedges=vedges, nthreads=nthreads,
                         compensated=False, rows=rows)

        # ThreadedBackend restatement on the host's CPU share, SerialBackend on one thread; each
        # on a row sample of about cpu_seconds (min of two timings, benchmark/regimes.jl:31-41)
        thr = cpu_rate(oracle_rows, N, threads, args.cpu_seconds)
        ser = cpu_rate(oracle_rows, N, 1, args.cpu_seconds)
        cpu = {"value": round(thr["gpairs_s"], 4), "unit": "Gpairs/s", "cores": threads, "kind": "port",
               "sample": f"every {thr['stride']}th outer row i of the same workload ({thr['pairs']} pairs), "
                         f"C++ restatement of ThreadedBackend (round-robin rows, per-thread histograms), "
                         f"binary64, {threads} threads ({host['cpu']}, nproc {host['nproc']})",
               "serial_1_thread_gpairs_s": round(ser["gpairs_s"], 5),
               "serial_sample": f"every {ser['stride']}th outer row ({ser['pairs']} pairs), SerialBackend restatement"}
        # configuration 1 (the reference's CPU-runnable case) at its full size on one thread, and
        # the same call on the GPU
        cpu["cfg1_serial_full"] = cfg1_serial_vs_gpu(ctx, stream, dev)

    out = {
        "metric": METRIC,
        "value": round(gps, 3), "unit": "Gpairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPE[args.precision],
        "data": "synthetic (seeded Fourier-mode field on the sphere)",
        "config": {"workload": f"sphere N={N} Haversine(1) joint du_L x distance 64x96, {args.precision}",
                   "pairs_per_step": pairs, "precision": args.precision, "rebinned_pairs": int(rebinned)},
        # compute-bound on the vector ALU (binary32 arithmetic, no matrix instructions); 157.3
        # TFLOP/s is the published fp32 vector peak
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": peak,
                     "unit": "TFLOP/s", "frac": round(achieved / peak, 4),
                     "traffic": read_traffic(TRAFFIC_FILE), "pipe": "VALU fp32",
                     "kernel": "hist32<sphere,D=2,joint,du_L" + (",fixed-point cells>" if args.precision == "mixed_fixed" else ">")
                               if args.precision != "fp64" else "hist64<sphere,D=2,joint,du_L>",
                     "kernel_ms": round(kern_s * 1e3, 3),
                     "flops_per_pair": flops},
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out))


def fused_vs_separate(ctx, stream, dx, du, N, dedges, vedges, prec):
    """Every output the spherical configuration asks for -- du_L^2, du_T^2, du_L^3, du_T^3 over
    the 64 distance bins and the joint du_L histogram -- as one fused call (one evaluation of each
    pair) and as five separate calls; wall time of each (best of three), counts compared."""
    import numpy as np
    import torch

    from sfamd import _native

    nd, nv = len(dedges) - 1, len(vedges) - 1
    names = ["L2", "T2", "L3", "T3", "joint L1"]
    dev = dx.device
    bufs = [(torch.zeros(nd * (nv if n == "joint L1" else 1), dtype=torch.float64, device=dev),
             torch.zeros(nd * (nv if n == "joint L1" else 1), dtype=torch.int64, device=dev)) for n in names]
    outs = []
    for n in names:
        o = _native.SfOutput()
        o.op, o.nl, o.nt = OPS[n]
        if n == "joint L1":
            o.n_val, o.val_edges = nv, vedges.ctypes.data
        outs.append(o)
    p = _device_problem(_native, dx.data_ptr(), du.data_ptr(), N, 2, 2, 1, OPS["L2"], dedges, None, prec)

    def run_fused():
        ctx.hist_multi_device([p], outs, [b[0].data_ptr() for b in bufs], [b[1].data_ptr() for b in bufs], stream)

    def run_separate():
        for n, (bs, bc) in zip(names, bufs):
            q = _device_problem(_native, dx.data_ptr(), du.data_ptr(), N, 2, 2, 1, OPS[n], dedges,
                                vedges if n == "joint L1" else None, prec)
            ctx.hist_device([q], bs.data_ptr(), bc.data_ptr(), stream)

    res = {}
    counts = {}
    for label, fn in (("fused", run_fused), ("separate", run_separate)):
        best = None
        for r in range(4):
            for bs, bc in bufs:
                bs.zero_()
                bc.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            if r:
                best = dt if best is None else min(best, dt)
        res[f"{label}_ms"] = round(best * 1e3, 3)
        counts[label] = [bc.cpu().numpy().copy() for _, bc in bufs]
    pairs = N * (N - 1) // 2
    res["outputs"] = names
    res["fused_gpairs_s"] = round(pairs / res["fused_ms"] / 1e6, 3)
    res["speedup_fused_vs_separate"] = round(res["separate_ms"] / res["fused_ms"], 3)
    res["counts_equal"] = all(np.array_equal(a, b) for a, b in zip(counts["fused"], counts["separate"]))
    return res


def host_info():
    """CPU model, visible processors, and the threads a baseline may use: the process's CPU share
    (OMP_NUM_THREADS when the host sets it, else the affinity mask)."""
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = int(env) if env.isdigit() and int(env) > 0 else nproc
    return {"cpu": model, "nproc": nproc, "threads": max(1, min(threads, nproc))}


def cpu_rate(run_rows, N, nthreads, seconds):
    """Pairs per second of run_rows((0, N, stride), nthreads) on a row sample sized to about
    `seconds` (a short estimate first), min of two timings."""
    def pairs_of(stride):
        return sum(N - 1 - i for i in range(0, N, stride))

    t0 = time.perf_counter()
    est_stride = 4096
    run_rows((0, N, est_stride), nthreads)
    rate = pairs_of(est_stride) / max(time.perf_counter() - t0, 1e-6)
    total = N * (N - 1) // 2
    stride = max(1, int(total / max(rate * seconds / 2, 1.0)))
    best = None
    for _ in range(2):
        t0 = time.perf_counter()
        run_rows((0, N, stride), nthreads)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    p = pairs_of(stride)
    return {"gpairs_s": p / best / 1e9, "stride": stride, "pairs": p}


def cfg1_serial_vs_gpu(ctx, stream, dev):
    """Configuration 1 (N = 4096 flat 2D, du_L^2 over 32 log bins): the SerialBackend restatement on
    one thread over all 8,386,560 pairs, against one binary64 and one mixed device call."""
    import numpy as np
    import torch

    import _oracle
    from sfamd import _native, synthetic

    x, u, e = synthetic.config1()
    N = x.shape[1]
    pairs = N * (N - 1) // 2
    best = None
    for _ in range(2):
        t0 = time.perf_counter()
        s_o, c_o, _ = _oracle.hist(x, u, e, op=(0, 2, 0), nthreads=1, compensated=False)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    dx = torch.from_numpy(np.ascontiguousarray(x.T)).to(dev)
    du = torch.from_numpy(np.ascontiguousarray(u.T)).to(dev)
    out = {"pairs": pairs, "serial_s": round(best, 4), "serial_gpairs_s": round(pairs / best / 1e9, 4)}
    for prec in ("fp64", "mixed"):
        s = torch.zeros(len(e) - 1, dtype=torch.float64, device=dev)
        c = torch.zeros(len(e) - 1, dtype=torch.int64, device=dev)
        p = _device_problem(_native, dx.data_ptr(), du.data_ptr(), N, 2, 2, 0, (0, 2, 0), e, None, prec)
        gbest = None
        for r in range(3):
            s.zero_()
            c.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ctx.hist_device([p], s.data_ptr(), c.data_ptr(), stream)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            if r:
                gbest = dt if gbest is None else min(gbest, dt)
        out[f"gpu_{prec}_ms"] = round(gbest * 1e3, 4)
        out[f"gpu_{prec}_counts_equal_serial"] = bool(np.array_equal(c.cpu().numpy().astype(np.uint64), c_o))
    return out


def _device_problem(_native, xd, ud, N, D, W, geom, op, dedges, vedges, prec):
    p = _native.SfProblem()
    p.D, p.W, p.N = D, W, N
    p.x, p.u = xd, ud
    p.geometry, p.radius = geom, 1.0
    p.op, p.nl, p.nt = op
    p.n_dist, p.dist_edges = len(dedges) - 1, dedges.ctypes.data
    if vedges is None:
        p.n_val, p.val_edges = 0, None
    else:
        p.n_val, p.val_edges = len(vedges) - 1, vedges.ctypes.data
    p.precision = PREC_CODE[prec]
    p.count_bits = 64
    return p


def bench_single_pass_2d(args):
    """calculate_structure_functions_single_pass_2d on the headline's points and field (first
    --n points, default 65536 here): 64 log distance bins, per invariant 48 value bins (log for
    |du|^2, du_L^2, du_T^2, linear for the three signed invariants).  One device call evaluates
    every pair once per invariant (six plane histograms); the rate counts pairs of the call."""
    import numpy as np
    import torch

    from sfamd import _native, synthetic

    n = args.n if args.n != 262144 else 65536
    x, u, dedges, _ = synthetic.config4(N=n)
    N = x.shape[1]
    pairs = N * (N - 1) // 2
    m = 2.0 * float(np.sum(np.var(u, axis=1)))
    pos = np.concatenate([[0.0], np.geomspace(1e-4 * m, 8.0 * m, 48)])
    sgn = np.linspace(-4.0 * m ** 1.5, 4.0 * m ** 1.5, 49)
    ve6 = np.ascontiguousarray(np.concatenate([pos, pos, pos, sgn, sgn, sgn]))
    nd, nv = len(dedges) - 1, 48
    dev = torch.device("cuda:0")
    dx = torch.from_numpy(np.ascontiguousarray(x.T)).to(dev)
    du = torch.from_numpy(np.ascontiguousarray(u.T)).to(dev)
    ctx = _native.Context(0)
    stream = torch.cuda.current_stream().cuda_stream

    def run(prec, reps):
        p = _device_problem(_native, dx.data_ptr(), du.data_ptr(), N, 2, 2, 1,
                            (_native.OP_SINGLE_PASS, 0, 0), dedges, None, prec)
        p.n_val, p.val_edges = nv, ve6.ctypes.data
        s = torch.zeros(6 * nd * nv, dtype=torch.float64, device=dev)
        c = torch.zeros(6 * nd * nv, dtype=torch.int64, device=dev)
        best = None
        for r in range(reps + 1):
            s.zero_()
            c.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ctx.hist_device([p], s.data_ptr(), c.data_ptr(), stream)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            if r > 0:
                best = dt if best is None else min(best, dt)
        return best, s.cpu().numpy(), c.cpu().numpy()

    prec = args.precision if args.precision != "mixed_fixed" else "mixed"
    dt, s32, c32 = run(prec, max(1, args.steps))
    st32 = ctx.stats()
    dt64, s64, c64 = run("fp64", 1)
    big = c64 >= 1000
    rel = np.abs(s32[big] - s64[big]) / np.maximum(np.abs(s64[big]), 1e-300)
    out = {"metric": "Gpairs/s (N(N-1)/2 pairs of the call, six invariant planes each)",
           "value": round(pairs / dt / 1e9, 3), "unit": "Gpairs/s", "n_gpus": 1,
           "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True, "dtype": DTYPE[prec],
           "data": "synthetic (seeded Fourier-mode field on the sphere)",
           "config": {"workload": f"sphere N={N}, single pass 2D, 64 log distance bins x 48 value bins per invariant",
                      "pairs_per_step": pairs, "precision": prec},
           "fp64_gpairs_s": round(pairs / dt64 / 1e9, 3), "speedup_vs_fp64": round(dt64 / dt, 3),
           "counts_bit_exact": bool(np.array_equal(c32, c64)),
           "sums_median_rel_err": float(np.median(rel)) if rel.size else None,
           "rebinned_pairs": int(st32["rebinned"]), "kernel_ms": round(st32["hist_ns"] * 1e-6, 3)}
    print(json.dumps(out))


def bench_table(args):
    """Roofline of every benchmark configuration's kernels: for each (configuration, operator) the
    binary32 kernel's time from the library's HIP events (best of --steps calls after one warm-up)
    and the whole call's wall time; achieved = pairs x FLOP-equivalents per pair / kernel time,
    against the FP32 peak recomputed from the device's compute units and clock (256 FLOP per clock
    per CU) and against the published 157.3 TFLOP/s.  The FP64_ROWS cases are also timed with the
    binary64 drop-in kernel, against the FP64 peak (78.6 TFLOP/s published, half the device's
    FP32 figure).  One JSON line."""
    import numpy as np
    import torch

    from sfamd import _native, synthetic

    ctx = _native.Context(0)
    info = ctx.device_info()
    peak_dev = info["cus"] * 256 * info["clock_khz"] * 1e3 / 1e12
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    cases = []
    x1, u1, e1 = synthetic.config1()
    cases.append(("cfg1", "flat 2D N=4096", x1, u1, e1, None, 0, ["L2"]))
    x2, u2, e2 = synthetic.config2()
    cases.append(("cfg2", "flat 2D N=65536", x2, u2, e2, None, 0,
                  ["L1", "T1", "L2", "T2", "L3", "T3", "single pass", "fused L1 T1 L2 T2 L3 T3"]))
    x3, u3, e3, v3 = synthetic.config3()
    cases.append(("cfg3", "flat 3D N=131072, 48x128", x3, u3, e3, v3, 0, ["joint L1"]))
    x4, u4, e4, v4 = synthetic.config4()
    cases.append(("cfg4", "sphere N=262144 (degrees, Haversine(1)), 64 distance bins", x4, u4, e4, v4, 1,
                  ["L2", "T2", "L3", "T3", "joint L1", "single pass", "fused L2 T2 L3 T3 + joint L1"]))
    rows = []
    for cfg, desc, x, u, de, ve, geom, ops in cases:
        N, D, W = x.shape[1], u.shape[0], x.shape[0]
        pairs = N * (N - 1) // 2
        dx = torch.from_numpy(np.ascontiguousarray(x.T)).to(dev)
        du = torch.from_numpy(np.ascontiguousarray(u.T)).to(dev)
        for name in ops:
            fused = FUSED.get(name)
            joint = name.startswith("joint") or (fused is not None and "joint L1" in fused)
            vedges = ve if joint else None
            ncell = (len(de) - 1) * ((len(ve) - 1) if joint else (6 if name == "single pass" else 1))
            s = torch.zeros(ncell, dtype=torch.float64, device=dev)
            c = torch.zeros(ncell, dtype=torch.int64, device=dev)
            outs, bufs = [], []
            for o in fused or []:
                so = _native.SfOutput()
                so.op, so.nl, so.nt = OPS[o]
                if o == "joint L1":
                    so.n_val, so.val_edges = len(ve) - 1, ve.ctypes.data
                outs.append(so)
                n_o = (len(de) - 1) * ((len(ve) - 1) if o == "joint L1" else 1)
                bufs.append((torch.zeros(n_o, dtype=torch.float64, device=dev),
                             torch.zeros(n_o, dtype=torch.int64, device=dev)))
            precs = ["mixed_fixed", "mixed"] if joint else ["mixed"]
            if (cfg, name) in FP64_ROWS:
                precs = precs + ["fp64"]
            for prec in precs:
                p = _device_problem(_native, dx.data_ptr(), du.data_ptr(), N, D, W, geom,
                                    OPS[fused[0]] if fused else OPS[name], de, None if fused else vedges, prec)
                best_k = best_w = None
                for r in range(max(1, args.steps) + 1):
                    torch.cuda.synchronize()
                    t0 = time.perf_counter()
                    if fused:
                        ctx.hist_multi_device([p], outs, [b[0].data_ptr() for b in bufs],
                                              [b[1].data_ptr() for b in bufs], stream)
                    else:
                        ctx.hist_device([p], s.data_ptr(), c.data_ptr(), stream)
                    torch.cuda.synchronize()
                    wall = time.perf_counter() - t0
                    kern = ctx.stats()["hist_ns"] * 1e-9
                    if r > 0:
                        best_k = kern if best_k is None else min(best_k, kern)
                        best_w = wall if best_w is None else min(best_w, wall)
                f = TABLE_FLOPS[(cfg, name)]
                ach = pairs * f / best_k / 1e12
                # binary64 rows: the same FLOP-equivalents against the FP64 peak (half the FP32 one)
                peak_pub, peak_d = (PEAK_FP64_TFLOPS, peak_dev / 2) if prec == "fp64" else (PEAK_FP32_TFLOPS, peak_dev)
                rows.append({"config": cfg, "workload": desc, "operator": name, "precision": prec,
                             "pairs": pairs, "flops_per_pair": f,
                             "kernel_ms": round(best_k * 1e3, 3), "call_ms": round(best_w * 1e3, 3),
                             "kernel_gpairs_s": round(pairs / best_k / 1e9, 2),
                             "call_gpairs_s": round(pairs / best_w / 1e9, 2),
                             "achieved_tflops": round(ach, 2), "peak_tflops": peak_pub,
                             "frac_published": round(ach / peak_pub, 4),
                             "frac_device": round(ach / peak_d, 4)})
                print(json.dumps(rows[-1]), file=sys.stderr, flush=True)
        del dx, du
    out = {"metric": "per-configuration roofline of the binary32 kernels and the binary64 drop-in "
                     "(kernel-time FLOP-equivalents)",
           "unit": "TFLOP/s", "peak_published": PEAK_FP32_TFLOPS, "peak_device": round(peak_dev, 2),
           "device": info, "steps": args.steps, "data": "synthetic (seeded configurations 1-4)",
           "rows": rows}
    print(json.dumps(out))


def bench_aux(args):
    """Eight velocity fields on the headline's points (first --n points, default 65536 here): an
    auxiliary axis u (2, N, 8) with shared positions, du_L^2 over the 64 log distance bins.  The
    default launch (one pair pass per field) against the opt-in B-strip launch (geometry once per
    pair, four fields per strip) and the binary64 kernel; the rate counts field-pairs
    (8 N(N-1)/2 per call)."""
    import numpy as np
    import torch

    from sfamd import _native, synthetic

    n = args.n if args.n != 262144 else 65536
    x, u0, dedges, _ = synthetic.config4(N=n)
    N = x.shape[1]
    B = 8
    rng = np.random.default_rng(20261203)
    lon, lat = np.radians(x[0]), np.radians(x[1])
    Pt = np.stack([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)])
    E = np.stack([-np.sin(lon), np.cos(lon), np.zeros_like(lon)])
    Nn = np.stack([-np.sin(lat) * np.cos(lon), -np.sin(lat) * np.sin(lon), np.cos(lat)])
    us = []
    for b in range(B):
        U = synthetic._flat_field(rng, Pt, 128, 1.0, 256.0, -5 / 6)
        us.append(np.stack([np.sum(U * E, axis=0), np.sum(U * Nn, axis=0)]))
    pairs = B * (N * (N - 1) // 2)
    nd = len(dedges) - 1
    dev = torch.device("cuda:0")
    dx = torch.from_numpy(np.ascontiguousarray(x.T)).to(dev)
    dxs = [torch.from_numpy(np.ascontiguousarray(x.T)).to(dev) for _ in range(B)]
    dus = [torch.from_numpy(np.ascontiguousarray(v.T)).to(dev) for v in us]
    ctx = _native.Context(0)
    stream = torch.cuda.current_stream().cuda_stream
    s = torch.zeros(B * nd, dtype=torch.float64, device=dev)
    c = torch.zeros(B * nd, dtype=torch.int64, device=dev)

    def probs(shared, prec):
        out = []
        for b in range(B):
            p = _device_problem(_native, (dx if shared else dxs[b]).data_ptr(), dus[b].data_ptr(), N, 2, 2, 1,
                                (0, 2, 0), dedges, None, prec)
            p.semantics = 1
            out.append(p)
        return out

    def run(ps, reps):
        best = None
        for r in range(reps + 1):
            s.zero_()
            c.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ctx.hist_device(ps, s.data_ptr(), c.data_ptr(), stream)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            if r:
                best = dt if best is None else min(best, dt)
        return best, s.cpu().numpy(), c.cpu().numpy()

    prec = "mixed" if args.precision == "mixed_fixed" else args.precision
    # default launch: one pair pass per field (the per-slice 1D kernel)
    t_def, s1, c1 = run(probs(True, prec), max(1, args.steps))
    items_def = ctx.stats()["items"]
    # opt-in B-strips (tuning variant 3): geometry once per pair for four fields
    ctx.set_tuning(0, 0, 0, 3)
    t_strip, s2, c2 = run(probs(True, prec), max(1, args.steps))
    items_strip = ctx.stats()["items"]
    ctx.set_tuning(0, 0, 0, 0)
    t64, s3, c3 = run(probs(True, "fp64"), 1)
    out = {"metric": "Gpairs/s (field-pairs: 8 fields x N(N-1)/2 per call)",
           "value": round(pairs / t_def / 1e9, 3), "unit": "Gpairs/s", "n_gpus": 1,
           "ms_per_step": round(t_def * 1e3, 3), "higher_is_better": True, "dtype": DTYPE[prec],
           "data": "synthetic (eight seeded Fourier-mode fields on the sphere points)",
           "config": {"workload": f"sphere N={N}, u (2, N, 8) shared positions, du_L^2 over 64 log distance bins",
                      "pairs_per_step": pairs, "items_default": items_def, "items_strip": items_strip},
           "strip_gpairs_s": round(pairs / t_strip / 1e9, 3), "strip_vs_default": round(t_def / t_strip, 3),
           "fp64_gpairs_s": round(pairs / t64 / 1e9, 3),
           "counts_equal_strip": bool(np.array_equal(c1, c2)), "counts_equal_fp64": bool(np.array_equal(c1, c3))}
    print(json.dumps(out))


def bench_batch(args):
    """256 independent 2D slices of ragged N (log-uniform in [2000, 40000]); per slice the
    longitudinal second-order SF over 48 log distance bins and the joint 48 x 64 (distance, du_L)
    histogram.  Batched: one fused plan over all slices (both outputs from one evaluation of each
    pair) executed as a captured hipGraph, beside two one-operator plans whose launch shapes come
    from the autotuner (it times one-operator launches; a fused launch keeps the built-in shapes);
    baseline: one binary64 device call per slice and operator (512 launches)."""
    import numpy as np
    import torch

    from sfamd import _native, synthetic

    xs, us, dedges, vbins = synthetic.config5()
    dev = torch.device("cuda:0")
    Ns = [x.shape[1] for x in xs]
    pairs = sum(n * (n - 1) // 2 for n in Ns)
    dx = [torch.from_numpy(np.ascontiguousarray(x.T)).to(dev) for x in xs]
    du = [torch.from_numpy(np.ascontiguousarray(u.T)).to(dev) for u in us]
    nd, nv = len(dedges) - 1, len(vbins[0]) - 1
    S = len(xs)
    out1_s = torch.zeros(S * nd, dtype=torch.float64, device=dev)
    out1_c = torch.zeros(S * nd, dtype=torch.int64, device=dev)
    outj_s = torch.zeros(S * nd * nv, dtype=torch.float64, device=dev)
    outj_c = torch.zeros(S * nd * nv, dtype=torch.int64, device=dev)
    ctx = _native.Context(0)
    stream = torch.cuda.current_stream().cuda_stream
    keep = [dedges] + list(vbins)

    def probs(op, joint, prec):
        return [_device_problem(_native, dx[k].data_ptr(), du[k].data_ptr(), Ns[k], 2, 2, 0, op,
                                dedges, vbins[k] if joint else None, prec) for k in range(S)]

    prec = args.precision
    # the batched form: one fused plan (both outputs from one evaluation of each pair; the joint
    # output takes each slice's own value edges) replayed as one hipGraph
    outs = []
    for op, jv in (((0, 2, 0), False), ((0, 1, 0), True)):
        o = _native.SfOutput()
        o.op, o.nl, o.nt = op
        if jv:
            o.n_val, o.val_edges = nv, None
        outs.append(o)
    # the autotuner's shapes for the one-operator plans (sf_autotune on the joint problems of all
    # slices and on the L2 ones)
    t0 = time.perf_counter()
    tuned_joint = ctx.autotune(probs((0, 1, 0), True, prec))
    tuned_l2 = ctx.autotune(probs((0, 2, 0), False, prec))
    t_tune = time.perf_counter() - t0
    t0 = time.perf_counter()
    planf = _native.Plan(ctx, probs((0, 1, 0), True, prec), outs)
    t_build = time.perf_counter() - t0
    plan1 = _native.Plan(ctx, probs((0, 2, 0), False, prec))
    planj = _native.Plan(ctx, probs((0, 1, 0), True, prec))

    def step_batched():
        planf.execute_multi([out1_s.data_ptr(), outj_s.data_ptr()], [out1_c.data_ptr(), outj_c.data_ptr()],
                            stream, graph=True)
        planf.sync()

    def step_two_plans():
        plan1.execute(out1_s.data_ptr(), out1_c.data_ptr(), stream, graph=True)
        planj.execute(outj_s.data_ptr(), outj_c.data_ptr(), stream, graph=True)
        plan1.sync()
        planj.sync()

    def timed(fn, steps):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    dt = timed(step_batched, args.steps)
    # Gpairs/s: N(N-1)/2 summed over the slices per wall second (each pair feeds both outputs)
    rate = pairs / dt / 1e9
    dt_two = timed(step_two_plans, max(1, args.steps // 2))
    # one execution of each form from zeroed outputs: the counts must agree
    res = []
    for fn in (step_batched, step_two_plans):
        for t in (out1_s, out1_c, outj_s, outj_c):
            t.zero_()
        fn()
        torch.cuda.synchronize()
        res.append((out1_c.cpu().numpy().copy(), outj_c.cpu().numpy().copy()))
    two_counts_equal = all(np.array_equal(a, b) for a, b in zip(res[0], res[1]))

    # sequential binary64 launches, one per slice and operator
    seq_s = torch.zeros(nd * nv, dtype=torch.float64, device=dev)
    seq_c = torch.zeros(nd * nv, dtype=torch.int64, device=dev)

    seq1 = probs((0, 2, 0), False, "fp64")
    seqj = probs((0, 1, 0), True, "fp64")

    def step_sequential():
        for k in range(S):
            ctx.hist_device([seq1[k]], seq_s.data_ptr(), seq_c.data_ptr(), stream)
            ctx.hist_device([seqj[k]], seq_s.data_ptr(), seq_c.data_ptr(), stream)
        torch.cuda.synchronize()

    step_sequential()
    t0 = time.perf_counter()
    step_sequential()
    dt_seq = time.perf_counter() - t0
    rate_seq = pairs / dt_seq / 1e9

    # binary64 in the same batched plan form
    p64_1 = _native.Plan(ctx, probs((0, 2, 0), False, "fp64"))
    p64_j = _native.Plan(ctx, probs((0, 1, 0), True, "fp64"))

    def step64():
        p64_1.execute(out1_s.data_ptr(), out1_c.data_ptr(), stream, graph=True)
        p64_j.execute(outj_s.data_ptr(), outj_c.data_ptr(), stream, graph=True)
        p64_1.sync()
        p64_j.sync()

    step64()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    step64()
    torch.cuda.synchronize()
    rate_b64 = pairs / (time.perf_counter() - t0) / 1e9

    # ThreadedBackend restatement on the host's CPU share: both outputs of every 16th slice
    cpu = None
    if not args.no_cpu:
        import _oracle
        host = host_info()
        sample = list(range(0, S, 16))
        t0 = time.perf_counter()
        for k in sample:
            _oracle.hist(xs[k], us[k], dedges, op=(0, 2, 0), nthreads=host["threads"], compensated=False)
            _oracle.hist(xs[k], us[k], dedges, op=(0, 1, 0), vedges=vbins[k], nthreads=host["threads"],
                         compensated=False)
        dt_cpu = time.perf_counter() - t0
        p_cpu = sum(Ns[k] * (Ns[k] - 1) // 2 for k in sample)
        cpu = {"value": round(p_cpu / dt_cpu / 1e9, 4), "unit": "Gpairs/s", "cores": host["threads"],
               "kind": "port",
               "sample": f"slices {sample[0]}, {sample[1]}, ... (every 16th, {p_cpu} pairs, both outputs), "
                         f"C++ restatement of ThreadedBackend, binary64, {host['threads']} threads "
                         f"({host['cpu']}, nproc {host['nproc']})"}

    out = {"metric": "Gpairs/s, batched ragged time slices (N(N-1)/2 summed over slices per second; "
                     "two outputs per slice)",
           "value": round(rate, 3), "pair_evaluations_per_s_G": round(2 * rate, 3), "unit": "Gpairs/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True,
           "dtype": DTYPE[prec], "data": "synthetic (256 seeded 2D Fourier-mode fields)",
           "config": {"workload": "256 slices, N log-uniform [2000, 40000], L2 1D 48 bins + joint du_L 48x64",
                      "slices": S, "pairs_per_operator": pairs, "plan_build_s": round(t_build, 3)},
           "autotuned": {"joint_shape": list(tuned_joint), "l2_shape": list(tuned_l2),
                         "shape_fields": ["block", "j_chunk", "blocks_per_cu", "variant", "copies", "layout"],
                         "tune_s": round(t_tune, 2),
                         "applied_to": "the two one-operator plans (the fused plan keeps the built-in shapes)"},
           "two_plans_gpairs_s": round(pairs / dt_two / 1e9, 3),
           "speedup_fused_vs_two_plans": round(dt_two / dt, 3),
           "two_plans_counts_equal_fused": bool(two_counts_equal),
           "sequential_fp64_gpairs_s": round(rate_seq, 3),
           "batched_fp64_gpairs_s": round(rate_b64, 3),
           "speedup_vs_sequential_fp64": round(rate / rate_seq, 3),
           "cpu_baseline": cpu}
    del keep
    print(json.dumps(out))


def bench_local(args):
    """The N = 262144 spherical field with 64 log distance bins over [1e-5, 0.05] rad and the
    joint du_L histogram: one-shot device calls (records sorted along a Morton curve, work items
    of tiles farther apart than 0.05 rad dropped) against the same problem through a plan."""
    import numpy as np
    import torch

    from sfamd import _native, synthetic

    x, u, _, vedges = synthetic.config4(N=args.n)
    dedges = np.geomspace(1e-5, 0.05, 65)
    N = x.shape[1]
    pairs = N * (N - 1) // 2
    dev = torch.device("cuda:0")
    dx = torch.from_numpy(np.ascontiguousarray(x.T)).to(dev)
    du = torch.from_numpy(np.ascontiguousarray(u.T)).to(dev)
    nd, nv = len(dedges) - 1, len(vedges) - 1
    sums = torch.zeros(nd * nv, dtype=torch.float64, device=dev)
    counts = torch.zeros(nd * nv, dtype=torch.int64, device=dev)
    ctx = _native.Context(0)
    stream = torch.cuda.current_stream().cuda_stream
    prob = _device_problem(_native, dx.data_ptr(), du.data_ptr(), N, 2, 2, 1, (0, 1, 0), dedges, vedges,
                           args.precision)

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        best = None
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        return best

    def oneshot():
        ctx.hist_device([prob], sums.data_ptr(), counts.data_ptr(), stream)

    t_cull = timed(oneshot, args.steps)
    st = ctx.stats()
    items_cull = st["items"]
    kern_cull_ms = st["hist_ns"] / 1e6
    sums.zero_()
    counts.zero_()
    oneshot()
    torch.cuda.synchronize()
    c_cull = counts.cpu().numpy().copy()
    plan = _native.Plan(ctx, [prob])

    def planned():
        plan.execute(sums.data_ptr(), counts.data_ptr(), stream, graph=False)
        plan.sync()

    t_full = timed(planned, max(1, args.steps // 2))
    sums.zero_()
    counts.zero_()
    planned()
    torch.cuda.synchronize()
    c_full = counts.cpu().numpy().copy()
    out = {"metric": "Gpairs/s (all N(N-1)/2 pairs of the call), distance bins up to 0.05 rad",
           "value": round(pairs / t_cull / 1e9, 3), "unit": "Gpairs/s", "n_gpus": 1,
           "ms_per_step": round(t_cull * 1e3, 3), "higher_is_better": True, "dtype": DTYPE[args.precision],
           "data": "synthetic (seeded Fourier-mode field on the sphere)",
           "config": {"workload": f"sphere N={N}, 64 log distance bins [1e-5, 0.05] rad x 96 du_L bins",
                      "pairs_per_step": pairs, "items_culled_call": items_cull,
                      "kernel_ms_culled_call": round(kern_cull_ms, 3)},
           "without_culling_gpairs_s": round(pairs / t_full / 1e9, 3),
           "speedup_from_culling": round(t_full / t_cull, 3),
           "counts_equal": bool(np.array_equal(c_cull, c_full)),
           "pairs_in_range": int(c_full.sum())}
    print(json.dumps(out))


def report_errors(args):
    """Per-bin error of the binary32 kernels against the binary64 kernel, per operator order:
    flat 2D (N = 65536, 64 log bins; L and T of orders 1..3) and the sphere (N = 262144; L2, T2,
    L3, T3).  For each bin with at least 1000 pairs: |s32 - s64| / |s64| (meaningful for the
    even, one-signed orders) and |s32 - s64| / sum ||du||^order (the magnitude scale, which does
    not cancel; from FullVector{order} in binary64)."""
    import numpy as np
    import torch

    from sfamd import _native, synthetic

    ctx = _native.Context(0)
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    rows = []
    cases = [("flat2 N=65536", synthetic.config2(), 0,
              [(0, 1, 0), (0, 2, 0), (0, 3, 0), (0, 0, 1), (0, 0, 2), (0, 0, 3)]),
             ("sphere N=262144", synthetic.config4()[:3], 1, [(0, 2, 0), (0, 0, 2), (0, 3, 0), (0, 0, 3)])]
    for name, (x, u, dedges), geom, ops in cases:
        N, D, W = x.shape[1], u.shape[0], x.shape[0]
        dx = torch.from_numpy(np.ascontiguousarray(x.T)).to(dev)
        du = torch.from_numpy(np.ascontiguousarray(u.T)).to(dev)
        nd = len(dedges) - 1
        sums = torch.zeros(nd, dtype=torch.float64, device=dev)
        counts = torch.zeros(nd, dtype=torch.int64, device=dev)

        def run(op, prec):
            sums.zero_()
            counts.zero_()
            ctx.hist_device([_device_problem(_native, dx.data_ptr(), du.data_ptr(), N, D, W, geom, op,
                                             dedges, None, prec)], sums.data_ptr(), counts.data_ptr(), stream)
            torch.cuda.synchronize()
            return sums.cpu().numpy().copy(), counts.cpu().numpy().copy()

        for op in ops:
            order = op[1] + op[2]
            s64, c64 = run(op, "fp64")
            s32, c32 = run(op, "mixed")
            mag, _ = run((5, order, 0), "fp64")
            big = c64 >= 1000
            d = np.abs(s32 - s64)[big]
            rel = d / np.maximum(np.abs(s64[big]), 1e-300)
            relm = d / np.maximum(mag[big], 1e-300)
            label = ("L" * op[1] + "T" * op[2]) if op[1] + op[2] > 0 else "1"
            rows.append({"case": name, "operator": f"{'L' if op[1] else ''}{op[1] or ''}"
                                                    f"{'T' if op[2] else ''}{op[2] or ''}",
                         "order": order, "counts_bit_exact": bool(np.array_equal(c32, c64)),
                         "bins": int(big.sum()),
                         "max_rel_to_sum": float(rel.max()), "median_rel_to_sum": float(np.median(rel)),
                         "max_rel_to_mag": float(relm.max()), "median_rel_to_mag": float(np.median(relm))})
            del label
    print(json.dumps({"report": "binary32 (mixed) vs binary64 per-bin error", "rows": rows}))


if __name__ == "__main__":
    main()
