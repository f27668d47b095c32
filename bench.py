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

# FLOP-equivalents per pair (FMA = 2, add/mul/cmp/cvt = 1, binary32 transcendental = 4) of the
# algorithmic pair computation, per workload (derivation in DESIGN.md)
FLOPS_PER_PAIR = {"sphere_joint_L1": 60, "flat2_L2": 28}
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
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one device copy of 1 GiB (counter calibration for HBM traffic)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding "
                                   "is a design note only)"}))
        sys.exit(2)

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
        # du_L is odd: bins of a joint (distance, du_L) histogram hold values of one sign, so the
        # per-bin relative error |s32 - s64| / |s64| is well conditioned
        big = c64 >= 1000
        precs = [args.precision] + (["mixed"] if args.precision == "mixed_fixed" else [])
        for prec in precs:
            s32, c32 = result(prec)
            rel = np.abs(s32 - s64)[big] / np.maximum(np.abs(s64[big]), 1e-300)
            extra[f"{prec}_vs_fp64"] = {
                "counts_bit_exact": bool(np.array_equal(c32, c64)),
                "sums_max_rel_err": float(rel.max()) if rel.size else None,
                "sums_median_rel_err": float(np.median(rel)) if rel.size else None,
                "bins_compared": int(big.sum())}
        if args.precision == "mixed_fixed":
            extra["mixed_gpairs_s"] = round(rate("mixed"), 3)

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
        import _oracle
        threads = min(16, os.cpu_count() or 1)
        # estimate the rate on a small row sample, then time a sample of about cpu_seconds
        def time_rows(stride):
            t = time.perf_counter()
            _oracle.hist(x, u, dedges, op=(0, 1, 0), geometry="deg", vedges=vedges, nthreads=threads,
                         compensated=False, rows=(0, N, stride))
            return time.perf_counter() - t, sum(N - 1 - i for i in range(0, N, stride))
        t_est, p_est = time_rows(4096)
        rate = p_est / max(t_est, 1e-6)
        stride = max(1, int(pairs / max(rate * args.cpu_seconds, 1.0)))
        t_s, p_s = time_rows(stride)
        cpu = {"value": round(p_s / t_s / 1e9, 4), "unit": "Gpairs/s", "cores": threads, "kind": "port",
               "sample": f"every {stride}th outer row i of the same workload ({p_s} pairs), "
                         f"C++ restatement of SerialBackend/ThreadedBackend, binary64, {threads} threads"}

    out = {
        "metric": METRIC,
        "value": round(gps, 3), "unit": "Gpairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPE[args.precision],
        "data": "synthetic (seeded Fourier-mode field on the sphere)",
        "config": {"workload": f"sphere N={N} Haversine(1) joint du_L x distance 64x96, {args.precision}",
                   "pairs_per_step": pairs, "precision": args.precision, "rebinned_pairs": int(rebinned)},
        # compute-bound: binary32 vector arithmetic (no matrix instructions); 157.3 TFLOP/s is the
        # published fp32 peak, the same figure for the vector and matrix pipes
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak,
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


if __name__ == "__main__":
    main()
