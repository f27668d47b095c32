"""Fused outputs (sf_hist_multi / sf_hist_multi_device): several operators and at most one joint
histogram from one evaluation of every pair.

Each output must equal its own call: counts bit-identical to the oracle (and so to the separate
GPU call), sums within that precision's bar against the compensated binary64 oracle (the same
bounds as tests/test_gpu_parity.py: 1e-12 relative to sum |value| for binary64; 1e-6 plus the
binary32 rounding of the pair arithmetic for the mixed kernel; plus 3 2^-22 sum |value| for the
fixed-point joint cells).
"""
import numpy as np
import pytest

import _oracle
import sfamd
from sfamd import _native
from test_gpu_parity import MAG, TOL, field, gpu_hist

pytestmark = pytest.mark.gpu

PREC = {"fp64": _native.PREC_FP64, "mixed": _native.PREC_MIXED, "fixed": _native.PREC_MIXED_FIXED}
ONE_D = [(0, 2, 0), (0, 0, 2), (0, 3, 0), (0, 0, 3), (1, 0, 0), (5, 1, 0), (2, 0, 0), (0, 1, 2)]
JOINT = (0, 1, 0)


def _problems(slices, dedges, geometry, prec, keep):
    out = []
    for x, u in slices:
        xc = np.asfortranarray(x, dtype=np.float64)
        uc = np.asfortranarray(u, dtype=np.float64)
        keep += [xc, uc]
        p = _native.SfProblem()
        p.D, p.W, p.N = uc.shape[0], xc.shape[0], xc.shape[1]
        p.x, p.u = xc.ctypes.data, uc.ctypes.data
        p.geometry = _oracle.GEOM[geometry]
        p.radius = 1.0
        p.op, p.nl, p.nt = ONE_D[0]
        p.n_dist, p.dist_edges = len(dedges) - 1, dedges.ctypes.data
        p.precision = PREC[prec]
        p.count_bits = 64
        out.append(p)
    return out


def _outputs(ops, vedges, keep):
    outs = []
    for op in ops:
        o = _native.SfOutput()
        if op == "joint":
            o.op, o.nl, o.nt = JOINT
            o.n_val, o.val_edges = len(vedges) - 1, vedges.ctypes.data
        else:
            o.op, o.nl, o.nt = op
        outs.append(o)
    return outs


def fused(ctx, slices, dedges, ops, geometry, prec, vedges=None):
    """One sf_hist_multi call; returns per output (sums, counts) of shape (n_slices, ncell)."""
    keep = []
    de = np.ascontiguousarray(dedges, dtype=np.float64)
    ve = None if vedges is None else np.ascontiguousarray(vedges, dtype=np.float64)
    keep += [de, ve]
    probs = _problems(slices, de, geometry, prec, keep)
    outs = _outputs(ops, ve, keep)
    nd = len(de) - 1
    ncell = [nd * (len(ve) - 1 if op == "joint" else 1) for op in ops]
    sums = [np.zeros(len(slices) * c) for c in ncell]
    counts = [np.zeros(len(slices) * c, dtype=np.uint64) for c in ncell]
    ctx.hist_multi(probs, outs, sums, counts)
    return [(s.reshape(len(slices), -1), c.reshape(len(slices), -1)) for s, c in zip(sums, counts)]


def _check(gs, gc, x, u, dedges, op, geometry, prec, vedges, label):
    os_, oc, oa, og = _oracle.hist(x, u, dedges, op=op, geometry=geometry, vedges=vedges, nthreads=8,
                                   with_mag=True)
    os_, oc, oa, og = (a.reshape(-1, order="F") for a in (os_, oc, oa, og))
    assert np.array_equal(gc, oc), f"{label}: counts differ at {np.argwhere(gc != oc)[:5].tolist()}"
    nan_o = np.isnan(os_)
    assert np.array_equal(np.isnan(gs), nan_o), f"{label}: NaN pattern differs"
    ok = ~nan_o
    tol = TOL["fp64" if prec == "fp64" else "mixed"]
    bound = tol * np.maximum(oa, 1e-300) + MAG["fp64" if prec == "fp64" else "mixed"] * og
    if prec == "fixed" and vedges is not None:
        bound = bound + 3.0 * 2.0 ** -22 * np.maximum(oa, 1e-300)
    err = np.abs(gs - os_)
    assert np.all(err[ok] <= bound[ok]), f"{label}: max excess {np.max((err - bound)[ok]):.3g}"


@pytest.mark.parametrize("geometry,D", [("flat", 2), ("flat", 3), ("deg", 2), ("rad", 2), ("deg", 3)])
@pytest.mark.parametrize("prec", ["fp64", "mixed", "fixed"])
def test_fused_matches_oracle(ctx, geometry, D, prec):
    """Eight 1D operators and the joint du_L histogram in one call."""
    x, u, dedges = field(geometry, D, 700, seed=71 + D)
    ve = np.linspace(-3, 3, 31)
    ops = ONE_D + ["joint"]
    res = fused(ctx, [(x, u)], dedges, ops, geometry, prec, ve)
    for op, (gs, gc) in zip(ops, res):
        oop = JOINT if op == "joint" else op
        _check(gs[0], gc[0], x, u, dedges, oop, geometry, prec, ve if op == "joint" else None,
               f"{geometry} D={D} {prec} {op}")


@pytest.mark.parametrize("prec", ["mixed", "fixed"])
def test_fused_equals_separate_calls(ctx, prec):
    """Counts of each fused output are those of its own call, cell for cell."""
    x, u, dedges = field("deg", 2, 1500, seed=5)
    ve = np.linspace(-3, 3, 49)
    ops = [(0, 2, 0), (0, 0, 2), "joint"]
    res = fused(ctx, [(x, u)], dedges, ops, "deg", prec, ve)
    for op, (gs, gc) in zip(ops, res):
        if op == "joint":
            ss, sc = gpu_hist(ctx, x, u, dedges, JOINT, geometry="deg", vedges=ve, prec=prec)
            ss, sc = ss.reshape(-1, order="F"), sc.reshape(-1, order="F")
        else:
            ss, sc = gpu_hist(ctx, x, u, dedges, op, geometry="deg", prec=prec)
        assert np.array_equal(gc[0], sc), f"{op}: counts differ from the separate call"
        scale = np.maximum(np.abs(ss), 1e-300)
        assert np.all(np.abs(gs[0] - ss) <= 1e-5 * scale + 1e-12), f"{op}: sums differ from the separate call"


def test_fused_ragged_batch(ctx):
    """Slices of N = 0, 1, 300 and 900 in one fused launch: output k of slice s at s * ncell_k."""
    rng = np.random.default_rng(9)
    slices = []
    for n in (0, 1, 300, 900):
        x, u, dedges = field("flat", 2, max(n, 4), seed=int(rng.integers(1 << 30)))
        slices.append((x[:, :n], u[:, :n]))
    ve = np.linspace(-3, 3, 25)
    ops = [(0, 2, 0), (0, 3, 0), "joint"]
    res = fused(ctx, slices, dedges, ops, "flat", "mixed", ve)
    for k, op in enumerate(ops):
        for s, (x, u) in enumerate(slices):
            gs, gc = res[k][0][s], res[k][1][s]
            if x.shape[1] < 2:
                assert not gc.any() and not gs.any()
                continue
            oop = JOINT if op == "joint" else op
            _check(gs, gc, x, u, dedges, oop, "flat", "mixed", ve if op == "joint" else None,
                   f"slice {s} {op}")


def test_fused_only_1d_outputs(ctx):
    x, u, dedges = field("deg", 2, 800, seed=3)
    ops = [(0, 2, 0), (0, 0, 2)]
    res = fused(ctx, [(x, u)], dedges, ops, "deg", "mixed")
    for op, (gs, gc) in zip(ops, res):
        _check(gs[0], gc[0], x, u, dedges, op, "deg", "mixed", None, f"1D only {op}")


def test_fused_public_api_aux_axes():
    """calculate_structure_functions_fused equals calculate_structure_function per operator,
    for a field with an auxiliary axis (three time slices sharing positions)."""
    rng = np.random.default_rng(21)
    N, B = 600, 3
    x = rng.uniform(0, 1, (2, N))
    u = rng.normal(size=(2, N, B))
    bins = sfamd.LogBinEdges(np.geomspace(1e-3, 1.5, 21))
    vb = sfamd.LinearBinEdges(-4.0, 4.0, 32)
    be = sfamd.MI355XBackend(precision="mixed")
    ops = [sfamd.L2SF, sfamd.T2SF, (sfamd.L1SF, vb)]
    got = sfamd.calculate_structure_functions_fused(ops, x, u, bins, np.uint64, backend=be,
                                                    output_type=sfamd.StructureFunctionSumsAndCounts)
    for e, g in zip(ops, got):
        sf, v = e if isinstance(e, tuple) else (e, None)
        if v is None:
            want = sfamd.calculate_structure_function(sf, x, u, bins, np.uint64, backend=be,
                                                      output_type=sfamd.StructureFunctionSumsAndCounts)
        else:
            want = sfamd.calculate_structure_function(sf, x, u, bins, v, np.uint64, backend=be)
        assert g.counts.shape == want.counts.shape
        assert np.array_equal(g.counts, want.counts)
        assert np.allclose(g.sums, want.sums, rtol=1e-5, atol=1e-9)


def test_fused_rejects(ctx):
    x, u, dedges = field("flat", 2, 50, seed=1)
    with pytest.raises(sfamd.ArgumentError):
        fused(ctx, [(x, u)], dedges, ONE_D + [(0, 1, 1)], "flat", "mixed")  # nine 1D outputs
    ve = np.linspace(-1, 1, 5)
    with pytest.raises(sfamd.ArgumentError):
        fused(ctx, [(x, u)], dedges, ["joint", "joint"], "flat", "mixed", ve)
    with pytest.raises(sfamd.ArgumentError):
        fused(ctx, [(x, u)], dedges, [(_native.OP_SINGLE_PASS, 0, 0)], "flat", "mixed")


@pytest.mark.parametrize("prec", ["mixed", "fixed"])
def test_fused_plan_graph_replays(ctx, prec):
    """A fused plan (du_L^2 1D and the joint du_L histogram of ragged 2D slices, the batched
    configuration's outputs) replayed as one hipGraph after the fields were rewritten in place and
    after unrelated eager work: every replay equals the fused eager call on the same data, and the
    first the oracle."""
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(17)
    sizes = [900, 1, 1500, 64]
    e = np.geomspace(1e-3, 1.5, 25)
    ve = np.linspace(-3, 3, 33)
    nd, nv = len(e) - 1, len(ve) - 1
    xs = [torch.zeros((n, 2), dtype=torch.float64, device=dev) for n in sizes]
    us = [torch.zeros((n, 2), dtype=torch.float64, device=dev) for n in sizes]

    def problems():
        out = []
        for x, u in zip(xs, us):
            p = _native.SfProblem()
            p.D, p.W, p.N = 2, 2, x.shape[0]
            p.x, p.u = x.data_ptr(), u.data_ptr()
            p.geometry, p.radius, p.op, p.nl, p.nt = 0, 1.0, 0, 2, 0
            p.n_dist, p.dist_edges = nd, e.ctypes.data
            p.precision, p.count_bits = PREC[prec], 64
            out.append(p)
        return out

    keep = []
    outs = _outputs([(0, 2, 0), "joint"], ve, keep)
    plan = _native.Plan(ctx, problems(), outs)
    stream = torch.cuda.current_stream().cuda_stream
    B = len(sizes)
    bufs = [(torch.zeros(B * nd, dtype=torch.float64, device=dev), torch.zeros(B * nd, dtype=torch.int64, device=dev)),
            (torch.zeros(B * nd * nv, dtype=torch.float64, device=dev),
             torch.zeros(B * nd * nv, dtype=torch.int64, device=dev))]
    ebufs = [(torch.zeros_like(a), torch.zeros_like(b)) for a, b in bufs]
    for rep in range(3):
        hx = [rng.uniform(0, 1, (n, 2)) for n in sizes]
        hu = [rng.normal(size=(n, 2)) for n in sizes]
        for x, u, a, b in zip(xs, us, hx, hu):
            x.copy_(torch.from_numpy(a))
            u.copy_(torch.from_numpy(b))
        junk = torch.randn(1 << 20, device=dev).sum()
        for a, b in bufs + ebufs:
            a.zero_()
            b.zero_()
        plan.execute_multi([a.data_ptr() for a, _ in bufs], [b.data_ptr() for _, b in bufs], stream, graph=True)
        assert not plan.sync()
        ctx.hist_multi_device(problems(), outs, [a.data_ptr() for a, _ in ebufs], [b.data_ptr() for _, b in ebufs],
                              stream)
        torch.cuda.synchronize()
        for (gs, gc), (es, ec) in zip(bufs, ebufs):
            assert np.array_equal(gc.cpu().numpy(), ec.cpu().numpy()), rep
            np.testing.assert_allclose(gs.cpu().numpy(), es.cpu().numpy(), rtol=1e-12, atol=1e-12)
        if rep == 0:
            for k, n in enumerate(sizes):
                if n < 2:
                    continue
                for o, (op, vv, nc) in enumerate([((0, 2, 0), None, nd), (JOINT, ve, nd * nv)]):
                    gs = bufs[o][0].cpu().numpy()[k * nc:(k + 1) * nc]
                    gc = bufs[o][1].cpu().numpy()[k * nc:(k + 1) * nc].astype(np.uint64)
                    _check(gs, gc, hx[k].T, hu[k].T, e, op, "flat", prec, vv, f"replay slice {k} output {o}")
        float(junk)
    plan.close()


def test_fused_plan_rejects(ctx):
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    x = torch.rand((100, 2), dtype=torch.float64, device=dev)
    e = np.geomspace(1e-3, 1.5, 9)
    for prec, ops in [("fp64", [(0, 2, 0)]), ("mixed", [(0, 2, 0), (5, 1, 0)])]:
        p = _native.SfProblem()
        p.D, p.W, p.N = 2, 2, 100
        p.x, p.u = x.data_ptr(), x.data_ptr()
        p.geometry, p.radius, p.op, p.nl, p.nt = 0, 1.0, 0, 2, 0
        p.n_dist, p.dist_edges = len(e) - 1, e.ctypes.data
        p.precision, p.count_bits = PREC[prec], 64
        keep = []
        with pytest.raises(sfamd.ArgumentError):
            _native.Plan(ctx, [p], _outputs(ops, None, keep))
