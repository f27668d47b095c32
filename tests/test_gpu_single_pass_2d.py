"""GPU parity of the single pass's 2D form (SF_OP_SINGLE_PASS with value bins) against the oracle.

calculate_structure_functions_single_pass_2d[!] (src/Calculations/serial_single_pass.jl:712-728,
1161-1181): every counted pair's six invariants |du|^2, du_L^2, du_T^2, du_L |du|^2, du_L^3,
du_L du_T^2 are each binned by distance and by their own value edges into a (6, n_dist, n_val)
accumulator.  Bar: counts bit-identical; sums within 1e-12 (binary64 kernel) or 1e-6 (mixed) of
sum |value| per cell, the mixed bound adding 24 u32 m^order per pair as in the single-operator
parity tests, with m = ||du|| (flat) or (||u_i|| + ||u_j||)(1 + c^2/sin sigma) on the sphere: the
geodesic frame transport mixes the two velocities, so binary32 errors there are absolute in the
velocities' size (a small |du|^2 between fast, far-apart points keeps them); the fixed-point cells
add at most 2^-23 of the value bin's width per pair.
"""
import ctypes

import numpy as np
import pytest

import _oracle
import sfamd
from sfamd import _native
from test_gpu_parity import CASES, MAG, TOL, field

pytestmark = pytest.mark.gpu

PREC = {"fp64": 0, "mixed": 1, "fixed": 2}


def value_edges(u, n=24):
    """Per-invariant edges from the field's scale: log-spaced (with 0) for the non-negative
    invariants, linear for the signed ones."""
    m = 2.0 * float(np.sum(np.var(u, axis=1)))
    pos = np.concatenate([[0.0], np.geomspace(1e-4 * m, 6.0 * m, n)])
    sgn = np.linspace(-3.0 * m ** 1.5, 3.0 * m ** 1.5, n + 1)
    return [pos, pos, pos, sgn, sgn, sgn]


def gpu_sp2d(ctx, x, u, dedges, vlist, geometry="flat", prec="fp64"):
    keep = []
    xc = np.asfortranarray(x, dtype=np.float64)
    uc = np.asfortranarray(u, dtype=np.float64)
    de = np.ascontiguousarray(dedges, dtype=np.float64)
    nv = len(vlist[0]) - 1
    ve6 = np.ascontiguousarray(np.concatenate([np.asarray(v, dtype=np.float64)[:nv + 1] for v in vlist]))
    keep += [xc, uc, de, ve6]
    p = _native.SfProblem()
    p.D, p.W, p.N = uc.shape[0], xc.shape[0], xc.shape[1]
    p.x, p.u = xc.ctypes.data, uc.ctypes.data
    p.geometry = _oracle.GEOM[geometry]
    p.radius = 1.0
    p.op, p.nl, p.nt = _native.OP_SINGLE_PASS, 0, 0
    p.n_dist, p.dist_edges = len(de) - 1, de.ctypes.data
    p.n_val, p.val_edges = nv, ve6.ctypes.data
    p.precision = PREC[prec]
    p.count_bits = 64
    nd = len(de) - 1
    s = np.zeros(6 * nd * nv)
    c = np.zeros(6 * nd * nv, dtype=np.uint64)
    ctx.hist([p], s, c)
    return s.reshape((6, nd, nv), order="F"), c.reshape((6, nd, nv), order="F")


def check2d(gs, gc, os_, oc, oa, og, prec, label, vlist=None):
    assert np.array_equal(gc, oc), f"{label}: counts differ at {np.argwhere(gc != oc)[:5].tolist()}"
    tol = TOL["mixed" if prec == "fixed" else prec]
    bound = tol * np.maximum(oa, 1e-300)
    if prec != "fp64":
        bound = bound + MAG["mixed"] * og
    if prec == "fixed":
        for t in range(6):
            w = np.diff(np.asarray(vlist[t], dtype=np.float64))[: gs.shape[2]]
            bound[t] += oc[t] * (2.0 ** -22) * w[None, :]
    err = np.abs(gs - os_)
    ratio = err / np.maximum(bound, 1e-300)
    worst = np.unravel_index(np.argmax(ratio), ratio.shape)
    assert np.all(err <= bound), (f"{label}: max err/bound {ratio[worst]:.3g} at (invariant, d, v) = "
                                  f"{tuple(int(k) for k in worst)}, count {int(oc[worst])}, "
                                  f"err {err[worst]:.3g}, sum {os_[worst]:.3g}, sum|v| {oa[worst]:.3g}, "
                                  f"per-plane max {[float(np.max(ratio[t])) for t in range(6)]}")


@pytest.mark.parametrize("geometry,D", CASES)
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_single_pass_2d_parity(ctx, geometry, D, prec):
    x, u, dedges = field(geometry, D, 700, seed=61 + D)
    vl = value_edges(u)
    os_, oc, oa, og = _oracle.single_pass_2d(x, u, dedges, vl, geometry=geometry, nthreads=8, with_mag=True)
    assert oc.sum() > 0
    gs, gc = gpu_sp2d(ctx, x, u, dedges, vl, geometry, prec)
    check2d(gs, gc, os_, oc, oa, og, prec, f"{geometry} D={D} {prec}")


@pytest.mark.parametrize("geometry,D", [("flat", 2), ("deg", 2)])
def test_single_pass_2d_fixed_point(ctx, geometry, D):
    """Uniform value edges in all six vectors: the planes run with fixed-point LDS cells."""
    x, u, dedges = field(geometry, D, 900, seed=71 + D)
    m = 2.0 * float(np.sum(np.var(u, axis=1)))
    pos = np.linspace(-0.1 * m, 5.0 * m, 33)
    sgn = np.linspace(-3.0 * m ** 1.5, 3.0 * m ** 1.5, 33)
    vl = [pos, pos, pos, sgn, sgn, sgn]
    os_, oc, oa, og = _oracle.single_pass_2d(x, u, dedges, vl, geometry=geometry, nthreads=8, with_mag=True)
    gs, gc = gpu_sp2d(ctx, x, u, dedges, vl, geometry, "fixed")
    check2d(gs, gc, os_, oc, oa, og, "fixed", f"{geometry} fixed", vlist=vl)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_single_pass_2d_marginals_match_1d(ctx, prec):
    """Value edges wide enough for every value: summing over value bins gives the 1D form's
    counts exactly (count row of every invariant) on the GPU too."""
    x, u, dedges = field("flat", 2, 800, seed=5)
    wide = np.array([-1e300, -1.0, 0.0, 0.5, 2.0, 1e300])
    gs, gc = gpu_sp2d(ctx, x, u, dedges, [wide] * 6, "flat", prec)
    from test_gpu_single_pass import gpu_sp
    s1, c1 = gpu_sp(ctx, x, u, dedges, "flat", prec)
    for t in range(6):
        assert np.array_equal(gc[t].sum(axis=1), c1[0])
    scale = np.abs(s1).max()
    np.testing.assert_allclose(gs.sum(axis=2)[[0, 1, 3, 4]], s1[[0, 1, 3, 4]], rtol=0, atol=1e-6 * scale)


def test_single_pass_2d_edge_cases(ctx):
    """Coincident points (r = 0 under a leading -Inf edge: du_L is NaN, so of the six values only
    |du|^2 is binned), a shared single edge vector, and value bins that leave most values out."""
    rng = np.random.default_rng(8)
    x = rng.uniform(0, 1, (2, 300))
    x[:, 150:160] = x[:, 140:150]              # ten exact duplicates
    u = rng.normal(size=(2, 300))
    dedges = np.concatenate([[-np.inf], np.geomspace(1e-3, 1.5, 12)])
    narrow = np.linspace(-0.5, 0.5, 9)
    for prec in ("fp64", "mixed"):
        os_, oc, oa, og = _oracle.single_pass_2d(x, u, dedges, narrow, nthreads=8, with_mag=True)
        gs, gc = gpu_sp2d(ctx, x, u, dedges, [narrow] * 6, "flat", prec)
        check2d(gs, gc, os_, oc, oa, og, prec, f"edge cases {prec}")


def test_single_pass_2d_batch_and_cull(ctx):
    """Ragged batch (two slices, own edges) and a one-shot call whose distance edges reach 2 % of
    the domain (tile culling active): each slice equals the oracle."""
    rng = np.random.default_rng(12)
    xs = [rng.uniform(0, 1, (2, n)) for n in (1500, 900)]
    us = [rng.normal(size=(2, n)) for n in (1500, 900)]
    de = [np.geomspace(1e-3, 0.02, 9), np.geomspace(2e-3, 0.02, 9)]
    vls = [value_edges(u, 16) for u in us]
    keep = []
    probs = []
    for x, u, d, vl in zip(xs, us, de, vls):
        xc, uc = np.asfortranarray(x), np.asfortranarray(u)
        ve6 = np.ascontiguousarray(np.concatenate(vl))
        keep += [xc, uc, d, ve6]
        p = _native.SfProblem()
        p.D, p.W, p.N = 2, 2, x.shape[1]
        p.x, p.u = xc.ctypes.data, uc.ctypes.data
        p.geometry, p.radius = 0, 1.0
        p.op, p.nl, p.nt = _native.OP_SINGLE_PASS, 0, 0
        p.n_dist, p.dist_edges = len(d) - 1, d.ctypes.data
        p.n_val, p.val_edges = 16, ve6.ctypes.data
        p.precision, p.count_bits = 1, 64
        probs.append(p)
    ncell = 6 * 8 * 16
    s = np.zeros(2 * ncell)
    c = np.zeros(2 * ncell, dtype=np.uint64)
    ctx.hist(probs, s, c)
    for k in range(2):
        os_, oc, oa, og = _oracle.single_pass_2d(xs[k], us[k], de[k], vls[k], nthreads=8, with_mag=True)
        gs = s[k * ncell:(k + 1) * ncell].reshape((6, 8, 16), order="F")
        gc = c[k * ncell:(k + 1) * ncell].reshape((6, 8, 16), order="F")
        check2d(gs, gc, os_, oc, oa, og, "mixed", f"batch slice {k}")
        assert oc.sum() > 0


def test_single_pass_2d_public_api(ctx):
    """sfamd.calculate_structure_functions_single_pass_2d returns the six invariants as
    StructureFunction2DSumsAndCounts (serial_single_pass.jl:1139-1150); the in-place form
    accumulates; an auxiliary axis with shared positions runs each field separately."""
    x, u, dedges = field("flat", 2, 500, seed=3)
    vl = value_edges(u, 12)
    res = sfamd.calculate_structure_functions_single_pass_2d(x, u, dedges, tuple(vl),
                                                             backend=sfamd.MI355XBackend(precision="fp64"),
                                                             count_eltype=np.uint64)
    assert list(res) == ["S2", "L2", "T2", "S3", "L3", "L1T2"]
    os_, oc, oa, og = _oracle.single_pass_2d(x, u, dedges, vl, nthreads=8, with_mag=True)
    for t, name in enumerate(res):
        r = res[name]
        assert isinstance(r, sfamd.StructureFunction2DSumsAndCounts)
        assert r.sums.shape == (len(dedges) - 1, 12)
        assert np.array_equal(r.counts, oc[t])
        assert np.all(np.abs(r.sums - os_[t]) <= 1e-12 * np.maximum(oa[t], 1e-300))
    sums = np.zeros((6, len(dedges) - 1, 12))
    counts = np.zeros((6, len(dedges) - 1, 12), dtype=np.uint64)
    for _ in range(2):
        sfamd.calculate_structure_functions_single_pass_2d_inplace(
            sums, counts, x, u, dedges, tuple(vl), backend=sfamd.MI355XBackend(precision="fp64"))
    assert np.array_equal(counts, 2 * oc)
    # auxiliary axis, shared positions: u (D, N, 2)
    u2 = np.stack([u, 0.5 * u[::-1]], axis=-1)
    res2 = sfamd.calculate_structure_functions_single_pass_2d(
        x, u2, dedges, vl[0], backend=sfamd.MI355XBackend(precision="fp64"), count_eltype=np.uint64)
    for b in range(2):
        ob, cb = _oracle.single_pass_2d(x, u2[..., b], dedges, vl[0], nthreads=8)
        for t, name in enumerate(res2):
            assert np.array_equal(res2[name].counts[..., b], cb[t])
    with pytest.raises(sfamd.DimensionMismatch):
        sfamd.calculate_structure_functions_single_pass_2d(x, u, dedges, (vl[0],) * 5 + (vl[0][:5],))


def test_single_pass_2d_queue_growth(ctx):
    """Log value bins reaching down to 1e-4 of the mean |du|^2 leave ~1 % of the sphere's plane
    pairs undecidable in binary32: more than the re-bin queue's first capacity (2^22).  The call
    grows the queue and reruns the binary32 kernel instead of redoing everything in binary64:
    counts equal the oracle, and the sums are the binary32 path's (not bit-identical to the
    binary64 kernel's)."""
    from sfamd import synthetic
    x, u, dedges, _ = synthetic.config4(N=12000)
    m = 2.0 * float(np.sum(np.var(u, axis=1)))
    pos = np.concatenate([[0.0], np.geomspace(1e-4 * m, 8.0 * m, 48)])
    sgn = np.linspace(-4.0 * m ** 1.5, 4.0 * m ** 1.5, 49)
    vl = [pos, pos, pos, sgn, sgn, sgn]
    gs, gc = gpu_sp2d(ctx, x, u, dedges, vl, "deg", "mixed")
    queued = ctx.stats()["rebinned"]
    assert queued > (1 << 22), queued
    os_, oc, oa, og = _oracle.single_pass_2d(x, u, dedges, vl, geometry="deg", nthreads=8, with_mag=True)
    check2d(gs, gc, os_, oc, oa, og, "mixed", "queue growth")
    s64, c64 = gpu_sp2d(ctx, x, u, dedges, vl, "deg", "fp64")
    assert np.array_equal(c64, gc)
    assert not np.array_equal(s64, gs)
