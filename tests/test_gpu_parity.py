"""GPU parity: libsfmi355x.so (through its C ABI) against the CPU oracle.

Bar (from the reference's own GPU-vs-serial tests, test/test_gpu_parity.jl:37-38 and
test/test_gpu_tiled_parity.jl): counts bit-identical; sums within 1e-12 (binary64 kernel) or
1e-6 (mixed kernel) of the compensated binary64 oracle, relative to sum |value| per bin (odd
orders cancel, so |sum| is not a usable scale).  The mixed bound adds the binary32 rounding of
the pair arithmetic itself, 24 u32 * ||du||^order per pair (u32 = 2^-24): a value that is small
only through cancellation (du nearly perpendicular to r for du_L) keeps that absolute error.
"""
import numpy as np
import pytest

import _oracle
from sfamd import _native, synthetic

pytestmark = pytest.mark.gpu

OPS = [(0, 1, 0), (0, 2, 0), (0, 3, 0), (0, 0, 1), (0, 0, 2), (0, 0, 3), (0, 2, 1), (0, 1, 2),
       (0, 0, 0), (1, 0, 0), (2, 0, 0), (3, 0, 0), (4, 0, 0), (5, 1, 0), (5, 3, 0)]
TOL = {"fp64": 1e-12, "mixed": 1e-6}
MAG = {"fp64": 0.0, "mixed": 24 * 2.0 ** -24}


def gpu_hist(ctx, x, u, dedges, op, geometry="flat", radius=1.0, vedges=None, prec="fp64"):
    keep = []
    xc = np.asfortranarray(x, dtype=np.float64)
    uc = np.asfortranarray(u, dtype=np.float64)
    de = np.ascontiguousarray(dedges, dtype=np.float64)
    keep += [xc, uc, de]
    p = _native.SfProblem()
    p.D, p.W, p.N = uc.shape[0], xc.shape[0], xc.shape[1]
    p.x, p.u = xc.ctypes.data, uc.ctypes.data
    p.geometry = _oracle.GEOM[geometry]
    p.radius = radius
    p.op, p.nl, p.nt = op
    p.n_dist, p.dist_edges = len(de) - 1, de.ctypes.data
    if vedges is None:
        p.n_val, p.val_edges = 0, None
        ncell = len(de) - 1
    else:
        ve = np.ascontiguousarray(vedges, dtype=np.float64)
        keep.append(ve)
        p.n_val, p.val_edges = len(ve) - 1, ve.ctypes.data
        ncell = (len(de) - 1) * (len(ve) - 1)
    p.precision = {"fp64": 0, "mixed": 1, "fixed": 2}[prec]
    p.count_bits = 64
    s = np.zeros(ncell)
    c = np.zeros(ncell, dtype=np.uint64)
    ctx.hist([p], s, c)
    if vedges is not None:
        nd, nv = len(de) - 1, len(vedges) - 1
        return s.reshape((nd, nv), order="F"), c.reshape((nd, nv), order="F")
    return s, c


def check(gs, gc, os_, oc, oa, tol, label="", og=None, mag=0.0):
    assert np.array_equal(gc, oc), f"{label}: counts differ at {np.argwhere(gc != oc)[:5].tolist()}"
    nan_o = np.isnan(os_)
    assert np.array_equal(np.isnan(gs), nan_o), f"{label}: NaN pattern differs"
    ok = ~nan_o
    err = np.abs(gs[ok] - os_[ok])
    bound = tol * np.maximum(oa[ok], 1e-300)
    if og is not None:
        bound = bound + mag * og[ok]
    assert np.all(err <= bound), f"{label}: max rel err {np.max(err / np.maximum(oa[ok], 1e-300)):.3g}"


def field(geometry, D, N, seed):
    rng = np.random.default_rng(seed)
    if geometry == "flat":
        x = rng.uniform(0, 1, (D, N))
        u = synthetic._flat_field(rng, x, 32, 2 * np.pi, 2 * np.pi * 32, -5 / 6)
        return x, u, np.geomspace(1e-3, np.sqrt(D) * 1.01, 25)
    xd, _ = synthetic.sphere_points(rng, N, n_tracks=2, per_track=min(256, N // 4))
    x = xd if geometry == "deg" else np.radians(xd)
    u = rng.normal(size=(D, N))
    return x, u, np.geomspace(1e-4, np.pi, 25)


CASES = [("flat", 2), ("flat", 3), ("deg", 2), ("rad", 2), ("deg", 3)]


def test_device_math_bitexact(ctx):
    """sincosd / sincos / atan2 / sqrt / divide on gfx950 equal the host's bit for bit."""
    rng = np.random.default_rng(3)
    n = 200000
    a = np.concatenate([rng.uniform(-720, 720, n), [0.0, -0.0, 90.0, -90, 180, -180, 270, 360, 45, 30, 1e-300, 1e300]])
    out = ctx.eval_math(0, a)
    ref = np.array([_oracle.sincosd(v) for v in a])
    assert np.array_equal(out.view(np.uint64), ref.view(np.uint64))
    r = np.concatenate([rng.uniform(-10, 10, n), rng.uniform(-1e5, 1e5, 1000), [0.0, -0.0, 1e-300]])
    out = ctx.eval_math(1, r)
    ref = np.array([_oracle.sincos(v) for v in r])
    assert np.array_equal(out.view(np.uint64), ref.view(np.uint64))
    y = np.abs(rng.normal(size=n)) * 10.0 ** rng.uniform(-12, 1, n)
    xx = np.abs(rng.normal(size=n)) * 10.0 ** rng.uniform(-12, 1, n)
    y[::3] *= -1
    xx[::5] *= -1
    out = ctx.eval_math(2, y, xx)
    ref = np.array([_oracle.atan2(a_, b_) for a_, b_ in zip(y, xx)])
    assert np.array_equal(out.view(np.uint64), ref.view(np.uint64))
    q = np.abs(rng.normal(size=n)) * 10.0 ** rng.uniform(-300, 300, n)
    assert np.array_equal(ctx.eval_math(3, q).view(np.uint64), np.sqrt(q).view(np.uint64))
    d = rng.normal(size=n) * 10.0 ** rng.uniform(-100, 100, n)
    assert np.array_equal(ctx.eval_math(4, q, d).view(np.uint64), (q / d).view(np.uint64))


@pytest.mark.parametrize("geometry,D", CASES)
@pytest.mark.parametrize("joint", [False, True])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_parity_all_operators(ctx, geometry, D, joint, prec):
    N = 700
    x, u, dedges = field(geometry, D, N, seed=11 + D)
    vedges = np.linspace(-3, 3, 31) if joint else None
    for op in OPS:
        if joint and op[0] == 0 and op[1] + op[2] >= 2:
            ve = np.linspace(-8, 8, 41) if op[1] % 2 or op[2] % 2 else np.linspace(0, 12, 37)
        else:
            ve = vedges
        os_, oc, oa, og = _oracle.hist(x, u, dedges, op=op, geometry=geometry, vedges=ve, nthreads=8,
                                       with_mag=True)
        gs, gc = gpu_hist(ctx, x, u, dedges, op, geometry=geometry, vedges=ve, prec=prec)
        check(gs, gc, os_, oc, oa, TOL[prec], f"{geometry} D={D} op={op} joint={joint} {prec}",
              og=og, mag=MAG[prec])


@pytest.mark.parametrize("prec", ["mixed", "fixed"])
@pytest.mark.parametrize("geometry", ["deg", "flat"])
def test_nearly_coincident_pairs(ctx, geometry, prec):
    """Pairs 1e-10 .. 1e-5 apart (sphere: radians; flat: in the unit square), binned down to
    1e-11 against 800 value bins.  Below a separation of 2^-22 the binary32 records' hi + lo
    positions (2^-48 relative) would turn the direction by up to 3.5e-5 rad, so those pairs are
    re-binned in binary64: counts bit-identical and sums within the mixed bar (which a pair
    evaluated in binary32 at 1e-10 would exceed 30-fold)."""
    rng = np.random.default_rng(41)
    m = 1500
    sep = 10.0 ** rng.uniform(-10, -5, m)
    ang = rng.uniform(0, 2 * np.pi, m)
    if geometry == "deg":
        lon = rng.uniform(-180, 180, m)
        lat = np.degrees(np.arcsin(rng.uniform(-0.95, 0.95, m)))
        near = np.stack([lon + np.degrees(sep * np.cos(ang)) / np.cos(np.radians(lat)),
                         lat + np.degrees(sep * np.sin(ang))])
        x = np.concatenate([np.stack([lon, lat]), near], axis=1)
    else:
        base = rng.uniform(0, 1, (2, m))
        x = np.concatenate([base, base + sep * np.stack([np.cos(ang), np.sin(ang)])], axis=1)
    u = rng.normal(size=(2, 2 * m))
    dedges = np.geomspace(1e-11, 1e-4, 8)
    ve = np.linspace(-4, 4, 801)
    os_, oc, oa, og = _oracle.hist(x, u, dedges, op=(0, 1, 0), geometry=geometry, vedges=ve, nthreads=8,
                                   with_mag=True)
    gs, gc = gpu_hist(ctx, x, u, dedges, (0, 1, 0), geometry=geometry, vedges=ve, prec=prec)
    assert oc.sum() >= 0.98 * m, "the close pairs lie in the binned range (a few values lie outside)"
    assert np.array_equal(gc, oc), f"{geometry} {prec}: counts differ"
    ok = ~np.isnan(os_)
    err = np.abs(gs - os_)[ok]
    bound = TOL["mixed"] * np.maximum(oa, 1e-300) + MAG["mixed"] * og
    if prec == "fixed":
        bound = bound + 1.5 * 2.0 ** -23 * (ve[1] - ve[0]) * oc.astype(float)
    assert np.all(err <= bound[ok]), f"{geometry} {prec}: max excess {np.max(err - bound[ok]):.3g}"


@pytest.mark.parametrize("geometry,D", CASES)
def test_parity_fixed_point_joint(ctx, geometry, D):
    """SF_PREC_MIXED_FIXED: joint cells of uniform value bins as 64-bit fixed-point words.
    Counts bit-identical; each pair adds at most 0.75 2^-22 value-bin widths of absolute error on
    top of the mixed kernel's bound, and only values of at least a quarter bin width enter the
    fixed-point words (smaller ones go to binary64 cells), so that error is at most 3 2^-22 |v|."""
    x, u, dedges = field(geometry, D, 700, seed=31 + D)
    for op in [(0, 1, 0), (0, 2, 0), (0, 0, 3), (1, 0, 0), (5, 1, 0)]:
        ve = np.linspace(-3, 3, 31) if op in ((0, 1, 0), (0, 0, 3)) else np.linspace(0, 12, 37)
        os_, oc, oa, og = _oracle.hist(x, u, dedges, op=op, geometry=geometry, vedges=ve, nthreads=8,
                                       with_mag=True)
        gs, gc = gpu_hist(ctx, x, u, dedges, op, geometry=geometry, vedges=ve, prec="fixed")
        assert np.array_equal(gc, oc), f"{geometry} D={D} op={op}: counts differ"
        ok = ~np.isnan(os_)
        err = np.abs(gs - os_)[ok]
        bound = (TOL["mixed"] * np.maximum(oa, 1e-300) + MAG["mixed"] * og
                 + 3.0 * 2.0 ** -22 * np.maximum(oa, 1e-300))[ok]
        assert np.all(err <= bound), f"{geometry} D={D} op={op}: max excess {np.max(err - bound):.3g}"
