"""GPU: the public API end to end, edge cases, batches, device pointers, determinism, and
full-size invariants, checked against the CPU oracle (counts identical, sums to tolerance)."""
import math

import numpy as np
import pytest

import _oracle
import sfamd
from sfamd import _native, synthetic
from test_gpu_parity import MAG, TOL, check, gpu_hist

pytestmark = pytest.mark.gpu
EARTH_R = 6.371e6
PRECS = ["fp64", "mixed"]


def be(prec):
    return sfamd.MI355XBackend(0, prec)


@pytest.mark.parametrize("prec", PRECS)
def test_public_api_1d_mean_and_raw(prec):
    x, u, e = synthetic.config1(N=600)
    s, c, a = _oracle.hist(x, u, e, op=(0, 2, 0), nthreads=8)
    r = sfamd.calculate_structure_function(sfamd.L2SF, x, u, e, backend=be(prec))
    assert isinstance(r, sfamd.StructureFunction)
    mean = np.where(c > 0, s / np.maximum(c, 1), np.nan)
    np.testing.assert_allclose(r.values, mean, rtol=10 * TOL[prec], equal_nan=True)
    raw = sfamd.calculate_structure_function(sfamd.L2SF, x, u, e, backend=be(prec),
                                             output_type=sfamd.StructureFunctionSumsAndCounts)
    assert raw.counts.dtype == np.uint32 and np.array_equal(raw.counts, c.astype(np.uint32))


@pytest.mark.parametrize("prec", PRECS)
def test_joint_marginalises_to_1d(prec):
    x, u, e, ve = synthetic.config3(N=800)
    wide = np.array([-1e30, -1.0, 0.0, 1.0, 1e30])
    j = sfamd.calculate_structure_function(sfamd.L1SF, x, u, e, wide, backend=be(prec))
    one = sfamd.calculate_structure_function(sfamd.L1SF, x, u, e, backend=be(prec),
                                             output_type=sfamd.StructureFunctionSumsAndCounts)
    m = sfamd.marginalize(j)
    assert np.array_equal(m.counts, one.counts)
    np.testing.assert_allclose(m.sums, one.sums, rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("prec", PRECS)
def test_inplace_accumulates(prec):
    x, u, e = synthetic.config1(N=300)
    sums = np.zeros(len(e) - 1)
    counts = np.zeros(len(e) - 1, dtype=np.uint64)
    sfamd.calculate_structure_function_inplace(sums, counts, sfamd.L3SF, x, u, e, backend=be(prec))
    sfamd.calculate_structure_function_inplace(sums, counts, sfamd.L3SF, x, u, e, backend=be(prec))
    s, c, a = _oracle.hist(x, u, e, op=(0, 3, 0))
    assert np.array_equal(counts, 2 * c)
    assert np.all(np.abs(sums - 2 * s) <= 2 * TOL[prec] * a + 1e-300)
    # joint in-place, which the reference's GPU path does not support (dispatch.jl:455-459)
    ve = np.linspace(-2, 2, 9)
    s2 = np.zeros((len(e) - 1, 8))
    c2 = np.zeros((len(e) - 1, 8), dtype=np.uint64)
    sfamd.calculate_structure_function_inplace(s2, c2, sfamd.L1SF, x, u, e, ve, backend=be(prec))
    os_, oc, oa = _oracle.hist(x, u, e, op=(0, 1, 0), vedges=ve)
    assert np.array_equal(c2, oc)


@pytest.mark.parametrize("prec", PRECS)
def test_ragged_batch_matches_per_slice(prec):
    rng = np.random.default_rng(5)
    sizes = [0, 1, 2, 37, 500, 1500, 257]
    xs = [rng.uniform(0, 1, (2, n)) for n in sizes]
    us = [rng.normal(size=(2, n)) for n in sizes]
    e = np.geomspace(1e-3, 1.5, 20)
    ves = [np.linspace(-3 - k, 3 + k, 13) for k in range(len(sizes))]
    res = sfamd.calculate_structure_function_batch(sfamd.L1SF, xs, us, e, ves, backend=be(prec))
    for x, u, ve, r in zip(xs, us, ves, res):
        s, c, a, g = _oracle.hist(x, u, e, op=(0, 1, 0), vedges=ve, with_mag=True)
        check(r.sums, r.counts.astype(np.uint64), s, c, a, TOL[prec], f"N={x.shape[1]}", og=g, mag=MAG[prec])
    res1 = sfamd.calculate_structure_function_batch(sfamd.L2SF, xs, us, e, backend=be(prec))
    for x, u, r in zip(xs, us, res1):
        s, c, a = _oracle.hist(x, u, e, op=(0, 2, 0))
        assert np.array_equal(r.counts.astype(np.uint64), c)


def test_auxiliary_axes_match_per_slice():
    rng = np.random.default_rng(6)
    N = 300
    x = rng.uniform(0, 1, (2, N))
    u = rng.normal(size=(2, N, 3))
    e = np.geomspace(1e-2, 1.5, 10)
    r = sfamd.calculate_structure_function(sfamd.T2SF, x, u, e, backend=be("fp64"),
                                           output_type=sfamd.StructureFunctionSumsAndCounts)
    assert r.sums.shape == (9, 3)
    # fields with auxiliary axes follow the reference's batch-leading kernels
    for b in range(3):
        s, c, a, g = _oracle.hist(x, u[:, :, b], e, op=(0, 0, 2), batch_leading=True, with_mag=True)
        assert np.array_equal(r.counts[:, b].astype(np.uint64), c)
        assert np.all(np.abs(r.sums[:, b] - s) <= 1e-12 * a + 1e-14 * g + 1e-300)
    xv = rng.uniform(0, 1, (2, N, 2))
    uv = rng.normal(size=(2, N, 2))
    r = sfamd.calculate_structure_function(sfamd.L2SF, xv, uv, e, backend=be("mixed"),
                                           output_type=sfamd.StructureFunctionSumsAndCounts)
    for b in range(2):
        s, c, a = _oracle.hist(xv[:, :, b], uv[:, :, b], e, op=(0, 2, 0), batch_leading=True)
        assert np.array_equal(r.counts[:, b].astype(np.uint64), c)


EDGE_CASES = {}


def case(f):
    EDGE_CASES[f.__name__] = f
    return f


@case
def duplicates_flat():
    rng = np.random.default_rng(1)
    x = np.repeat(rng.uniform(0, 1, (2, 40)), 3, axis=1)
    u = rng.normal(size=(2, 120))
    return x, u, np.array([0.0, 1e-9, 0.1, 0.5, 2.0]), "flat", 1.0, None


@case
def lattice_on_edges():
    g = np.arange(8.0)
    X, Y = np.meshgrid(g, g)
    x = np.stack([X.ravel(), Y.ravel()])
    u = np.stack([np.sin(x[0]), np.cos(x[1])])
    return x, u, np.array([0.0, 1.0, 2.0, math.sqrt(5.0), 3.0, 5.0, 10.0]), "flat", 1.0, None


@case
def infpadded_flat():
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(0, 1, (2, 200)), np.zeros((2, 3))], axis=1)
    u = rng.normal(size=(2, 203))
    return x, u, np.array([-np.inf, 0.05, 0.3, np.inf]), "flat", 1.0, None


@case
def infpadded_sphere():
    rng = np.random.default_rng(3)
    x = np.stack([rng.uniform(-180, 180, 300), rng.uniform(-89, 89, 300)])
    u = rng.normal(size=(2, 300))
    return x, u, np.array([-np.inf, 0.1, 1.0, np.inf]), "deg", 1.0, None


@case
def signed_zero_value_edges():
    rng = np.random.default_rng(4)
    x = rng.uniform(0, 1, (2, 150))
    u = np.round(rng.normal(size=(2, 150)))  # many identical increments -> du_L = +-0.0 exactly
    return x, u, np.geomspace(1e-3, 1.5, 8), "flat", 1.0, np.array([-2.0, -0.0, 2.0])


@case
def nan_inputs():
    rng = np.random.default_rng(5)
    x = rng.uniform(0, 1, (2, 100))
    u = rng.normal(size=(2, 100))
    u[0, 7] = np.nan
    return x, u, np.geomspace(1e-3, 1.5, 8), "flat", 1.0, None


@case
def nan_inputs_joint():
    x, u, e, _, _, _ = nan_inputs()
    return x, u, e, "flat", 1.0, np.linspace(-3, 3, 7)


@case
def nonuniform_value_edges():
    rng = np.random.default_rng(6)
    x = rng.uniform(0, 1, (3, 200))
    u = rng.normal(size=(3, 200))
    v = np.geomspace(1e-3, 4, 9)
    return x, u, np.geomspace(1e-3, 2.0, 10), "flat", 1.0, np.concatenate([-v[::-1], v])


@case
def offset_coordinates():
    rng = np.random.default_rng(7)
    x = 1.0e4 + rng.uniform(0, 1, (2, 300))
    u = rng.normal(size=(2, 300))
    return x, u, np.geomspace(1e-3, 1.5, 12), "flat", 1.0, np.linspace(-3, 3, 13)


@case
def earth_radius_radians_joint():
    rng = np.random.default_rng(8)
    lon = rng.uniform(-np.pi, np.pi, 250)
    lat = np.arcsin(rng.uniform(-1, 1, 250))
    u = rng.normal(size=(2, 250))
    return np.stack([lon, lat]), u, np.geomspace(1e-4, np.pi, 16), "rad", 1.0, np.linspace(-4, 4, 17)


@case
def poles_and_antipodes():
    lon = np.array([0.0, 90.0, 180.0, -90.0, 0.0, 10.0, 0.0, 180.0, 45.0, -135.0])
    lat = np.array([90.0, 90.0, -90.0, -90.0, 0.0, 0.0, 0.0, 0.0, 30.0, -30.0])
    u = np.stack([np.linspace(-1, 1, 10), np.linspace(1, -1, 10)])
    return np.stack([lon, lat]), u, np.array([0.0, 0.01, 1.0, 3.0, math.pi, 4.0]), "deg", 1.0, None


@pytest.mark.parametrize("name", sorted(EDGE_CASES))
@pytest.mark.parametrize("prec", PRECS)
def test_edge_cases(ctx, name, prec):
    x, u, e, geom, radius, ve = EDGE_CASES[name]()
    for op in [(0, 1, 0), (0, 2, 0), (0, 0, 3), (1, 0, 0)]:
        s, c, a, g = _oracle.hist(x, u, e, op=op, geometry=geom, radius=radius, vedges=ve, with_mag=True)
        gs, gc = gpu_hist(ctx, x, u, e, op, geometry=geom, radius=radius, vedges=ve, prec=prec)
        check(gs, gc, s, c, a, TOL[prec], f"{name} op={op} {prec}", og=g, mag=MAG[prec])


def test_device_pointer_entry_matches_host(ctx):
    torch = pytest.importorskip("torch")
    x, u, e, ve = synthetic.config4(N=3000)
    hs, hc = gpu_hist(ctx, x, u, e, (0, 1, 0), geometry="deg", vedges=ve, prec="mixed")
    dev = torch.device("cuda:0")
    dx = torch.from_numpy(np.ascontiguousarray(x.T)).to(dev)
    du = torch.from_numpy(np.ascontiguousarray(u.T)).to(dev)
    nd, nv = len(e) - 1, len(ve) - 1
    sums = torch.zeros(nd * nv, dtype=torch.float64, device=dev)
    counts = torch.zeros(nd * nv, dtype=torch.int64, device=dev)
    p = _native.SfProblem()
    p.D, p.W, p.N = 2, 2, x.shape[1]
    p.x, p.u = dx.data_ptr(), du.data_ptr()
    p.geometry, p.radius, p.op, p.nl, p.nt = 1, 1.0, 0, 1, 0
    p.n_dist, p.dist_edges, p.n_val, p.val_edges = nd, e.ctypes.data, nv, ve.ctypes.data
    p.precision, p.count_bits = 1, 64
    for _ in range(2):  # accumulates
        ctx.hist_device([p], sums.data_ptr(), counts.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    c = counts.cpu().numpy().reshape((nd, nv), order="F").astype(np.uint64)
    s = sums.cpu().numpy().reshape((nd, nv), order="F")
    assert np.array_equal(c, 2 * hc)
    np.testing.assert_allclose(s, 2 * hs, rtol=1e-12, atol=1e-12)
    st = ctx.stats()
    assert st["pairs"] == x.shape[1] * (x.shape[1] - 1) // 2 and st["hist_ns"] > 0


@pytest.mark.parametrize("prec", PRECS)
def test_deterministic_counts_across_launch_shapes(ctx, prec):
    x, u, e, ve = synthetic.config4(N=2500)
    ref = None
    shapes = [(0, 0, 0, 0), (256, 0, 0, 0), (512, 256, 0, 0), (1024, 0, 1, 0), (64, 64, 0, 0)]
    if prec == "mixed":
        shapes += [(0, 0, 0, 1), (0, 0, 0, 2), (256, 512, 0, 1)]
    try:
        for blk, jc, bpc, var in shapes:
            ctx.set_tuning(blk, jc, bpc, var)
            s, c = gpu_hist(ctx, x, u, e, (0, 1, 0), geometry="deg", vedges=ve, prec=prec)
            if ref is None:
                ref = (s, c)
            assert np.array_equal(c, ref[1]), (blk, jc, bpc, var)
            np.testing.assert_allclose(s, ref[0], rtol=1e-11, atol=1e-11)
    finally:
        ctx.set_tuning(0, 0, 0, 0)


def test_full_size_mixed_counts_equal_fp64(ctx):
    """At the benchmark's full size (N = 262144 on the sphere, 3.4e10 pairs) the oracle is too slow
    to run here, so check size-independent properties: the binary32 kernel's counts equal the
    binary64 kernel's bit for bit; the joint histogram with catch-all value bins marginalises to
    the 1D histogram; with catch-all distance bins every pair is counted once."""
    x, u, e, ve = synthetic.config4()
    N = x.shape[1]
    s64, c64 = gpu_hist(ctx, x, u, e, (0, 1, 0), geometry="deg", vedges=ve, prec="fp64")
    s32, c32 = gpu_hist(ctx, x, u, e, (0, 1, 0), geometry="deg", vedges=ve, prec="mixed")
    assert np.array_equal(c32, c64)
    # every value of cell (k, m) lies in value bin m, so sum |v| <= count * max |edge of bin m|
    vmax = np.maximum(np.abs(ve[:-1]), np.abs(ve[1:]))[None, :]
    assert np.all(np.abs(s32 - s64) <= 1e-6 * c64 * vmax + 1e-300)
    l2_64, c2_64 = gpu_hist(ctx, x, u, e, (0, 2, 0), geometry="deg", prec="fp64")
    l2_32, c2_32 = gpu_hist(ctx, x, u, e, (0, 2, 0), geometry="deg", prec="mixed")
    assert np.array_equal(c2_32, c2_64)
    assert np.all(np.abs(l2_32 - l2_64) <= 1e-6 * np.abs(l2_64))
    wide = np.array([-np.inf, 0.0, np.inf])
    sj, cj = gpu_hist(ctx, x, u, e, (0, 2, 0), geometry="deg", vedges=wide, prec="mixed")
    s1, c1 = gpu_hist(ctx, x, u, e, (0, 2, 0), geometry="deg", prec="mixed")
    assert np.array_equal(cj.sum(axis=1), c1)
    ip = np.array([-np.inf, 1e-3, 1.0, np.inf])
    sa, ca = gpu_hist(ctx, x, u, ip, (0, 2, 0), geometry="deg", prec="mixed")
    assert int(ca.sum()) == N * (N - 1) // 2
