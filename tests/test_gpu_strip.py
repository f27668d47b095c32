"""B-strips: slices that share positions (a field with auxiliary axes) and bin 1D histograms are
run by a kernel that evaluates each pair's geometry once for up to four velocity fields
(batch_leading.jl:105-140), opt-in through tuning variant 3.  Each field must still equal the
oracle's batch-leading evaluation of its own slice: counts bit-identical, sums within the mixed
bound."""
import numpy as np
import pytest

import _oracle
import sfamd
from sfamd import _native

pytestmark = pytest.mark.gpu

OPS = [(sfamd.L1SF, (0, 1, 0)), (sfamd.L2SF, (0, 2, 0)), (sfamd.S2SF, (1, 0, 0)), (sfamd.T2SF, (0, 0, 2))]


@pytest.fixture(autouse=True)
def strips_on():
    """B-strips are opt-in (tuning variant 3): the public API's context runs them for every test
    here, and goes back to the default afterwards."""
    ctx = sfamd.MI355XBackend().context()
    ctx.set_tuning(0, 0, 0, 3)
    yield
    ctx.set_tuning(0, 0, 0, 0)


def check_fields(r, x, u, e, op, geometry, prec, label):
    for b in range(u.shape[2]):
        s, c, a, g = _oracle.hist(x, u[:, :, b], e, op=op, geometry=geometry, nthreads=8,
                                  batch_leading=True, with_mag=True)
        assert np.array_equal(r.counts[:, b], c), (label, b)
        tol, mag = (1e-12, 0.0) if prec == "fp64" else (1e-6, 24 * 2.0 ** -24)
        assert np.all(np.abs(r.sums[:, b] - s) <= tol * a + (mag + 1e-14) * g + 1e-300), (label, b)


@pytest.mark.parametrize("sf,op", OPS)
@pytest.mark.parametrize("prec", ["mixed", "mixed_fixed"])
def test_strip_flat_lattice(sf, op, prec):
    """A grid (batch-leading rounding exercised) with six fields: a strip of four and one of two."""
    g = np.arange(28) * 0.1
    X, Y = np.meshgrid(g, g)
    x = np.stack([X.ravel(), Y.ravel()])
    u = np.random.default_rng(3).normal(size=(2, x.shape[1], 6))
    e = np.arange(1, 12) * 0.1
    r = sfamd.calculate_structure_function(sf, x, u, e, backend=sfamd.MI355XBackend(precision=prec),
                                           count_eltype=np.uint64,
                                           output_type=sfamd.StructureFunctionSumsAndCounts)
    check_fields(r, x, u, e, op, "flat", prec, f"flat lattice {op}")


@pytest.mark.parametrize("geometry,D,B", [("flat", 3, 4), ("deg", 2, 5), ("deg", 3, 3), ("rad", 2, 2)])
@pytest.mark.parametrize("sf,op", OPS)
def test_strip_geometries(geometry, D, B, sf, op):
    rng = np.random.default_rng(10 + D + B)
    N = 900
    if geometry == "flat":
        x = rng.uniform(0, 1, (D, N))
        e = np.geomspace(1e-2, 1.6, 14)
        metric = sfamd.Euclidean()
    else:
        lon = rng.uniform(-180, 180, N)
        lat = np.degrees(np.arcsin(rng.uniform(-1, 1, N)))
        x = np.stack([lon, lat]) if geometry == "deg" else np.radians(np.stack([lon, lat]))
        e = np.geomspace(1e-2, 3.1, 14)
        metric = sfamd.Haversine(1.0) if geometry == "deg" else sfamd.SphericalAngle()
    u = rng.normal(size=(D, N, B))
    r = sfamd.calculate_structure_function(sf, x, u, e, backend=sfamd.MI355XBackend(precision="mixed"),
                                           count_eltype=np.uint64, distance_metric=metric,
                                           output_type=sfamd.StructureFunctionSumsAndCounts)
    check_fields(r, x, u, e, op, geometry, "mixed", f"{geometry} D={D} B={B} {op}")


def test_strip_runs_one_item_set_per_strip():
    """Eight fields sharing x run as two strips: a quarter of the work items of the same problems
    with distinct position arrays, and the same results."""
    rng = np.random.default_rng(5)
    N = 3000
    x = rng.uniform(0, 1, (2, N))
    u = rng.normal(size=(2, N, 8))
    e = np.geomspace(1e-2, 1.4, 17)
    ctx = _native.Context(0)
    ctx.set_tuning(0, 0, 0, 3)
    keep = []

    def problems(shared):
        out = []
        xs = np.asfortranarray(x)
        for b in range(8):
            xb = xs if shared else np.asfortranarray(x.copy())
            ub = np.asfortranarray(u[:, :, b])
            keep.extend([xb, ub])
            p = _native.SfProblem()
            p.D, p.W, p.N = 2, 2, N
            p.x, p.u = xb.ctypes.data, ub.ctypes.data
            p.geometry, p.radius = 0, 1.0
            p.op, p.nl, p.nt = 0, 2, 0
            p.n_dist, p.dist_edges = len(e) - 1, e.ctypes.data
            p.n_val, p.val_edges = 0, None
            p.precision, p.count_bits, p.semantics = 1, 64, 1
            out.append(p)
        return out

    nd = len(e) - 1
    s1, c1 = np.zeros(8 * nd), np.zeros(8 * nd, dtype=np.uint64)
    ctx.hist(problems(True), s1, c1)
    items_strip = ctx.stats()["items"]
    s2, c2 = np.zeros(8 * nd), np.zeros(8 * nd, dtype=np.uint64)
    ctx.hist(problems(False), s2, c2)
    items_slices = ctx.stats()["items"]
    assert items_strip * 4 == items_slices, (items_strip, items_slices)
    assert np.array_equal(c1, c2)
    np.testing.assert_allclose(s1, s2, rtol=1e-12, atol=0)


def test_strip_plan_binary64_redo():
    """A plan whose re-bin queue overflows is redone by sf_plan_sync with the binary64 kernel, one
    slice at a time: on an integer lattice with an edge at every integer squared distance up to
    400, every pair within 20 is undecidable in binary32 (more than the 2^22 first capacity)."""
    import torch
    g = np.arange(100, dtype=np.float64)
    X, Y = np.meshgrid(g, g)
    x = np.stack([X.ravel(), Y.ravel()])
    N = x.shape[1]
    u = np.random.default_rng(9).normal(size=(2, N, 2))
    e = np.concatenate([[0.5], np.sqrt(np.arange(1, 401, dtype=np.float64))])
    nd = len(e) - 1
    dev = torch.device("cuda:0")
    dx = torch.from_numpy(np.ascontiguousarray(x.T)).to(dev)
    dus = [torch.from_numpy(np.ascontiguousarray(u[:, :, b].T)).to(dev) for b in range(2)]
    ctx = _native.Context(0)
    ctx.set_tuning(0, 0, 0, 3)
    probs = []
    for b in range(2):
        p = _native.SfProblem()
        p.D, p.W, p.N = 2, 2, N
        p.x, p.u = dx.data_ptr(), dus[b].data_ptr()
        p.geometry, p.radius = 0, 1.0
        p.op, p.nl, p.nt = 0, 1, 0
        p.n_dist, p.dist_edges = nd, e.ctypes.data
        p.n_val, p.val_edges = 0, None
        p.precision, p.count_bits, p.semantics = 1, 64, 1
        probs.append(p)
    plan = _native.Plan(ctx, probs)
    s = torch.zeros(2 * nd, dtype=torch.float64, device=dev)
    c = torch.zeros(2 * nd, dtype=torch.int64, device=dev)
    plan.execute(s.data_ptr(), c.data_ptr(), torch.cuda.current_stream().cuda_stream, graph=False)
    assert plan.sync()
    sh, ch = s.cpu().numpy(), c.cpu().numpy().astype(np.uint64)
    for b in range(2):
        os_, oc, oa = _oracle.hist(x, u[:, :, b], e, op=(0, 1, 0), nthreads=8, batch_leading=True)
        assert np.array_equal(ch[b * nd:(b + 1) * nd], oc), b
        assert np.all(np.abs(sh[b * nd:(b + 1) * nd] - os_) <= 1e-12 * oa + 1e-300), b
    plan.close()
