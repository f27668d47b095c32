"""GPU parity of the single-pass invariants (SF_OP_SINGLE_PASS) against the oracle.

calculate_structure_functions_single_pass[!] (src/Calculations/serial_single_pass.jl:404-421,
602-700): one pair pass, rows S2, L2, T2, S3, L3, L1T2 of a (6, n_bins) accumulator, counts shared.
Bar: counts bit-identical; each row within the tolerance of the single-operator parity tests,
scaled by sum |value| of that row's operator (for the differenced rows T2 and L1T2, of the two
rows they are differenced from).
"""
import numpy as np
import pytest

import _oracle
import sfamd
from sfamd import _native
from test_gpu_parity import CASES, MAG, TOL, check, field

pytestmark = pytest.mark.gpu

ROW_OPS = [(1, 0, 0), (0, 2, 0), (0, 0, 2), (2, 0, 0), (0, 3, 0), (0, 1, 2)]


def problem(x, u, dedges, geometry, prec, keep):
    xc = np.asfortranarray(x, dtype=np.float64)
    uc = np.asfortranarray(u, dtype=np.float64)
    de = np.ascontiguousarray(dedges, dtype=np.float64)
    keep += [xc, uc, de]
    p = _native.SfProblem()
    p.D, p.W, p.N = uc.shape[0], xc.shape[0], xc.shape[1]
    p.x, p.u = xc.ctypes.data, uc.ctypes.data
    p.geometry = _oracle.GEOM[geometry]
    p.radius = 1.0
    p.op, p.nl, p.nt = _native.OP_SINGLE_PASS, 0, 0
    p.n_dist, p.dist_edges = len(de) - 1, de.ctypes.data
    p.n_val, p.val_edges = 0, None
    p.precision = 0 if prec == "fp64" else 1
    p.count_bits = 64
    return p


def gpu_sp(ctx, x, u, dedges, geometry="flat", prec="fp64"):
    keep = []
    p = problem(x, u, dedges, geometry, prec, keep)
    nd = len(dedges) - 1
    s = np.zeros(6 * nd)
    c = np.zeros(6 * nd, dtype=np.uint64)
    ctx.hist([p], s, c)
    return s.reshape((6, nd), order="F"), c.reshape((6, nd), order="F")


def row_scales(x, u, dedges, geometry):
    """sum |value| and sum ||du||^order per bin for each row's operator (oracle)."""
    oa, og = [], []
    for op in ROW_OPS:
        _, _, a, g = _oracle.hist(x, u, dedges, op=op, geometry=geometry, nthreads=8, with_mag=True)
        oa.append(a)
        og.append(g)
    # differenced rows: the scale of the two rows they come from
    oa[2] = oa[0] + oa[1]
    og[2] = og[0] + og[1]
    oa[5] = oa[3] + oa[4]
    og[5] = og[3] + og[4]
    return oa, og


def check_sp(gs, gc, os_, oc, oa, og, prec, label):
    for r in range(6):
        check(gs[r], gc[r], os_[r], oc[r], oa[r], TOL[prec], f"{label} row {r}", og=og[r], mag=MAG[prec])


@pytest.mark.parametrize("geometry,D", CASES)
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_single_pass_parity(ctx, geometry, D, prec):
    x, u, dedges = field(geometry, D, 700, seed=41 + D)
    os_, oc = _oracle.single_pass(x, u, dedges, geometry=geometry, nthreads=8)
    gs, gc = gpu_sp(ctx, x, u, dedges, geometry, prec)
    oa, og = row_scales(x, u, dedges, geometry)
    check_sp(gs, gc, os_, oc, oa, og, prec, f"{geometry} D={D} {prec}")
    for r in range(1, 6):
        assert np.array_equal(gc[r], gc[0])
    assert np.array_equal(gs[2], gs[0] - gs[1])
    assert np.array_equal(gs[5], gs[3] - gs[4])


def test_single_pass_rows_match_single_operator_calls(ctx):
    """Each accumulated row equals the single-operator kernel on the same pairs (same bins,
    same counts; sums to binary64 summation-order differences)."""
    x, u, dedges = field("flat", 2, 900, seed=5)
    gs, gc = gpu_sp(ctx, x, u, dedges)
    from test_gpu_parity import gpu_hist
    for r, op in enumerate(ROW_OPS):
        s1, c1 = gpu_hist(ctx, x, u, dedges, op)
        assert np.array_equal(gc[r], c1)
        _, _, oa = _oracle.hist(x, u, dedges, op=op, nthreads=8)
        scale = oa if r not in (2, 5) else None
        if scale is not None:
            assert np.all(np.abs(gs[r] - s1) <= 1e-12 * np.maximum(scale, 1e-300))


def test_single_pass_edge_cases(ctx):
    """N = 0, 1, 2; coincident points under a leading -Inf edge (r = 0 is binned, du_L is NaN in
    the reference: 0 * inv(0)); an empty in-range set."""
    rng = np.random.default_rng(8)
    edges = np.array([0.1, 0.2, 0.4])
    for N in (0, 1, 2):
        x = rng.uniform(0, 1, (2, N))
        u = rng.normal(size=(2, N))
        os_, oc = _oracle.single_pass(x, u, edges)
        for prec in ("fp64", "mixed"):
            gs, gc = gpu_sp(ctx, x, u, edges, prec=prec)
            assert np.array_equal(gc, oc)
            assert np.allclose(gs, os_, rtol=1e-12 if prec == "fp64" else 1e-5, atol=1e-12)
    x = rng.uniform(0, 1, (2, 50))
    x[:, 10] = x[:, 3]
    u = rng.normal(size=(2, 50))
    ie = np.array([-np.inf, 0.05, 0.3, 2.0])
    os_, oc = _oracle.single_pass(x, u, ie)
    for prec in ("fp64", "mixed"):
        gs, gc = gpu_sp(ctx, x, u, ie, prec=prec)
        assert np.array_equal(gc, oc)
        assert np.array_equal(np.isnan(gs), np.isnan(os_)), prec
        ok = ~np.isnan(os_)
        assert np.allclose(gs[ok], os_[ok], rtol=1e-6, atol=1e-12)
    far = np.array([5.0, 6.0])
    gs, gc = gpu_sp(ctx, x, u, far)
    assert gc.sum() == 0 and np.all(gs == 0)


def test_single_pass_batch_and_public_api(ctx):
    """Ragged batch through the C ABI, and the Python mirror of the reference's entry points:
    allocating form (dict keyed by invariant) and the accumulating ! form."""
    rng = np.random.default_rng(21)
    dedges = np.geomspace(2e-3, 1.5, 20)
    fields = [field("flat", 3, n, seed=60 + n)[:2] for n in (300, 1, 517, 64)]
    keep = []
    probs = [problem(x, u, dedges, "flat", "mixed", keep) for x, u in fields]
    nd = len(dedges) - 1
    s = np.zeros(len(probs) * 6 * nd)
    c = np.zeros(len(probs) * 6 * nd, dtype=np.uint64)
    ctx.hist(probs, s, c)
    for k, (x, u) in enumerate(fields):
        os_, oc = _oracle.single_pass(x, u, dedges, nthreads=8)
        blk = slice(k * 6 * nd, (k + 1) * 6 * nd)
        gs = s[blk].reshape((6, nd), order="F")
        gc = c[blk].reshape((6, nd), order="F")
        oa, og = row_scales(x, u, dedges, "flat")
        check_sp(gs, gc, os_, oc, oa, og, "mixed", f"slice {k}")

    x, u = fields[2]
    be = sfamd.MI355XBackend(0, "fp64")
    res = sfamd.calculate_structure_functions_single_pass(
        x, u, dedges, backend=be, output_type=sfamd.StructureFunctionSumsAndCounts, count_eltype=np.uint64)
    assert list(res) == ["S2", "L2", "T2", "S3", "L3", "L1T2"]
    os_, oc = _oracle.single_pass(x, u, dedges, nthreads=8)
    for r, name in enumerate(res):
        assert np.array_equal(res[name].counts, oc[r])
        assert res[name].operator == sfamd.SINGLE_PASS_OPERATORS[name]
    avg = sfamd.calculate_structure_functions_single_pass(x, u, dedges, backend=be)
    with np.errstate(invalid="ignore", divide="ignore"):
        np.testing.assert_allclose(avg["L2"].values, res["L2"].sums / res["L2"].counts, rtol=1e-15)

    sums = np.zeros((6, nd))
    counts = np.zeros((6, nd), dtype=np.uint64)
    sums[2] = 99.0
    for _ in range(2):
        sfamd.calculate_structure_functions_single_pass_inplace(sums, counts, x, u, dedges, backend=be)
    for r in range(6):
        assert np.array_equal(counts[r], 2 * oc[0])
    np.testing.assert_array_equal(sums[2], sums[0] - sums[1])
    np.testing.assert_allclose(sums[0], 2 * res["S2"].sums, rtol=1e-14)

    # shared positions with an auxiliary axis: (6, n_bins, B)
    uu = np.stack([u, 2.0 * u, -u], axis=-1)
    ra = sfamd.calculate_structure_functions_single_pass(
        x, uu, dedges, backend=be, output_type=sfamd.StructureFunctionSumsAndCounts, count_eltype=np.uint64)
    assert ra["S2"].sums.shape == (nd, 3)
    np.testing.assert_allclose(ra["S2"].sums[:, 1], 4.0 * res["S2"].sums, rtol=1e-13)
    # du_L^3 is odd in u; the auxiliary axis runs the batch-leading pair (du_L = du . (dx / r)),
    # the single field the point-field pair (du . dx / r), in another accumulation order, so the
    # two agree to rounding relative to sum |du_L^3| of the bin, not to the (cancelling) sum
    oa, _ = row_scales(x, u, dedges, "flat")
    assert np.all(np.abs(ra["L3"].sums[:, 2] + res["L3"].sums) <= 1e-13 * oa[4] + 1e-300)


def test_single_pass_full_size_mixed_equals_fp64(ctx):
    """N = 65536 sphere: the mixed kernel's counts equal the binary64 kernel's bit for bit and its
    sums agree to the mixed bound (a size-independent property; the oracle would take minutes)."""
    x, u, dedges = field("deg", 2, 65536, seed=99)
    s64, c64 = gpu_sp(ctx, x, u, dedges, "deg", "fp64")
    s32, c32 = gpu_sp(ctx, x, u, dedges, "deg", "mixed")
    assert np.array_equal(c32, c64)
    assert c64[0].sum() > 0
    for r in (0, 1, 3, 4):
        scale = np.abs(s64[r]) + 1e-3 * np.abs(s64[0])
        assert np.all(np.abs(s32[r] - s64[r]) <= 1e-5 * np.maximum(scale, 1e-300)), r
