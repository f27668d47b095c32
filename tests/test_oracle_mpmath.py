"""Brute force at small N in 40-digit arithmetic from independent formulas (mp_reference.py)
pins the oracle's operators, geometry and binning: counts identical, sums within 1e-13 of
sum |value|."""
import ctypes

import numpy as np
import pytest

import _oracle
import mp_reference

OPS = [(0, 1, 0), (0, 2, 0), (0, 3, 0), (0, 0, 1), (0, 0, 2), (0, 0, 3), (0, 2, 1), (0, 1, 2),
       (1, 0, 0), (2, 0, 0), (3, 0, 0), (4, 0, 0), (5, 1, 0), (5, 3, 0)]


def case(geometry, D, N, seed):
    rng = np.random.default_rng(seed)
    if geometry == "flat":
        x = rng.uniform(0, 1, (D, N))
        edges = np.geomspace(0.02, 1.8, 7)
    else:
        lon = rng.uniform(-180, 180, N)
        lat = np.degrees(np.arcsin(rng.uniform(-1, 1, N)))
        x = np.stack([lon, lat])
        if geometry == "rad":
            x = np.radians(x)
        edges = np.geomspace(0.05, 3.1, 7)
    u = rng.normal(size=(D, N))
    return x, u, edges


@pytest.mark.parametrize("geometry,D", [("flat", 2), ("flat", 3), ("deg", 2), ("rad", 2), ("deg", 3)])
@pytest.mark.parametrize("op", OPS)
def test_oracle_vs_mpmath(geometry, D, op):
    x, u, edges = case(geometry, D, 14, seed=100 + D + 7 * OPS.index(op))
    ms, mc, ma, margin = mp_reference.brute(x, u, edges, op, geometry)
    assert margin > 1e-9, "test data too close to an edge"
    s, c, a = _oracle.hist(x, u, edges, op=op, geometry=geometry)
    assert np.array_equal(c, mc)
    assert np.all(np.abs(s - ms) <= 1e-13 * np.maximum(ma, 1e-300) + 1e-300)


@pytest.mark.parametrize("geometry,D", [("flat", 2), ("flat", 3), ("deg", 2)])
def test_oracle_joint_vs_mpmath(geometry, D):
    x, u, edges = case(geometry, D, 16, seed=77 + D)
    vedges = np.array([-2.5, -1.1, -0.3, 0.35, 1.2, 2.7])
    ms, mc, ma, margin = mp_reference.brute(x, u, edges, (0, 1, 0), geometry, vedges=vedges)
    assert margin > 1e-9
    s, c, a = _oracle.hist(x, u, edges, op=(0, 1, 0), geometry=geometry, vedges=vedges)
    assert np.array_equal(c, mc)
    assert np.all(np.abs(s - ms) <= 1e-13 * np.maximum(ma, 1e-300) + 1e-300)


@pytest.mark.parametrize("geometry,D", [("flat", 2), ("flat", 3), ("deg", 2), ("deg", 3)])
def test_single_pass_vs_mpmath_and_single_operators(geometry, D):
    """Single-pass invariants (serial_single_pass.jl:602-700): each row equals the brute force of
    its operator, the count row is shared, and S2 = L2 + T2, S3 = L3 + L1T2 per bin."""
    x, u, edges = case(geometry, D, 14, seed=300 + D)
    s, c = _oracle.single_pass(x, u, edges, geometry=geometry)
    rows = [(1, 0, 0), (0, 2, 0), (0, 0, 2), (2, 0, 0), (0, 3, 0), (0, 1, 2)]
    for r, op in enumerate(rows):
        ms, mc, ma, margin = mp_reference.brute(x, u, edges, op, geometry)
        assert margin > 1e-9
        assert np.array_equal(c[r], mc)
        assert np.all(np.abs(s[r] - ms) <= 1e-12 * np.maximum(ma, 1e-300) + 1e-300), r
    np.testing.assert_array_equal(s[2], s[0] - s[1])
    np.testing.assert_array_equal(s[5], s[3] - s[4])


def test_single_pass_accumulates_then_derives():
    """calculate_structure_functions_single_pass! semantics (serial_single_pass.jl:404-421,
    208-219): a second call adds rows S2, L2, S3, L3 and count row 1 and re-derives the rest, so
    a buffer whose derived rows were tampered with is repaired."""
    x, u, edges = case("flat", 2, 12, seed=77)
    s1, c1 = _oracle.single_pass(x, u, edges)
    p, keep = _oracle.make(x, u, edges, (1, 0, 0), "flat", 1.0, None)
    nd = p.n_dist
    s = np.zeros(6 * nd)
    c = np.zeros(6 * nd, dtype=np.uint64)
    s[2::6] = 123.0
    c[3::6] = 7
    for _ in range(2):
        assert _oracle.lib().sfo_hist_single_pass(ctypes.byref(p), 1, 0, p.N, 1, s.ctypes.data, c.ctypes.data) == 0
    s = s.reshape((6, nd), order="F")
    c = c.reshape((6, nd), order="F")
    for r in (0, 1, 3, 4):
        np.testing.assert_allclose(s[r], 2 * s1[r], rtol=1e-15)
    np.testing.assert_array_equal(s[2], s[0] - s[1])
    np.testing.assert_array_equal(s[5], s[3] - s[4])
    for r in range(6):
        np.testing.assert_array_equal(c[r], 2 * c1[0])


SP_ROWS = [(1, 0, 0), (0, 2, 0), (0, 0, 2), (2, 0, 0), (0, 3, 0), (0, 1, 2)]


def sp2d_value_edges(geometry):
    """Per-invariant value edges: log-spaced for the non-negative invariants, linear for the
    signed ones (the pairing the reference's docs suggest, batch_api.jl:108-112)."""
    pos = np.concatenate([[0.0], np.geomspace(0.05, 12.0, 6)])
    sgn = np.array([-9.0, -2.2, -0.4, 0.3, 1.9, 8.5, 40.0])
    return [pos, pos, pos, sgn, sgn, sgn]


@pytest.mark.parametrize("geometry,D", [("flat", 2), ("flat", 3), ("deg", 2), ("deg", 3)])
def test_single_pass_2d_vs_mpmath(geometry, D):
    """Single pass, 2D form (serial_single_pass.jl:1161-1181): invariant t's (distance, value)
    histogram equals the 40-digit brute force of its operator binned by its own value edges."""
    x, u, edges = case(geometry, D, 15, seed=500 + D)
    ve = sp2d_value_edges(geometry)
    s, c = _oracle.single_pass_2d(x, u, edges, ve, geometry=geometry)
    for t, op in enumerate(SP_ROWS):
        ms, mc, ma, margin = mp_reference.brute(x, u, edges, op, geometry, vedges=ve[t])
        assert margin > 1e-9, (t, margin)
        assert np.array_equal(c[t], mc), t
        assert np.all(np.abs(s[t] - ms) <= 1e-12 * np.maximum(ma, 1e-300) + 1e-300), t


def test_single_pass_2d_marginals_match_1d():
    """With value edges wide enough to hold every value, summing the 2D form over value bins
    gives the 1D single pass: counts exactly, sums to rounding (T2 and L1T2 are accumulated per
    pair here and derived from the other rows in 1D)."""
    x, u, edges = case("flat", 2, 40, seed=9)
    wide = np.array([-1e300, -1.0, 0.0, 0.5, 2.0, 1e300])
    s2, c2 = _oracle.single_pass_2d(x, u, edges, wide)
    s1, c1 = _oracle.single_pass(x, u, edges)
    for t in range(6):
        np.testing.assert_array_equal(c2[t].sum(axis=1), c1[t])
    scale = np.abs(s1).max()
    np.testing.assert_allclose(s2.sum(axis=2), s1, rtol=0, atol=1e-12 * scale)
    # one shared vector == six copies of it
    s6, c6 = _oracle.single_pass_2d(x, u, edges, [wide] * 6)
    np.testing.assert_array_equal(c6, c2)
    np.testing.assert_array_equal(s6, s2)


def test_single_pass_2d_threads_and_rows():
    """Round-robin threads give the same counts and sums to rounding; row subsets add up."""
    x, u, edges = case("deg", 2, 60, seed=11)
    ve = sp2d_value_edges("deg")
    s1, c1 = _oracle.single_pass_2d(x, u, edges, ve, geometry="deg")
    s4, c4 = _oracle.single_pass_2d(x, u, edges, ve, geometry="deg", nthreads=4)
    np.testing.assert_array_equal(c1, c4)
    np.testing.assert_allclose(s1, s4, rtol=1e-13, atol=1e-300)
    sa, ca = _oracle.single_pass_2d(x, u, edges, ve, geometry="deg", rows=(0, 60, 2))
    sb, cb = _oracle.single_pass_2d(x, u, edges, ve, geometry="deg", rows=(1, 60, 2))
    np.testing.assert_array_equal(ca + cb, c1)
