"""The oracle's LinearBinEdges rule (searchsortedfirst(::LinearBinEdges), src/BinEdges.jl:114-131)
pinned by the reference's own known answers (test/test_bin_edges.jl:29-52, 112-123, 150-184) and
by an exact-arithmetic restatement of the rule (rationals; a fused multiply-add is one rounding of
the exact a*b + c), over queries one and two ulps either side of every edge, where the rule's
decision points depart from the edge values.  The restatement distinguishes a fused from an
unfused muladd on these queries, so the oracle's use of fused multiply-adds is checked too."""
import bisect
import math
from fractions import Fraction

import numpy as np
import pytest

import _oracle
from sfamd import bins


def _fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def _ssf_exact(r, x, fused=True):
    """searchsortedfirst(::LinearBinEdges, x) with exact rational arithmetic."""
    f, l, n = r["first_edge"], r["last_edge"], r["n"]
    if x <= f:
        return 1
    if x > l:
        return n + 1
    c = (-f) * r["inv_step"]
    t = _fma(x, r["inv_step"], c) if fused else x * r["inv_step"] + c
    idx = min(max(math.floor(t) + 1, 1), n)
    u = _fma(idx - 1, r["step_val"], f) if fused else (idx - 1) * r["step_val"] + f
    return idx + 1 if u < x else idx


def _digitize_exact(r, x, fused=True):
    if r["padded"]:
        if x <= -math.inf:
            return 0
        if x > r["last_edge"]:
            return r["n"] + 1
        return _ssf_exact(r, x, fused)
    return _ssf_exact(r, x, fused) - 1


def test_reference_kat_linear_matches_range_search():
    """test_bin_edges.jl:29-45: for range(1, 10, length=10) and 100 queries range(0, 11, length=100),
    the rule equals searchsortedfirst on the range itself."""
    lin = bins.LinearBinEdges(1.0, 10.0, 10)
    r = lin.rule()
    e = list(lin.edges())
    assert e == [float(k) for k in range(1, 11)]
    for q in bins.LinearBinEdges(0.0, 11.0, 100).edges():
        assert _oracle.rule_digitize(r, q) == bisect.bisect_left(e, q), q


def test_reference_kat_padded_linear():
    """test_bin_edges.jl:112-123: InfPaddedBinEdges(LinearBinEdges(1:10)) over range(-1, 12, length=100):
    searchsortedfirst = inner + 1 (1 at -Inf, length + 2 above the last edge)."""
    lin = bins.LinearBinEdges(1.0, 10.0, 10)
    pad = bins.InfPaddedBinEdges(lin)
    r = bins.rule_of(pad)
    assert r["padded"] and len(pad.edges()) == 12
    e = list(lin.edges())
    for q in list(bins.LinearBinEdges(-1.0, 12.0, 100).edges()) + [-math.inf, math.inf]:
        if q <= -math.inf:
            ref = 1
        elif q > e[-1]:
            ref = len(e) + 2
        else:
            ref = bisect.bisect_left(e, q) + 1 + 1
        assert _oracle.rule_digitize(r, q) == ref - 1, q


def test_reference_kat_hot_loop_random_points():
    """test_bin_edges.jl:150-184: with 40 random points, LinearBinEdges(range(0.01, 2, length=11))
    and the collected vector give the same counts (no pair lies within ulps of an edge)."""
    rng = np.random.default_rng(42)
    x = rng.uniform(size=(2, 40))
    u = rng.normal(size=(2, 40))
    lin = bins.LinearBinEdges(0.01, 2.0, 11)
    s1, c1, _ = _oracle.hist(x, u, lin, op=(0, 2, 0))
    s2, c2, _ = _oracle.hist(x, u, lin.edges(), op=(0, 2, 0))
    assert np.array_equal(c1, c2) and int(c1.sum()) > 0
    np.testing.assert_allclose(s1, s2, rtol=1e-14)


def _edge_queries(e):
    q = []
    for v in e:
        a = float(v)
        q += [a, float(np.nextafter(a, -np.inf)), float(np.nextafter(a, np.inf)),
              float(np.nextafter(np.nextafter(a, -np.inf), -np.inf)),
              float(np.nextafter(np.nextafter(a, np.inf), np.inf))]
    return q


def test_rule_matches_exact_restatement_near_edges():
    rng = np.random.default_rng(7)
    n_fused_differs = n_search_differs = 0
    for trial in range(120):
        a = float(rng.uniform(-5, 5))
        b = a + float(10.0 ** rng.uniform(-2, 1))
        n = int(rng.integers(2, 120))
        lin = bins.LinearBinEdges(a, b, n)
        e = lin.edges()
        queries = _edge_queries(e) + list(rng.uniform(a - 1, b + 1, 50)) + [math.nan]
        for padded in (False, True):
            r = lin.rule(padded=padded)
            nb = n + 1 if padded else n - 1
            for q in queries:
                got = _oracle.rule_digitize(r, q)
                if q != q:
                    assert got == nb + 1  # NaN (rejected by the reference) is not counted
                    continue
                want = _digitize_exact(r, q)
                assert got == want, (a, b, n, padded, q)
                n_fused_differs += want != _digitize_exact(r, q, fused=False)
                if not padded:
                    n_search_differs += want != bisect.bisect_left(list(e), q)
    # the queries discriminate: an unfused muladd, or a search of the edge values, would fail
    assert n_fused_differs > 0 and n_search_differs > 0


def test_rule_is_monotone():
    lin = bins.LinearBinEdges(-0.766735510274243, 7.51201340199197, 52)
    r = lin.rule()
    q = np.sort(np.concatenate([_edge_queries(lin.edges()), np.linspace(-1, 8, 2001)]))
    d = [_oracle.rule_digitize(r, float(v)) for v in q]
    assert all(d1 <= d2 for d1, d2 in zip(d, d[1:]))


@pytest.mark.parametrize("geometry", ["flat", "deg"])
def test_histogram_uses_rule(geometry):
    """A histogram over a LinearBinEdges axis is the per-pair rule applied to each pair's distance
    (brute force over sfo_pair's distances)."""
    if geometry == "flat":
        g = np.arange(0, 1.0001, 0.1)
        X, Y = np.meshgrid(g, g)
        x = np.stack([X.ravel(), Y.ravel()])
        lin = bins.LinearBinEdges(0.0, 2.0, 21)
    else:
        k = np.arange(0, 60)
        x = np.concatenate([np.stack([0.5 * k, 0 * k]), np.stack([0 * k + 10.0, 0.5 * k - 20.0])], axis=1).astype(float)
        lin = bins.LinearBinEdges(0.0, np.pi / 4, 46)
    u = np.random.default_rng(1).normal(size=x.shape)
    s, c, _ = _oracle.hist(x, u, lin, op=(0, 2, 0), geometry=geometry)
    r = lin.rule()
    N = x.shape[1]
    want = np.zeros(len(c), dtype=np.uint64)
    for i in range(N):
        for j in range(i + 1, N):
            ok, dist, _, _, _ = _oracle.pair(x, u, lin.edges(), i, j, op=(0, 2, 0), geometry=geometry)
            b = _digitize_exact(r, dist)
            if ok and 1 <= b <= len(c):
                want[b - 1] += 1
    assert np.array_equal(c, want)
