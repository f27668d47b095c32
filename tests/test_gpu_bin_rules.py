"""LinearBinEdges axes on the GPU: the reference digitizes them with its O(1) rule
(searchsortedfirst(::LinearBinEdges), src/BinEdges.jl:114-131; SquaredLinearPlan's sqrt(r^2) key on
the flat point-field path, :506, 516), not by searching the edge values.  On a lattice the two differ
for many pairs (the rule's decision points sit a few ulps off the edge values), so these cases pin
that the library applies the rule: counts bit-identical to the oracle with the rule, in every
precision, on both geometries, padded or not, and on the batch-leading path."""
import numpy as np
import pytest

import _oracle
import sfamd
from sfamd import bins

pytestmark = pytest.mark.gpu

PRECS = ["fp64", "mixed", "mixed_fixed"]


def lattice():
    g = np.arange(0, 1.0001, 0.1)
    X, Y = np.meshgrid(g, g)
    x = np.stack([X.ravel(), Y.ravel()])
    rng = np.random.default_rng(5)
    u = rng.integers(-8, 9, size=x.shape) * 0.25
    return x, u


def _check(r, s_ref, c_ref, a_ref, label, tol):
    c = np.asarray(r.counts).astype(np.uint64)
    assert np.array_equal(c, c_ref), f"{label}: counts differ in {int((c != c_ref).sum())} cells"
    err = np.abs(np.asarray(r.sums) - s_ref)
    assert np.all(err <= tol * np.maximum(a_ref, 1e-300) + 1e-12), f"{label}: sums"


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("padded", [False, True])
def test_flat_lattice_linear_axes(prec, padded):
    x, u = lattice()
    lin = bins.LinearBinEdges(0.0, 2.0, 21)
    de = bins.InfPaddedBinEdges(lin) if padded else lin
    ve = bins.LinearBinEdges(-4.0, 4.0, 33)
    be = sfamd.MI355XBackend(0, prec)
    for op, sf in (((0, 1, 0), sfamd.L1SF), ((0, 2, 0), sfamd.L2SF)):
        s, c, a = _oracle.hist(x, u, de, op=op, vedges=ve)
        s_srch, c_srch, _ = _oracle.hist(x, u, de.edges(), op=op, vedges=ve.edges())
        assert not np.array_equal(c, c_srch), "the lattice must have pairs where the rule and a search differ"
        r = sfamd.calculate_structure_function(sf, x, u, de, ve, np.uint64, backend=be)
        _check(r, s, c, a, f"joint {op} {prec} padded={padded}", 1e-6 if prec != "fp64" else 1e-12)
        s1, c1, a1 = _oracle.hist(x, u, de, op=op)
        r1 = sfamd.calculate_structure_function(sf, x, u, de, np.uint64, backend=be,
                                                output_type=sfamd.StructureFunctionSumsAndCounts)
        _check(r1, s1, c1, a1, f"1D {op} {prec} padded={padded}", 1e-6 if prec != "fp64" else 1e-12)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_flat_lattice_linear_axes_batch_leading(prec):
    """Auxiliary axis (u (2, N, 2)): the batch-leading loop bins r = sqrt(r^2) with the rule."""
    x, u0 = lattice()
    u = np.stack([u0, -0.5 * u0], axis=-1)
    de = bins.LinearBinEdges(0.0, 2.0, 21)
    be = sfamd.MI355XBackend(0, prec)
    r = sfamd.calculate_structure_function(sfamd.L2SF, x, u, de, np.uint64, backend=be,
                                           output_type=sfamd.StructureFunctionSumsAndCounts)
    for b in range(2):
        s, c, a = _oracle.hist(x, u[:, :, b], de, op=(0, 2, 0), batch_leading=True)
        got = np.asarray(r.counts)[:, b].astype(np.uint64)
        assert np.array_equal(got, c), f"slice {b} {prec}"
        assert np.all(np.abs(np.asarray(r.sums)[:, b] - s) <= 1e-6 * np.maximum(a, 1e-300) + 1e-12)


@pytest.mark.parametrize("prec", PRECS)
def test_sphere_linear_distance_axis(prec):
    """Points at multiples of 0.5 degree along the equator and a meridian: central angles fall on
    the edges of range(0, pi/4, length=46) (0.5 degree steps) up to rounding."""
    k = np.arange(0, 90)
    x = np.concatenate([np.stack([0.5 * k, 0 * k]), np.stack([0 * k + 10.0, 0.5 * k - 20.0])], axis=1).astype(float)
    rng = np.random.default_rng(9)
    u = rng.normal(size=(2, x.shape[1]))
    de = bins.LinearBinEdges(0.0, np.pi / 4, 46)
    ve = bins.LinearBinEdges(-4.0, 4.0, 41)
    be = sfamd.MI355XBackend(0, prec)
    s, c, a = _oracle.hist(x, u, de, op=(0, 1, 0), geometry="deg", vedges=ve)
    r = sfamd.calculate_structure_function(sfamd.L1SF, x, u, de, ve, np.uint64, backend=be,
                                           distance_metric=sfamd.Haversine(1.0))
    _check(r, s, c, a, f"sphere joint {prec}", 1e-6 if prec != "fp64" else 1e-12)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("rule", [True, False])
def test_sphere_values_on_value_edges(prec, rule):
    """Values that sit exactly on a value edge: a third of the points carry zero velocity, so the
    pairs among them have du_L = +-0, on the edge at 0 of range(-2, 2, length=41) (as the rule, or
    as the equally spaced edge vector); quarter-unit velocities along the equator put many more
    values on edges.  The kernels read uniform value bins from a binary32 t = (v - v0) / w and
    must hand every such value to the exact rule or search: counts bit-identical to the oracle."""
    rng = np.random.default_rng(17)
    n = 600
    lon = np.concatenate([rng.uniform(-180, 180, n - 200), np.arange(200) * 1.5 - 150.0])
    lat = np.concatenate([np.degrees(np.arcsin(rng.uniform(-1, 1, n - 200))), np.zeros(200)])
    x = np.stack([lon, lat])
    u = rng.integers(-8, 9, size=(2, n)) * 0.25
    u[:, : n // 3] = 0.0
    u[1, n - 200:] = 0.0
    lin = bins.LinearBinEdges(-2.0, 2.0, 41)
    ve = lin if rule else lin.edges()
    assert 0.0 in np.asarray(lin.edges())
    de = np.array([0.0, 0.01, 0.1, 0.5, 1.0, 2.0, np.pi])
    be = sfamd.MI355XBackend(0, prec)
    s, c, a = _oracle.hist(x, u, de, op=(0, 1, 0), geometry="deg", vedges=ve)
    r = sfamd.calculate_structure_function(sfamd.L1SF, x, u, de, ve, np.uint64, backend=be,
                                           distance_metric=sfamd.Haversine(1.0))
    _check(r, s, c, a, f"sphere values on edges {prec} rule={rule}", 1e-6 if prec != "fp64" else 1e-12)
