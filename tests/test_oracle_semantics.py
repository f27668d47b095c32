"""The reference evaluates point fields with its point-field kernels and fields with auxiliary
axes with its batch-leading kernels; on the flat geometry they bin and evaluate differently in
the last bits (squared digitize plan and _sf_raw, against pair_frame's r = sqrt(dx.dx) and the
operator on rhat = dx / r: serial.jl:68-101, StructureFunctionTypes.jl:227-244 against
batch_leading.jl:105-330, HelperFunctions.jl:248-330).  These tests pin the oracle's two modes."""
import numpy as np
import pytest

import _oracle
import mp_reference
from test_oracle_mpmath import OPS, case


def test_known_answer_sqrt_rounding():
    """x_i = (0, 0), x_j = (1, 2^-26): dx.dx = 1 + 2^-52 exactly, which is above the squared edge
    1.0 (point kernels: bin (1, 2]) while sqrt(1 + 2^-52) rounds to exactly 1.0 (IEEE round to
    nearest: 1 + 2^-53 - 2^-107 is below the midpoint), which is not above the edge 1.0
    (batch-leading kernels: bin (0.5, 1])."""
    x = np.array([[0.0, 1.0], [0.0, 2.0 ** -26]])
    u = np.array([[0.0, 1.0], [0.0, 0.0]])
    e = np.array([0.5, 1.0, 2.0])
    assert 1.0 + 2.0 ** -52 == 1.0 * 1.0 + 2.0 ** -26 * 2.0 ** -26
    assert np.sqrt(1.0 + 2.0 ** -52) == 1.0
    assert _oracle.pair(x, u, e, 0, 1, op=(0, 1, 0))[3] == 2
    assert _oracle.pair(x, u, e, 0, 1, op=(0, 1, 0), batch_leading=True)[3] == 1
    s, c, _ = _oracle.hist(x, u, e, op=(1, 0, 0))
    sb, cb, _ = _oracle.hist(x, u, e, op=(1, 0, 0), batch_leading=True)
    assert c.tolist() == [0, 1] and cb.tolist() == [1, 0]


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("op", OPS)
def test_batch_leading_vs_mpmath(D, op):
    """Away from edges both modes equal the 40-digit brute force (the modes differ only in
    rounding).  Sums within 1e-13 of sum |value| plus 1e-14 of sum ||du||^order: the transverse
    component from rhat cancels when du is nearly longitudinal, leaving a binary64 error of the
    size of ||du|| (as the reference's own operator does)."""
    x, u, edges = case("flat", D, 14, seed=900 + D + 7 * OPS.index(op))
    ms, mc, ma, margin = mp_reference.brute(x, u, edges, op, "flat")
    assert margin > 1e-9
    s, c, a, g = _oracle.hist(x, u, edges, op=op, batch_leading=True, with_mag=True)
    assert np.array_equal(c, mc)
    assert np.all(np.abs(s - ms) <= 1e-13 * np.maximum(ma, 1e-300) + 1e-14 * g + 1e-300)


def test_lattice_bins_differ_between_modes():
    """On a grid with edges at multiples of the spacing the two modes disagree for about 1 % of
    the counted pairs (distances that are exact multiples in real arithmetic); the sphere has one
    mode."""
    g = np.arange(24) * 0.1
    X, Y = np.meshgrid(g, g)
    x = np.stack([X.ravel(), Y.ravel()])
    u = np.random.default_rng(1).normal(size=x.shape)
    e = np.arange(1, 11) * 0.1
    _, c, _ = _oracle.hist(x, u, e, op=(0, 2, 0), nthreads=8)
    _, cb, _ = _oracle.hist(x, u, e, op=(0, 2, 0), nthreads=8, batch_leading=True)
    assert c.sum() == cb.sum() or abs(int(c.sum()) - int(cb.sum())) < c.sum() // 50
    assert not np.array_equal(c, cb)
    s1, c1 = _oracle.single_pass(x, u, e, nthreads=8)
    s2, c2 = _oracle.single_pass(x, u, e, nthreads=8, batch_leading=True)
    assert not np.array_equal(c1, c2)
    xs = np.stack([np.linspace(-170, 170, 60), np.linspace(-60, 60, 60)])
    us = np.random.default_rng(2).normal(size=(2, 60))
    es = np.geomspace(0.05, 3.0, 9)
    for op in [(0, 1, 0), (0, 2, 0)]:
        a = _oracle.hist(xs, us, es, op=op, geometry="deg")
        b = _oracle.hist(xs, us, es, op=op, geometry="deg", batch_leading=True)
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[0], b[0])
