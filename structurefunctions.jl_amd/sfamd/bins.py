"""Bin edge specifications (src/BinEdges.jl) and how each reaches the C ABI.

Bins are half-open ``(e[k-1], e[k]]`` (src/HelperFunctions.jl:77-84).  The C ABI always takes
the axis as an explicit binary64 edge vector; a LinearBinEdges axis also passes its rule:

* plain vector           -> itself, searched (BinEdges, generic search, BinEdges.jl:49-61)
* LinearBinEdges(range)  -> the range's elements plus the rule fields (first_edge, last_edge,
                            inv_step, step_val, length): the reference digitizes it with
                            floor(muladd(x, inv_step, -first * inv_step)) + 1 and a one-sided
                            correction (BinEdges.jl:114-131), on the flat squared-distance plan
                            with the key sqrt(r^2) (SquaredLinearPlan, :412-417, 506, 516)
* LogBinEdges(edges)     -> exp.(range(log(first), log(last); length)) (physical_edges_vector,
                            BinEdges.jl:181-205), searched.  On the flat point-field path this is
                            exactly the reference (its SquaredLogPlan searches the squared
                            physical edges, :425-437, 528-541); elsewhere the reference digitizes
                            log(x) on the log grid (:214-217), which a search of the physical edges
                            matches except for queries within about an ulp of an edge (parity
                            unpinned: Julia's log is not available here)
* InfPaddedBinEdges(e)   -> [-Inf; e; +Inf] with the implicit catch-all bins (BinEdges.jl:270-345),
                            keeping e's rule when e is a LinearBinEdges.  In the flat geometry the
                            leading -Inf stands for SquaredInfPaddedPlan's underflow bin
                            (BinEdges.jl:434-437, 520, 553-558): it is below every squared
                            distance, including 0.

A Python LinearBinEdges(start, stop, length) builds what Julia's range(start, stop; length) holds:
elements and step rounded to nearest from exact rational arithmetic (Julia computes them in
TwicePrecision, which agrees except in rare ties), inv_step = 1 / step in binary64.
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np

from .errors import ArgumentError


class AbstractBinEdges:
    """src/BinEdges.jl:32."""

    def edges(self) -> np.ndarray:
        raise NotImplementedError

    def __len__(self):
        return len(self.edges())

    def __array__(self, dtype=None, copy=None):
        e = self.edges()
        return e.astype(dtype) if dtype is not None else e


class BinEdges(AbstractBinEdges):
    """Generic sorted edges (src/BinEdges.jl:49-61); BinEdges of an AbstractBinEdges is that object
    (BinEdges.jl:362)."""

    def __new__(cls, edges):
        if isinstance(edges, AbstractBinEdges):
            return edges
        return super().__new__(cls)

    def __init__(self, edges):
        if isinstance(edges, AbstractBinEdges):
            return
        self._e = np.asarray(edges, dtype=np.float64).ravel().copy()

    def edges(self):
        return self._e


def _range_elements(start: float, stop: float, length: int) -> np.ndarray:
    """Elements of range(start, stop; length) rounded to nearest (exact rationals)."""
    a, b = Fraction(start), Fraction(stop)
    if length == 1:
        return np.array([start])
    out = np.empty(length)
    for k in range(length):
        out[k] = float(a + (b - a) * k / (length - 1))
    out[0], out[-1] = start, stop
    return out


class LinearBinEdges(AbstractBinEdges):
    """Uniform edges from (start, stop, length) (src/BinEdges.jl:94-131).  Holds the reference's
    fields: first_edge, last_edge, step_val = step(range), inv_step = inv(step_val)."""

    def __init__(self, start, stop=None, length=None):
        if stop is None:  # from an explicit uniform vector (BinEdges.jl:102)
            v = np.asarray(start, dtype=np.float64).ravel()
            start, stop, length = v[0], v[-1], len(v)
        self.start, self.stop, self.length = float(start), float(stop), int(length)
        if self.length < 2:
            raise ArgumentError("LinearBinEdges needs at least 2 edges")
        self._e = _range_elements(self.start, self.stop, self.length)
        self.first_edge = self.start
        self.last_edge = self.stop
        self.step_val = float((Fraction(self.stop) - Fraction(self.start)) / (self.length - 1))
        self.inv_step = 1.0 / self.step_val if self.step_val != 0.0 else float("inf")

    def edges(self):
        return self._e

    def rule(self, padded=False):
        """The LinearBinEdges rule fields as the C ABI takes them (sf_bin_rule)."""
        return {"padded": bool(padded), "n": self.length, "first_edge": self.first_edge,
                "last_edge": self.last_edge, "inv_step": self.inv_step, "step_val": self.step_val}


class LogBinEdges(AbstractBinEdges):
    """Log-uniform edges (src/BinEdges.jl:171-205)."""

    def __init__(self, edges=None, *, log_edges=None):
        if log_edges is not None:
            le = np.asarray(log_edges, dtype=np.float64).ravel()
        else:
            e = np.asarray(edges, dtype=np.float64).ravel()
            if np.any(e <= 0):
                raise ArgumentError("Log-spaced bin edges must be strictly positive.")
            le = np.log(e)
        self.log_edges = np.linspace(le[0], le[-1], len(le))

    def edges(self):
        # physical_edges_vector (BinEdges.jl:203-205)
        return np.exp(self.log_edges)


def LogBinEdges_from_log_edges(log_edges):
    """src/BinEdges.jl:192-195."""
    return LogBinEdges(log_edges=log_edges)


class InfPaddedBinEdges(AbstractBinEdges):
    """Implicit -Inf / +Inf catch-all edges (src/BinEdges.jl:270-345).  A LinearBinEdges inside
    keeps its rule."""

    def __init__(self, edges):
        self.linear = edges if isinstance(edges, LinearBinEdges) else None
        e = edges.edges() if isinstance(edges, AbstractBinEdges) else np.asarray(edges, dtype=np.float64).ravel()
        if self.linear is None:
            if len(e) and np.isinf(e[0]):
                e = e[1:]
            if len(e) and np.isinf(e[-1]):
                e = e[:-1]
        self.inner = e

    def edges(self):
        return np.concatenate([[-np.inf], self.inner, [np.inf]])


def rule_of(bins):
    """The LinearBinEdges rule of a bin specification (dict of sf_bin_rule fields), or None when
    the axis is searched."""
    if isinstance(bins, LinearBinEdges):
        return bins.rule()
    if isinstance(bins, InfPaddedBinEdges) and bins.linear is not None:
        return bins.linear.rule(padded=True)
    if isinstance(bins, range) and len(bins) >= 2:  # an AbstractRange: BinEdges wraps it (BinEdges.jl:363)
        return LinearBinEdges(bins[0], bins[-1], len(bins)).rule()
    return None


def as_edges(bins) -> np.ndarray:
    """Explicit binary64 edge vector for any supported bin specification."""
    if isinstance(bins, AbstractBinEdges):
        e = bins.edges()
    elif isinstance(bins, range):
        e = np.asarray(list(bins), dtype=np.float64)
    else:
        e = np.asarray(bins, dtype=np.float64).ravel()
    e = np.ascontiguousarray(e, dtype=np.float64)
    if e.ndim != 1 or len(e) < 2:
        raise ArgumentError("bin edges must be a vector with at least 2 entries")
    return e


def n_histogram_bins(bins) -> int:
    """length(edges) - 1 (src/BinEdges.jl:352)."""
    return len(as_edges(bins)) - 1


def midpoints(bins) -> np.ndarray:
    """src/HelperFunctions.jl:91-94."""
    e = as_edges(bins)
    return (e[:-1] + e[1:]) / 2
