"""Bin edge specifications (src/BinEdges.jl) normalised to explicit binary64 edge vectors.

Bins are half-open ``(e[k-1], e[k]]`` (src/HelperFunctions.jl:77-84).  The C ABI takes explicit
binary64 edge vectors; each reference edge type maps to the edge values it represents:

* plain vector           -> itself (BinEdges, generic search, BinEdges.jl:49-61)
* LinearBinEdges(range)  -> the range's elements (BinEdges.jl:94-131)
* LogBinEdges(edges)     -> exp.(range(log(first), log(last); length)) (physical_edges_vector,
                            BinEdges.jl:181-205)
* InfPaddedBinEdges(e)   -> [-Inf; e; +Inf] with the implicit catch-all bins (BinEdges.jl:270-345).
  In the flat geometry the leading -Inf stands for SquaredInfPaddedPlan's underflow bin
  (BinEdges.jl:434-437, 520, 553-558): it is below every squared distance, including 0.

The reference digitizes LinearBinEdges/LogBinEdges with O(1) FMA/log lookups whose results can
differ from a search on these edge values for queries within an ulp of an edge; the explicit
vectors make the binning exact and identical on host and device.
"""
from __future__ import annotations

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
    """Generic sorted edges (src/BinEdges.jl:49-61)."""

    def __init__(self, edges):
        if isinstance(edges, AbstractBinEdges):
            self._e = edges.edges()
        else:
            self._e = np.asarray(edges, dtype=np.float64).ravel().copy()

    def edges(self):
        return self._e


class LinearBinEdges(AbstractBinEdges):
    """Uniform edges from (start, stop, length) (src/BinEdges.jl:94-131)."""

    def __init__(self, start, stop=None, length=None):
        if stop is None:  # from an explicit uniform vector (BinEdges.jl:102)
            v = np.asarray(start, dtype=np.float64).ravel()
            start, stop, length = v[0], v[-1], len(v)
        self.start, self.stop, self.length = float(start), float(stop), int(length)

    def edges(self):
        return np.linspace(self.start, self.stop, self.length)


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
    """Implicit -Inf / +Inf catch-all edges (src/BinEdges.jl:270-345)."""

    def __init__(self, edges):
        e = np.asarray(edges, dtype=np.float64).ravel()
        if len(e) and np.isinf(e[0]):
            e = e[1:]
        if len(e) and np.isinf(e[-1]):
            e = e[:-1]
        self.inner = e

    def edges(self):
        return np.concatenate([[-np.inf], self.inner, [np.inf]])


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
