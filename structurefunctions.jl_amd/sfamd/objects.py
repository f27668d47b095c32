"""Result containers (src/StructureFunctionObjects.jl:30-115, 246-279, 357-366)."""
from __future__ import annotations

import numpy as np

from .errors import ArgumentError, DimensionMismatch


class AbstractStructureFunction:
    pass


def _bin_average(sums, counts):
    """sums / counts with NaN for empty bins (src/Calculations/serial.jl:132-141)."""
    sums = np.asarray(sums)
    counts = np.asarray(counts)
    out = np.full(sums.shape, np.nan, dtype=sums.dtype)
    nz = counts != 0
    out[nz] = sums[nz] / counts[nz]
    return out


class StructureFunction(AbstractStructureFunction):
    """Averaged 1D result (StructureFunctionObjects.jl:30-42)."""

    def __init__(self, operator, distance, values):
        values = np.asarray(values)
        n = values.shape[0] if values.ndim > 1 else len(values)
        if len(distance) != n + 1:
            raise DimensionMismatch(
                "Flat distance edges must have length one greater than the leading value-bin axis "
                f"(got edges={len(distance)}, value bins={n})")
        self.operator, self.distance, self.values = operator, distance, values

    def __getitem__(self, i):
        return self.values[i]

    def __len__(self):
        return len(self.values)

    def __iter__(self):
        return iter(self.values)


class StructureFunctionSumsAndCounts(AbstractStructureFunction):
    """Raw 1D sums and counts (StructureFunctionObjects.jl:50-74)."""

    def __init__(self, operator, distance, sums, counts):
        sums, counts = np.asarray(sums), np.asarray(counts)
        ns = sums.shape[0] if sums.ndim > 1 else len(sums)
        nc = counts.shape[0] if counts.ndim > 1 else len(counts)
        if not (len(distance) == ns + 1 and ns == nc and sums.shape == counts.shape):
            raise DimensionMismatch(
                "Flat distance edges must satisfy length(distance) == size(sums,1) + 1 and "
                f"sums/counts must match shape (got edges={len(distance)}, sums={sums.shape}, "
                f"counts={counts.shape})")
        self.operator, self.distance, self.sums, self.counts = operator, distance, sums, counts

    def __len__(self):
        return len(self.sums)

    def __add__(self, other):
        if self.operator != other.operator:
            raise ArgumentError("Cannot add results with different operators")
        if not np.array_equal(np.asarray(self.distance), np.asarray(other.distance)):
            raise ArgumentError("Cannot add results with different binning")
        return StructureFunctionSumsAndCounts(self.operator, self.distance, self.sums + other.sums,
                                              self.counts + other.counts)


class StructureFunction2DSumsAndCounts(AbstractStructureFunction):
    """Raw joint distance x value sums and counts (StructureFunctionObjects.jl:88-115)."""

    def __init__(self, operator, distance_bins, value_bins, sums, counts):
        sums, counts = np.asarray(sums), np.asarray(counts)
        if sums.shape != counts.shape:
            raise DimensionMismatch(
                f"Sums and counts matrices must have identical shape (got sums: {sums.shape}, "
                f"counts: {counts.shape})")
        if sums.shape[0] != len(distance_bins) - 1:
            raise DimensionMismatch("distance_bins must have length size(sums,1)+1")
        if sums.shape[1] != len(value_bins) - 1:
            raise DimensionMismatch("value_bins must have length size(sums,2)+1")
        self.operator, self.distance_bins, self.value_bins = operator, distance_bins, value_bins
        self.sums, self.counts = sums, counts

    def __len__(self):
        return len(self.distance_bins) - 1

    def __add__(self, other):
        if self.operator != other.operator:
            raise ArgumentError("Cannot add results with different operators")
        if not np.array_equal(np.asarray(self.distance_bins), np.asarray(other.distance_bins)):
            raise ArgumentError("Cannot add results with different distance binning")
        if not np.array_equal(np.asarray(self.value_bins), np.asarray(other.value_bins)):
            raise ArgumentError("Cannot add results with different value binning")
        return StructureFunction2DSumsAndCounts(self.operator, self.distance_bins, self.value_bins,
                                                self.sums + other.sums, self.counts + other.counts)


def marginalize(sf2d: StructureFunction2DSumsAndCounts) -> StructureFunctionSumsAndCounts:
    """Sum over the value axis (StructureFunctionObjects.jl:357-366)."""
    return StructureFunctionSumsAndCounts(sf2d.operator, sf2d.distance_bins,
                                          sf2d.sums.sum(axis=1), sf2d.counts.sum(axis=1))
