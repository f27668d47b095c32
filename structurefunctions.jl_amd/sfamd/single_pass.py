"""The six isotropic invariants S2, L2, T2, S3, L3, L1T2 in one pair pass.

    calculate_structure_functions_single_pass(x, u, distance_bins; backend, output_type,
                                              count_eltype, distance_metric)
        src/Calculations/serial_single_pass.jl:663-700; returns a dict keyed by invariant
    calculate_structure_functions_single_pass_inplace(sums, counts, x, u, distance_bins; ...)
        calculate_structure_functions_single_pass! (serial_single_pass.jl:404-421): accumulates
        into (6, n_bins[, auxiliary...]) buffers

    calculate_structure_functions_single_pass_2d(x, u, distance_bins, value_bins; ...)
        src/Calculations/serial_single_pass.jl:1161-1181; dict keyed by invariant of
        StructureFunction2DSumsAndCounts (each invariant binned by distance and its own value)
    calculate_structure_functions_single_pass_2d_inplace(sums_3d, counts_3d, x, u, distance_bins,
                                                         value_bins; ...)
        calculate_structure_functions_single_pass_2d! (serial_single_pass.jl:712-728): accumulates
        into (6, n_bins, n_val[, auxiliary...]) buffers

One pass bins every pair by distance and accumulates |du|^2, du_L^2, du_L |du|^2 and du_L^3
(rows 1, 2, 4, 5); T2 = S2 - L2 and L1T2 = S3 - L3 are differenced per bin and the shared count is
replicated into every row (_sp1d_derive_rows!, serial_single_pass.jl:208-219).  The pair loop is
the library's SF_OP_SINGLE_PASS kernel.  The Helmholtz rows the reference appends for point-field
input (helmholtz_decompose_2d, serial_single_pass.jl:11-92) are post-processing and not produced.
"""
from __future__ import annotations

import numpy as np

from . import _native
from .bins import as_edges, rule_of
from .dispatch import (MI355XBackend, _assert_counts_representable, _finalize, _run, _semantics, _slices,
                       _validate_array_shape)
from .errors import DimensionMismatch
from .metrics import Euclidean
from .objects import StructureFunction, StructureFunction2DSumsAndCounts, StructureFunctionSumsAndCounts
from .operators import L1T2SF, L2SF, L3SF, S2SF, S3SF, T2SF

SINGLE_PASS_N = 6
# stacked-row order of the accumulator (serial_single_pass.jl:602-609)
SINGLE_PASS_OPERATORS = {"S2": S2SF, "L2": L2SF, "T2": T2SF, "S3": S3SF, "L3": L3SF, "L1T2": L1T2SF}


class _SinglePassOp:
    """Operator code of the library's single-pass kernel (not a user-facing operator)."""
    code = (_native.OP_SINGLE_PASS, 0, 0)


def _single_pass_raw(x, u, distance_bins, backend, distance_metric):
    """(sums, counts) shaped (6, n_bins, auxiliary...) for fresh (zeroed) buffers."""
    backend = backend or MI355XBackend()
    kind, _, _ = _validate_array_shape(x, u, distance_metric)
    de = as_edges(distance_bins)
    slices, aux = _slices(x, u, kind)
    sums, counts, nd, _ = _run(_SinglePassOp(), slices, [de] * len(slices), [None] * len(slices),
                               distance_metric, backend, _semantics(kind), [rule_of(distance_bins)] * len(slices))
    B = len(slices)
    s = sums.reshape(B, SINGLE_PASS_N * nd).T.reshape((SINGLE_PASS_N, nd) + aux, order="F")
    c = counts.reshape(B, SINGLE_PASS_N * nd).T.reshape((SINGLE_PASS_N, nd) + aux, order="F")
    return s, c


def calculate_structure_functions_single_pass(x, u, distance_bins, *, backend: MI355XBackend | None = None,
                                              output_type=StructureFunction, count_eltype=np.uint32,
                                              distance_metric=Euclidean()):
    """Six invariants in one pass, as {name: result} (serial_single_pass.jl:663-700)."""
    x = np.asarray(x, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)
    _validate_array_shape(x, u, distance_metric)
    _assert_counts_representable(count_eltype, x.shape[1])
    s, c = _single_pass_raw(x, u, distance_bins, backend, distance_metric)
    c = c.astype(count_eltype)
    out = {}
    for row, (name, op) in enumerate(SINGLE_PASS_OPERATORS.items()):
        raw = StructureFunctionSumsAndCounts(op, distance_bins, s[row], c[0])
        out[name] = _finalize(raw, output_type)
    return out


def calculate_structure_functions_single_pass_inplace(sums, counts, x, u, distance_bins, *,
                                                      backend: MI355XBackend | None = None,
                                                      distance_metric=Euclidean()):
    """calculate_structure_functions_single_pass! (serial_single_pass.jl:404-421): rows 1, 2, 4, 5
    and count row 1 accumulate into sums/counts; rows 3, 6 and count rows 2..6 are then rederived
    over the whole buffer."""
    x = np.asarray(x, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)
    kind, _, _ = _validate_array_shape(x, u, distance_metric)
    nd = len(as_edges(distance_bins)) - 1
    aux = () if kind == "point" else u.shape[2:]
    want = (SINGLE_PASS_N, nd) + aux
    if tuple(np.shape(sums)) != want:
        raise DimensionMismatch(f"sums must have shape {want}; got {np.shape(sums)}")
    if tuple(np.shape(counts)) != want:
        raise DimensionMismatch(f"counts must have shape {want}; got {np.shape(counts)}")
    s, c = _single_pass_raw(x, u, distance_bins, backend, distance_metric)
    for r in (0, 1, 3, 4):
        sums[r] += s[r]
    counts[0] += c[0].astype(counts.dtype)
    sums[2] = sums[0] - sums[1]
    sums[5] = sums[3] - sums[4]
    for r in range(1, SINGLE_PASS_N):
        counts[r] = counts[0]
    return sums, counts


# ---- 2D (joint) form ------------------------------------------------------------------------

def _value_bin_list(value_bins):
    """The six per-invariant edge vectors: one vector shared by all invariants, or a 6-tuple
    (SinglePass2DValueBins, batch_api.jl:101-104)."""
    if isinstance(value_bins, (tuple, list)) and len(value_bins) == SINGLE_PASS_N and \
            all(np.ndim(v) == 1 for v in value_bins):
        return list(value_bins)
    return [value_bins] * SINGLE_PASS_N


def _validate_value_bins(vlist, n_val):
    """_validate_value_bins! (batch_api.jl:132-150): every vector needs n_val + 1 edges or more."""
    for t, v in enumerate(vlist):
        if len(v) < n_val + 1:
            raise DimensionMismatch(
                f"value_bins[{t + 1}] needs at least {n_val + 1} edges for n_val={n_val} (got {len(v)})")


def _single_pass_2d_raw(x, u, distance_bins, vlist, n_val, backend, distance_metric):
    """(sums, counts) shaped (6, n_bins, n_val, auxiliary...) for fresh buffers.  Only the first
    n_val + 1 edges of each vector matter: a value above edge n_val + 1 falls in a bin > n_val,
    which the reference drops (_sp2d_scatter!, serial_single_pass.jl:852-864)."""
    backend = backend or MI355XBackend()
    kind, _, _ = _validate_array_shape(x, u, distance_metric)
    de = as_edges(distance_bins)
    ve6 = np.ascontiguousarray(np.concatenate(
        [np.asarray(as_edges(v), dtype=np.float64)[:n_val + 1] for v in vlist]))
    slices, aux = _slices(x, u, kind)
    sums, counts, nd, nv = _run(_SinglePassOp(), slices, [de] * len(slices), [ve6] * len(slices),
                                distance_metric, backend, _semantics(kind), [rule_of(distance_bins)] * len(slices))
    B = len(slices)
    ncell = SINGLE_PASS_N * nd * nv
    s = sums.reshape(B, ncell).T.reshape((SINGLE_PASS_N, nd, nv) + aux, order="F")
    c = counts.reshape(B, ncell).T.reshape((SINGLE_PASS_N, nd, nv) + aux, order="F")
    return s, c


def calculate_structure_functions_single_pass_2d(x, u, distance_bins, value_bins, *,
                                                 backend: MI355XBackend | None = None,
                                                 output_type=StructureFunction2DSumsAndCounts,
                                                 count_eltype=np.uint32, distance_metric=Euclidean()):
    """Six invariant joint (distance, value) histograms in one pass, as {name: result}
    (serial_single_pass.jl:1161-1181, collection :1139-1150).  Unlike the 1D form the counts
    differ per invariant: each invariant's value lands in its own value bin."""
    x = np.asarray(x, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)
    _validate_array_shape(x, u, distance_metric)
    _assert_counts_representable(count_eltype, x.shape[1])
    vlist = _value_bin_list(value_bins)
    n_val = len(vlist[0]) - 1
    _validate_value_bins(vlist, n_val)
    s, c = _single_pass_2d_raw(x, u, distance_bins, vlist, n_val, backend, distance_metric)
    c = c.astype(count_eltype)
    out = {}
    for t, (name, op) in enumerate(SINGLE_PASS_OPERATORS.items()):
        raw = StructureFunction2DSumsAndCounts(op, distance_bins, vlist[t], s[t], c[t])
        out[name] = _finalize(raw, output_type)
    return out


def calculate_structure_functions_single_pass_2d_inplace(sums_3d, counts_3d, x, u, distance_bins,
                                                         value_bins, *,
                                                         backend: MI355XBackend | None = None,
                                                         distance_metric=Euclidean()):
    """calculate_structure_functions_single_pass_2d! (serial_single_pass.jl:712-728, checks of
    _accumulate_single_pass_2d! :740-747): accumulates into (6, n_bins, n_val[, aux...]) buffers,
    n_val = size(sums_3d, 3)."""
    x = np.asarray(x, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)
    kind, _, _ = _validate_array_shape(x, u, distance_metric)
    nd = len(as_edges(distance_bins)) - 1
    shp = tuple(np.shape(sums_3d))
    if len(shp) < 3 or shp[0] != SINGLE_PASS_N or shp[1] != nd:
        raise DimensionMismatch(f"sums must have shape ({SINGLE_PASS_N}, n_bins, n_val); got {shp}")
    if tuple(np.shape(counts_3d)) != shp:
        raise DimensionMismatch("counts and sums must have the same shape")
    n_val = shp[2]
    aux = () if kind == "point" else u.shape[2:]
    if shp[3:] != tuple(aux):
        raise DimensionMismatch(f"sums must have shape {(SINGLE_PASS_N, nd, n_val) + tuple(aux)}; got {shp}")
    vlist = _value_bin_list(value_bins)
    _validate_value_bins(vlist, n_val)
    s, c = _single_pass_2d_raw(x, u, distance_bins, vlist, n_val, backend, distance_metric)
    sums_3d += s
    counts_3d += c.astype(counts_3d.dtype)
    return sums_3d, counts_3d
