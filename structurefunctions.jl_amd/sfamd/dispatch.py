"""Public entry points mirroring the reference's calculation API.

    calculate_structure_function(sf, x, u, distance_bins[, value_bins][, count_eltype];
                                 backend, output_type, distance_metric)
        src/Calculations/dispatch.jl:64-136
    calculate_structure_function_inplace(sums, counts, sf, x, u, distance_bins[, value_bins]; ...)
        calculate_structure_function!  (dispatch.jl:346-373), accumulates
    calculate_structure_function_batch(sf, xs, us, distance_bins[, value_bins]; ...)
        ragged batches of independent point fields in one persistent launch
    calculate_structure_functions_fused(outputs, x, u, distance_bins; ...)
        several operators (and at most one joint histogram) from one evaluation of every pair

Arrays follow the reference's shapes: ``x`` is (W, N) and ``u`` is (D, N) (or with trailing
auxiliary axes, shapes.jl:67-105).  The only backend is :class:`MI355XBackend`, which calls the
gfx950 library through its C ABI; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from .bins import as_edges, rule_of
from .errors import ArgumentError, DimensionMismatch
from .metrics import Euclidean, geometry_of
from .objects import (StructureFunction, StructureFunction2DSumsAndCounts,
                      StructureFunctionSumsAndCounts, _bin_average)
from .operators import AbstractDerivedStructureFunctionType, AbstractPairwiseStructureFunctionType

_CONTEXTS: dict = {}


class MI355XBackend:
    """Execution backend running the pair histogram on one MI355X (gfx950) GPU.

    ``precision="fp64"`` is the drop-in binary64 kernel (counts bit-identical to the oracle's
    restatement of SerialBackend; against Julia itself only up to the ulp-level differences of its
    sincosd/atan and LogBinEdges' log-grid lookup, which no test here can pin);
    ``precision="mixed"`` runs the pair arithmetic in binary32, re-bins every pair whose distance
    or value lies within the binary32 error bound of a bin edge in binary64 (counts still
    bit-identical), and accumulates in binary64.  ``precision="mixed_fixed"`` is "mixed" with
    joint cells of uniform value bins kept as 64-bit fixed-point words on the device: counts
    bit-identical, each pair's value entering its sum with an extra absolute error of at most
    2^-23 of its value bin's width.
    """

    _CODES = {"fp64": _native.PREC_FP64, "mixed": _native.PREC_MIXED,
              "mixed_fixed": _native.PREC_MIXED_FIXED}

    def __init__(self, device: int = 0, precision: str = "fp64"):
        if precision not in self._CODES:
            raise ArgumentError("precision must be 'fp64', 'mixed' or 'mixed_fixed'")
        self.device = int(device)
        self.precision = precision

    @property
    def precision_code(self):
        return self._CODES[self.precision]

    def context(self) -> _native.Context:
        ctx = _CONTEXTS.get(self.device)
        if ctx is None:
            ctx = _native.Context(self.device)
            _CONTEXTS[self.device] = ctx
        return ctx

    def __repr__(self):
        return f"MI355XBackend(device={self.device}, precision={self.precision!r})"


def _derived_error(sf):
    raise ArgumentError(
        f"{type(sf).__name__} is a derived structure-function quantity, not a pairwise operator. "
        "Use helmholtz_decompose_2d for 2D Helmholtz rotational/divergent quantities.")


def _validate_array_shape(x, u, metric):
    """src/Calculations/shapes.jl:67-105; returns ('point'|'shared'|'varying', D, W)."""
    if x.ndim < 2:
        raise DimensionMismatch(f"x must have shape (Dx, N) or (Dx, N, auxiliary...); got ndims(x)={x.ndim}")
    if u.ndim < 2:
        raise DimensionMismatch(f"u must have shape (D, N) or (D, N, auxiliary...); got ndims(u)={u.ndim}")
    D = u.shape[0]
    if D not in (2, 3):
        raise DimensionMismatch(f"expected spatial/velocity dimension D=2 or D=3 on axis 1; got D={D}")
    geom, radius, W = geometry_of(metric)
    W = D if W is None else W
    if x.shape[0] != W:
        raise DimensionMismatch(
            f"this geometry locates a point with {W} coordinate(s) on axis 1 of x, but got "
            f"size(x,1)={x.shape[0]} (velocity dimension D={D} from size(u,1))")
    if u.shape[1] != x.shape[1]:
        raise DimensionMismatch(
            f"x and u must share axis-2 point count N; got size(x,2)={x.shape[1]} and size(u,2)={u.shape[1]}")
    if x.ndim == 2 and u.ndim == 2:
        return "point", D, W
    if x.ndim == 2 and u.ndim >= 3:
        return "shared", D, W
    if x.ndim != u.ndim:
        raise DimensionMismatch(
            f"varying-position inputs must have the same rank; got ndims(x)={x.ndim} and ndims(u)={u.ndim}")
    if x.shape[2:] != u.shape[2:]:
        raise DimensionMismatch(
            f"varying-position inputs must share auxiliary axes; got {x.shape[2:]} and {u.shape[2:]}")
    return "varying", D, W


def _assert_counts_representable(count_dtype, n_points: int):
    """src/Calculations/serial.jl:111-120."""
    n_pairs = n_points * (n_points - 1) // 2
    tmax = int(np.iinfo(count_dtype).max)
    if n_pairs > tmax:
        raise ArgumentError(
            f"count_eltype={np.dtype(count_dtype).name} cannot represent the worst-case pair count "
            f"{n_pairs} for N={n_points} (typemax = {tmax}); pass count_eltype=UInt64 or Int64.")


def _attach_rules(p, drule, vrule, keep):
    """Point the problem at its axes' LinearBinEdges rules (None: searched axis)."""
    for name, fields in (("dist_rule", drule), ("val_rule", vrule)):
        r = _native.make_rule(fields)
        if r is not None:
            keep.append(r)
            setattr(p, name, ctypes.addressof(r))


def _problem(sf, x2, u2, dedges, vedges, metric, precision, keep, semantics=_native.SEM_POINT,
             drule=None, vrule=None):
    """Build one SfProblem for a (W, N) / (D, N) slice; keep references alive in `keep`."""
    geom, radius, _ = geometry_of(metric)
    xc = np.asfortranarray(x2, dtype=np.float64)
    uc = np.asfortranarray(u2, dtype=np.float64)
    keep += [xc, uc, dedges] + ([vedges] if vedges is not None else [])
    kind, nl, nt = sf.code
    p = _native.SfProblem()
    p.D, p.W, p.N = uc.shape[0], xc.shape[0], xc.shape[1]
    p.x, p.u = xc.ctypes.data, uc.ctypes.data
    p.geometry, p.radius = geom, radius
    p.op, p.nl, p.nt = kind, nl, nt
    p.n_dist, p.dist_edges = len(dedges) - 1, dedges.ctypes.data
    if vedges is None:
        p.n_val, p.val_edges = 0, None
    elif kind == _native.OP_SINGLE_PASS:
        # the single pass's 2D form: six edge vectors back to back
        p.n_val, p.val_edges = len(vedges) // 6 - 1, vedges.ctypes.data
    else:
        p.n_val, p.val_edges = len(vedges) - 1, vedges.ctypes.data
    p.precision = precision
    p.count_bits = 64
    p.semantics = semantics
    _attach_rules(p, drule, vrule, keep)
    return p


def _slices(x, u, kind):
    """Split auxiliary axes into (W,N)/(D,N) slices; returns (list of (x2,u2), aux shape)."""
    if kind == "point":
        return [(x, u)], ()
    aux = u.shape[2:]
    B = int(np.prod(aux))
    uu = u.reshape(u.shape[0], u.shape[1], B, order="F")
    if kind == "shared":
        return [(x, uu[:, :, b]) for b in range(B)], aux
    xx = x.reshape(x.shape[0], x.shape[1], B, order="F")
    return [(xx[:, :, b], uu[:, :, b]) for b in range(B)], aux


def _semantics(kind):
    """Fields with auxiliary axes go through the reference's batch-leading kernels
    (batch.jl:201-217, batch_leading.jl:105-330), point fields through its point-field kernels."""
    return _native.SEM_POINT if kind == "point" else _native.SEM_BATCH


def _run(sf, slices, dedges_list, vedges_list, metric, backend, semantics=_native.SEM_POINT,
         drules=None, vrules=None):
    keep = []
    n = len(slices)
    drules = drules or [None] * n
    vrules = vrules or [None] * n
    problems = [_problem(sf, xs, us, de, ve, metric, backend.precision_code, keep, semantics, dr, vr)
                for (xs, us), de, ve, dr, vr in zip(slices, dedges_list, vedges_list, drules, vrules)]
    nd = len(dedges_list[0]) - 1
    sp = sf.code[0] == _native.OP_SINGLE_PASS
    if vedges_list[0] is None:
        nv = 0
    else:
        nv = (len(vedges_list[0]) // 6 if sp else len(vedges_list[0])) - 1
    ncell = (6 if sp else 1) * nd * max(nv, 1)
    sums = np.zeros(len(slices) * ncell, dtype=np.float64)
    counts = np.zeros(len(slices) * ncell, dtype=np.uint64)
    backend.context().hist(problems, sums, counts)
    return sums, counts, nd, nv


def _finalize(raw, output_type):
    """src/Calculations/dispatch.jl:20-29."""
    if isinstance(raw, StructureFunctionSumsAndCounts):
        if output_type in (None, StructureFunction):
            return StructureFunction(raw.operator, raw.distance, _bin_average(raw.sums, raw.counts))
        if output_type is StructureFunctionSumsAndCounts:
            return raw
    if isinstance(raw, StructureFunction2DSumsAndCounts):
        if output_type in (None, StructureFunction2DSumsAndCounts):
            return raw
    raise ArgumentError(
        f"Cannot produce a {getattr(output_type, '__name__', output_type)} from this calculation "
        f"(got a {type(raw).__name__}). Check the `output_type` keyword.")


def _normalize_args(value_bins, count_eltype):
    # Julia dispatch: a Type in the value_bins position is the count_eltype (1D form)
    if value_bins is not None and isinstance(value_bins, type) and issubclass(value_bins, np.integer):
        return None, value_bins
    if value_bins is not None and isinstance(value_bins, np.dtype):
        return None, value_bins.type
    return value_bins, count_eltype


def calculate_structure_function(sf, x, u, distance_bins, value_bins=None, count_eltype=np.uint32, *,
                                 backend: MI355XBackend | None = None, output_type=None,
                                 distance_metric=Euclidean(), verbose=False, show_progress=False):
    """Pair histogram of ``sf`` over all unordered pairs of points (dispatch.jl:64-136)."""
    if isinstance(sf, AbstractDerivedStructureFunctionType):
        _derived_error(sf)
    if not isinstance(sf, AbstractPairwiseStructureFunctionType):
        raise ArgumentError(f"{sf!r} is not a structure-function operator")
    if isinstance(x, tuple) or isinstance(u, tuple):
        raise ArgumentError("tuple-of-component-vector inputs are not part of the stabilized public "
                            "API; pass arrays with shape (D, N) or (D, N, auxiliary...) instead")
    value_bins, count_eltype = _normalize_args(value_bins, count_eltype)
    backend = backend or MI355XBackend()
    x = np.asarray(x, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)
    kind, D, W = _validate_array_shape(x, u, distance_metric)
    _assert_counts_representable(count_eltype, x.shape[1])
    de = as_edges(distance_bins)
    ve = None if value_bins is None else as_edges(value_bins)
    dr = rule_of(distance_bins)
    vr = None if value_bins is None else rule_of(value_bins)
    slices, aux = _slices(x, u, kind)
    sums, counts, nd, nv = _run(sf, slices, [de] * len(slices), [ve] * len(slices), distance_metric, backend,
                                _semantics(kind), [dr] * len(slices), [vr] * len(slices))
    raw = _raw_result(sf, sums, counts, nd, nv, len(slices), aux, distance_bins, value_bins, count_eltype)
    return _finalize(raw, output_type)


def _raw_result(sf, sums, counts, nd, nv, B, aux, distance_bins, value_bins, count_eltype):
    """Slice-major flat sums/counts (B slices of nd or nd x nv cells) as the reference's raw
    result with the auxiliary axes last."""
    if value_bins is None:
        s = sums.reshape(B, nd).T.reshape((nd,) + aux, order="F")
        c = counts.reshape(B, nd).T.reshape((nd,) + aux, order="F").astype(count_eltype)
        return StructureFunctionSumsAndCounts(sf, distance_bins, s, c)
    s = np.stack([sums[b * nd * nv:(b + 1) * nd * nv].reshape((nd, nv), order="F") for b in range(B)], axis=-1)
    c = np.stack([counts[b * nd * nv:(b + 1) * nd * nv].reshape((nd, nv), order="F") for b in range(B)], axis=-1)
    s = s.reshape((nd, nv) + aux, order="F")
    c = c.reshape((nd, nv) + aux, order="F").astype(count_eltype)
    return StructureFunction2DSumsAndCounts(sf, distance_bins, value_bins, s, c)


def calculate_structure_functions_fused(outputs, x, u, distance_bins, count_eltype=np.uint32, *,
                                        backend: MI355XBackend | None = None, output_type=None,
                                        distance_metric=Euclidean()):
    """Several pair histograms of one field from a single evaluation of every pair.

    ``outputs`` lists operators; an entry ``(operator, value_bins)`` asks for that operator's joint
    distance x value histogram (at most one such entry).  The result is the list of what
    ``calculate_structure_function(operator, x, u, distance_bins[, value_bins])`` returns for each
    entry, in order: counts identical, sums to the backend precision's bound.  The reference
    computes each operator with its own call (dispatch.jl:64-136); this form shares the pair
    geometry, the longitudinal/transverse frame and the increments across the outputs
    (sfmi355x.h, sf_hist_multi).  At most 8 one-dimensional outputs plus one joint output.
    """
    entries = list(outputs)
    if not entries:
        raise ArgumentError("outputs must name at least one operator")
    ops, vbins = [], []
    for e in entries:
        sf, vb = (e if isinstance(e, tuple) else (e, None))
        if isinstance(sf, AbstractDerivedStructureFunctionType):
            _derived_error(sf)
        if not isinstance(sf, AbstractPairwiseStructureFunctionType):
            raise ArgumentError(f"{sf!r} is not a structure-function operator")
        if sf.code[0] == _native.OP_SINGLE_PASS:
            raise ArgumentError("the single pass is its own fused form; call "
                                "calculate_structure_functions_single_pass instead")
        ops.append(sf)
        vbins.append(vb)
    if sum(v is not None for v in vbins) > 1:
        raise ArgumentError("a fused call takes at most one joint (operator, value_bins) output")
    if sum(v is None for v in vbins) > _native.SF_MAX_FUSED:
        raise ArgumentError(f"a fused call takes at most {_native.SF_MAX_FUSED} one-dimensional outputs")
    backend = backend or MI355XBackend()
    x = np.asarray(x, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)
    kind, D, W = _validate_array_shape(x, u, distance_metric)
    _assert_counts_representable(count_eltype, x.shape[1])
    de = as_edges(distance_bins)
    dr = rule_of(distance_bins)
    slices, aux = _slices(x, u, kind)
    keep = []
    problems = [_problem(ops[0], xs, us, de, None, distance_metric, backend.precision_code, keep,
                         _semantics(kind), dr, None) for xs, us in slices]
    nd = len(de) - 1
    outs, sums, counts, nvs = [], [], [], []
    for sf, vb in zip(ops, vbins):
        o = _native.SfOutput()
        o.op, o.nl, o.nt = sf.code
        nv = 0
        if vb is not None:
            ve = as_edges(vb)
            keep.append(ve)
            nv = len(ve) - 1
            o.n_val, o.val_edges = nv, ve.ctypes.data
            r = _native.make_rule(rule_of(vb))
            if r is not None:
                keep.append(r)
                o.val_rule = ctypes.addressof(r)
        outs.append(o)
        nvs.append(nv)
        sums.append(np.zeros(len(slices) * nd * max(nv, 1), dtype=np.float64))
        counts.append(np.zeros(len(slices) * nd * max(nv, 1), dtype=np.uint64))
    backend.context().hist_multi(problems, outs, sums, counts)
    res = []
    for sf, vb, nv, s, c in zip(ops, vbins, nvs, sums, counts):
        raw = _raw_result(sf, s, c, nd, nv, len(slices), aux, distance_bins, vb, count_eltype)
        res.append(_finalize(raw, output_type if vb is None else None))
    return res


def calculate_structure_function_inplace(sums, counts, sf, x, u, distance_bins, value_bins=None, *,
                                         backend: MI355XBackend | None = None,
                                         distance_metric=Euclidean(), verbose=False, show_progress=False):
    """calculate_structure_function! (dispatch.jl:346-373): accumulates into sums/counts."""
    if isinstance(sf, AbstractDerivedStructureFunctionType):
        _derived_error(sf)
    backend = backend or MI355XBackend()
    x = np.asarray(x, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)
    kind, D, W = _validate_array_shape(x, u, distance_metric)
    _assert_counts_representable(np.asarray(counts).dtype.type, x.shape[1])
    de = as_edges(distance_bins)
    ve = None if value_bins is None else as_edges(value_bins)
    nd = len(de) - 1
    want = (nd,) if ve is None else (nd, len(ve) - 1)
    slices, aux = _slices(x, u, kind)
    if tuple(np.shape(sums)) != want + aux or tuple(np.shape(counts)) != want + aux:
        raise DimensionMismatch(f"sums/counts must have shape {want + aux}; got {np.shape(sums)}, {np.shape(counts)}")
    dr = rule_of(distance_bins)
    vr = None if value_bins is None else rule_of(value_bins)
    s, c, nd, nv = _run(sf, slices, [de] * len(slices), [ve] * len(slices), distance_metric, backend,
                        _semantics(kind), [dr] * len(slices), [vr] * len(slices))
    B = len(slices)
    ncell = nd * max(nv, 1)
    s = np.stack([s[b * ncell:(b + 1) * ncell].reshape(want, order="F") for b in range(B)], axis=-1)
    c = np.stack([c[b * ncell:(b + 1) * ncell].reshape(want, order="F") for b in range(B)], axis=-1)
    sums += s.reshape(want + aux, order="F")
    counts += c.reshape(want + aux, order="F").astype(counts.dtype)
    return None


def calculate_structure_function_batch(sf, xs, us, distance_bins, value_bins=None, *,
                                       count_eltype=np.uint64, backend: MI355XBackend | None = None,
                                       distance_metric=Euclidean()):
    """Independent point fields of different N in one persistent launch.

    ``distance_bins`` / ``value_bins`` are either shared or one edge vector per slice.  Returns
    one raw result (StructureFunctionSumsAndCounts or StructureFunction2DSumsAndCounts) per slice.
    """
    if isinstance(sf, AbstractDerivedStructureFunctionType):
        _derived_error(sf)
    backend = backend or MI355XBackend()
    n = len(xs)
    if len(us) != n:
        raise DimensionMismatch("xs and us must have the same number of slices")

    def per_slice(b):
        if b is None:
            return [None] * n, [None] * n
        if isinstance(b, (list, tuple)) and len(b) == n and (np.ndim(b[0]) == 1 or hasattr(b[0], "edges")):
            return [as_edges(e) for e in b], [rule_of(e) for e in b]
        e = as_edges(b)
        return [e] * n, [rule_of(b)] * n

    (des, drs), (ves, vrs) = per_slice(distance_bins), per_slice(value_bins)
    slices = []
    for x, u in zip(xs, us):
        x = np.asarray(x, dtype=np.float64)
        u = np.asarray(u, dtype=np.float64)
        kind, _, _ = _validate_array_shape(x, u, distance_metric)
        if kind != "point":
            raise DimensionMismatch("batch slices must be (W, N) / (D, N) point fields")
        _assert_counts_representable(count_eltype, x.shape[1])
        slices.append((x, u))
    if n == 0:
        return []
    sums, counts, nd, nv = _run(sf, slices, des, ves, distance_metric, backend, _native.SEM_POINT, drs, vrs)
    ncell = nd * max(nv, 1)
    out = []
    for b in range(n):
        s = sums[b * ncell:(b + 1) * ncell]
        c = counts[b * ncell:(b + 1) * ncell].astype(count_eltype)
        if ves[b] is None:
            out.append(StructureFunctionSumsAndCounts(sf, des[b], s, c))
        else:
            out.append(StructureFunction2DSumsAndCounts(sf, des[b], ves[b], s.reshape((nd, nv), order="F"),
                                                        c.reshape((nd, nv), order="F")))
    return out
