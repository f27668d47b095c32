"""ctypes wrapper of the CPU oracle (oracle/libsforacle.so) for tests and bench.py's CPU baseline.

The oracle is test infrastructure: a restatement of the reference's SerialBackend, used only to
check the gfx950 results and to time a CPU baseline.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PATH = os.path.join(ROOT, "oracle", "libsforacle.so")

GEOM = {"flat": 0, "deg": 1, "rad": 2}


class BinRule(ctypes.Structure):
    """sfo_bin_rule: the fields of a LinearBinEdges (src/BinEdges.jl:94-100), optionally padded."""
    _fields_ = [("kind", ctypes.c_int32), ("padded", ctypes.c_int32), ("n", ctypes.c_int64),
                ("first_edge", ctypes.c_double), ("last_edge", ctypes.c_double),
                ("inv_step", ctypes.c_double), ("step_val", ctypes.c_double)]


class Prob(ctypes.Structure):
    _fields_ = [
        ("D", ctypes.c_int32), ("W", ctypes.c_int32), ("N", ctypes.c_int64),
        ("x", ctypes.c_void_p), ("u", ctypes.c_void_p),
        ("geometry", ctypes.c_int32), ("radius", ctypes.c_double),
        ("op", ctypes.c_int32), ("nl", ctypes.c_int32), ("nt", ctypes.c_int32),
        ("n_dist", ctypes.c_int32), ("dist_edges", ctypes.c_void_p),
        ("n_val", ctypes.c_int32), ("val_edges", ctypes.c_void_p),
        ("batch_leading", ctypes.c_int32), ("reserved", ctypes.c_int32),
        ("dist_rule", ctypes.c_void_p), ("val_rule", ctypes.c_void_p),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(PATH)
        L.sfo_hist.argtypes = [ctypes.POINTER(Prob), ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                               ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                               ctypes.c_void_p, ctypes.c_void_p]
        L.sfo_hist.restype = ctypes.c_int
        L.sfo_hist_single_pass.argtypes = [ctypes.POINTER(Prob), ctypes.c_int, ctypes.c_int64,
                                           ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p,
                                           ctypes.c_void_p]
        L.sfo_hist_single_pass.restype = ctypes.c_int
        L.sfo_hist_single_pass_2d.argtypes = [ctypes.POINTER(Prob), ctypes.c_int, ctypes.c_int64,
                                              ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_void_p]
        L.sfo_hist_single_pass_2d.restype = ctypes.c_int
        L.sfo_rule_digitize.argtypes = [ctypes.POINTER(BinRule), ctypes.c_double]
        L.sfo_rule_digitize.restype = ctypes.c_int64
        L.sfo_atan2.argtypes = [ctypes.c_double, ctypes.c_double]
        L.sfo_atan2.restype = ctypes.c_double
        L.sfo_sincosd.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.sfo_sincos.argtypes = L.sfo_sincosd.argtypes
        L.sfo_pair.argtypes = [ctypes.POINTER(Prob), ctypes.c_int64, ctypes.c_int64,
                               ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                               ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
        L.sfo_pair.restype = ctypes.c_int
        _lib = L
    return _lib


def _rule(fields):
    if fields is None:
        return None
    return BinRule(1, 1 if fields["padded"] else 0, int(fields["n"]), float(fields["first_edge"]),
                   float(fields["last_edge"]), float(fields["inv_step"]), float(fields["step_val"]))


def rule_digitize(fields, x):
    """digitize(x, axis) of a LinearBinEdges axis (dict of rule fields, sfamd.bins.rule_of)."""
    r = _rule(fields)
    return int(lib().sfo_rule_digitize(ctypes.byref(r), float(x)))


def make(x, u, edges, op=(0, 2, 0), geometry="flat", radius=1.0, vedges=None, batch_leading=False):
    """batch_leading: evaluate flat pairs as the reference's batch-leading kernels (fields with
    auxiliary axes) rather than its point-field kernels.  edges / vedges may be sfamd bin
    specifications: a LinearBinEdges axis is digitized by its rule."""
    from sfamd.bins import as_edges, rule_of
    keep = []
    xc = np.asfortranarray(x, dtype=np.float64)
    uc = np.asfortranarray(u, dtype=np.float64)
    drule = _rule(rule_of(edges))
    vrule = None if vedges is None else _rule(rule_of(vedges))
    de = np.ascontiguousarray(as_edges(edges), dtype=np.float64)
    if vedges is not None:
        vedges = as_edges(vedges)
    keep += [xc, uc, de, drule, vrule]
    p = Prob()
    p.D, p.W, p.N = uc.shape[0], xc.shape[0], xc.shape[1]
    p.x, p.u = xc.ctypes.data, uc.ctypes.data
    p.geometry = GEOM[geometry] if isinstance(geometry, str) else geometry
    p.radius = radius
    p.op, p.nl, p.nt = op
    p.n_dist, p.dist_edges = len(de) - 1, de.ctypes.data
    if vedges is None:
        p.n_val, p.val_edges = 0, None
    else:
        ve = np.ascontiguousarray(vedges, dtype=np.float64)
        keep.append(ve)
        p.n_val, p.val_edges = len(ve) - 1, ve.ctypes.data
    p.batch_leading = 1 if batch_leading else 0
    if drule is not None:
        p.dist_rule = ctypes.addressof(drule)
    if vrule is not None:
        p.val_rule = ctypes.addressof(vrule)
    return p, keep


def hist(x, u, edges, op=(0, 2, 0), geometry="flat", radius=1.0, vedges=None, nthreads=1,
         compensated=True, rows=None, with_mag=False, batch_leading=False):
    """Oracle histogram: returns (sums, counts, abs_sums[, mag_sums]); joint results are
    (n_dist, n_val).  abs_sums = sum |value|, mag_sums = sum ||du||^order per bin."""
    p, keep = make(x, u, edges, op, geometry, radius, vedges, batch_leading)
    nd = p.n_dist
    nv = p.n_val
    ncell = nd * max(nv, 1)
    s = np.zeros(ncell)
    c = np.zeros(ncell, dtype=np.uint64)
    a = np.zeros(ncell)
    g = np.zeros(ncell)
    i0, i1, st = (0, p.N, 1) if rows is None else rows
    rc = lib().sfo_hist(ctypes.byref(p), nthreads, 1 if compensated else 0, i0, i1, st,
                        s.ctypes.data, c.ctypes.data, a.ctypes.data, g.ctypes.data)
    if rc != 0:
        raise ValueError(f"oracle rejected the problem (code {rc})")
    out = [s, c, a, g]
    if nv:
        out = [v.reshape((nd, nv), order="F") for v in out]
    return tuple(out) if with_mag else tuple(out[:3])


def single_pass(x, u, edges, geometry="flat", radius=1.0, nthreads=1, rows=None, batch_leading=False):
    """Six invariants in one pass: (sums, counts), each (6, n_dist) rows S2, L2, T2, S3, L3, L1T2."""
    p, keep = make(x, u, edges, (1, 0, 0), geometry, radius, None, batch_leading)
    nd = p.n_dist
    s = np.zeros(6 * nd)
    c = np.zeros(6 * nd, dtype=np.uint64)
    i0, i1, st = (0, p.N, 1) if rows is None else rows
    rc = lib().sfo_hist_single_pass(ctypes.byref(p), nthreads, i0, i1, st, s.ctypes.data, c.ctypes.data)
    if rc != 0:
        raise ValueError(f"oracle rejected the problem (code {rc})")
    return s.reshape((6, nd), order="F"), c.reshape((6, nd), order="F")


def single_pass_2d(x, u, edges, vedges, geometry="flat", radius=1.0, nthreads=1, rows=None,
                   with_mag=False, batch_leading=False):
    """Single pass, 2D form: (sums, counts), each (6, n_dist, n_val); with_mag adds sum |value| and
    sum ||du||^order per cell.  vedges is one edge vector shared by the six invariants or a
    sequence of six (each at least n_val + 1 long)."""
    if len(vedges) == 6 and np.ndim(vedges[0]) == 1:
        # n_val from the first vector; the others need at least n_val + 1 edges and only their
        # first n_val + 1 matter (_validate_value_bins!, batch_api.jl:132-150)
        nv = len(vedges[0]) - 1
        ve6 = np.ascontiguousarray(np.concatenate([np.asarray(v, dtype=np.float64)[:nv + 1] for v in vedges]))
    else:
        v = np.asarray(vedges, dtype=np.float64)
        nv = len(v) - 1
        ve6 = np.ascontiguousarray(np.tile(v, 6))
    assert ve6.size == 6 * (nv + 1)
    p, keep = make(x, u, edges, (1, 0, 0), geometry, radius, None, batch_leading)
    p.n_val = nv
    nd = p.n_dist
    s = np.zeros(6 * nd * nv)
    c = np.zeros(6 * nd * nv, dtype=np.uint64)
    a = np.zeros(6 * nd * nv)
    g = np.zeros(6 * nd * nv)
    i0, i1, st = (0, p.N, 1) if rows is None else rows
    rc = lib().sfo_hist_single_pass_2d(ctypes.byref(p), nthreads, i0, i1, st, ve6.ctypes.data,
                                       s.ctypes.data, c.ctypes.data, a.ctypes.data, g.ctypes.data)
    if rc != 0:
        raise ValueError(f"oracle rejected the problem (code {rc})")
    shp = (6, nd, nv)
    out = (s.reshape(shp, order="F"), c.reshape(shp, order="F"))
    if with_mag:
        out += (a.reshape(shp, order="F"), g.reshape(shp, order="F"))
    return out


def pair(x, u, edges, i, j, op=(0, 2, 0), geometry="flat", radius=1.0, vedges=None, batch_leading=False):
    p, keep = make(x, u, edges, op, geometry, radius, vedges, batch_leading)
    d, v = ctypes.c_double(), ctypes.c_double()
    db, vb = ctypes.c_int64(), ctypes.c_int64()
    ok = lib().sfo_pair(ctypes.byref(p), i, j, ctypes.byref(d), ctypes.byref(v), ctypes.byref(db), ctypes.byref(vb))
    if ok < 0:
        raise ValueError("bad pair")
    return bool(ok), d.value, v.value, db.value, vb.value


def sincosd(x):
    s, c = ctypes.c_double(), ctypes.c_double()
    lib().sfo_sincosd(x, ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def sincos(x):
    s, c = ctypes.c_double(), ctypes.c_double()
    lib().sfo_sincos(x, ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def atan2(y, x):
    return lib().sfo_atan2(y, x)
