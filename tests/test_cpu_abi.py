"""CPU checks of the C ABI: the library loads, exports what include/sfmi355x.h declares, and
validates inputs before any device work (sf_check needs no GPU)."""
import os
import re

import numpy as np
import pytest

from sfamd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "sfmi355x.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*(sf_\w+)\(", txt, re.M)))


def test_exports_match_header():
    L = _native.lib()
    names = declared_symbols()
    assert len(names) >= 10
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_native.EXPORTED)


def prob(**kw):
    x = np.random.default_rng(0).uniform(0, 1, (2, 10))
    u = np.zeros((2, 10))
    e = np.geomspace(0.01, 1.5, 9)
    d = dict(D=2, W=2, N=10, geometry=0, radius=1.0, op=0, nl=2, nt=0, precision=0, count_bits=32)
    d.update(kw)
    p = _native.SfProblem()
    keep = [np.asfortranarray(x), np.asfortranarray(u), e]
    p.x, p.u = keep[0].ctypes.data, keep[1].ctypes.data
    p.n_dist, p.dist_edges = len(e) - 1, e.ctypes.data
    p.n_val, p.val_edges = 0, None
    for k, v in d.items():
        setattr(p, k, v)
    return p, keep


def test_check_valid():
    p, keep = prob()
    assert _native.check_problems([p])[0] == 0


@pytest.mark.parametrize("kw,code", [
    (dict(D=4, W=4), _native.SF_E_DIM),
    (dict(W=3), _native.SF_E_DIM),
    (dict(geometry=1, W=3), _native.SF_E_DIM),
    (dict(geometry=7), _native.SF_E_ARG),
    (dict(geometry=1, radius=-1.0), _native.SF_E_ARG),
    (dict(op=9), _native.SF_E_ARG),
    (dict(nl=3, nt=1), _native.SF_E_UNSUPPORTED),
    (dict(op=5, nl=4), _native.SF_E_UNSUPPORTED),
    (dict(precision=5), _native.SF_E_ARG),
    (dict(count_bits=16), _native.SF_E_ARG),
    (dict(semantics=2), _native.SF_E_ARG),
])
def test_check_rejects(kw, code):
    p, keep = prob(**kw)
    rc, msg = _native.check_problems([p])
    assert rc == code, msg
    assert msg


def test_single_pass_checks():
    """SF_OP_SINGLE_PASS: distance bins only (1D form), or with value bins the 2D form, whose
    val_edges holds six edge vectors back to back, each validated (sorted, no NaN)."""
    p, keep = prob(op=_native.OP_SINGLE_PASS, nl=0)
    assert _native.check_problems([p])[0] == 0
    ve6 = np.concatenate([np.linspace(-1, 1, 5)] * 6)
    keep.append(ve6)
    p.n_val, p.val_edges = 4, ve6.ctypes.data
    assert _native.check_problems([p])[0] == 0
    bad = ve6.copy()
    bad[5 * 5 + 2] = np.nan            # the last invariant's vector
    keep.append(bad)
    p.val_edges = bad.ctypes.data
    rc, msg = _native.check_problems([p])
    assert rc == _native.SF_E_ARG and "NaN" in msg
    bad2 = ve6.copy()
    bad2[3 * 5 + 1], bad2[3 * 5 + 2] = 0.9, -0.9   # the fourth vector unsorted
    keep.append(bad2)
    p.val_edges = bad2.ctypes.data
    rc, msg = _native.check_problems([p])
    assert rc == _native.SF_E_ARG and "sorted" in msg


def test_count_width_guard():
    """UInt32 counts: N = 92682 passes, 92683 is rejected (test/test_core_correctness.jl:225-249)."""
    p, keep = prob(N=92682)
    assert _native.check_problems([p])[0] == 0
    p, keep = prob(N=92683)
    rc, msg = _native.check_problems([p])
    assert rc == _native.SF_E_COUNT and "UInt64" in msg
    p, keep = prob(N=92683, count_bits=64)
    assert _native.check_problems([p])[0] == 0


def test_unsorted_and_nan_edges_rejected():
    p, keep = prob()
    e = np.array([0.1, 0.5, 0.3])
    keep.append(e)
    p.n_dist, p.dist_edges = 2, e.ctypes.data
    assert _native.check_problems([p])[0] == _native.SF_E_ARG
    e2 = np.array([0.1, np.nan, 0.3])
    keep.append(e2)
    p.dist_edges = e2.ctypes.data
    assert _native.check_problems([p])[0] == _native.SF_E_ARG
    e3 = np.array([-0.5, 0.5, 1.0])
    keep.append(e3)
    p.dist_edges = e3.ctypes.data
    assert _native.check_problems([p])[0] == _native.SF_E_ARG


def test_batch_mismatch_rejected():
    p1, k1 = prob()
    p2, k2 = prob(nl=3)
    assert _native.check_problems([p1, p2])[0] == _native.SF_E_ARG
    p3, k3 = prob(semantics=_native.SEM_BATCH)
    assert _native.check_problems([p3])[0] == 0
    assert _native.check_problems([p1, p3])[0] == _native.SF_E_ARG
