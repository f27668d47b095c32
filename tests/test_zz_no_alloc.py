"""The allocation-free entry (sf_hist_batch_device_noalloc) against the allocating device entry
(sf_hist_batch_device) through the C ABI: two calls in a row, the second smaller, then a call too
large for the scratch array, which is refused before any device work."""
import numpy as np
import pytest

import _oracle
from sfamd import _native, synthetic

pytestmark = pytest.mark.gpu


def _problem(dx, du, N, D, W, geom, op, de, ve, prec):
    p = _native.SfProblem()
    p.D, p.W, p.N = D, W, N
    p.x, p.u = dx, du
    p.geometry, p.radius = geom, 1.0
    p.op, p.nl, p.nt = op
    p.n_dist, p.dist_edges = len(de) - 1, de.ctypes.data
    if ve is None:
        p.n_val, p.val_edges = 0, None
    else:
        p.n_val, p.val_edges = len(ve) - 1, ve.ctypes.data
    p.precision = prec
    p.count_bits = 64
    return p


def _both(ctx, torch, x, u, geom, op, de, ve, prec):
    dev = torch.device("cuda:0")
    dx = torch.from_numpy(np.ascontiguousarray(x.T)).to(dev)
    du = torch.from_numpy(np.ascontiguousarray(u.T)).to(dev)
    N, D, W = x.shape[1], u.shape[0], x.shape[0]
    ncell = (len(de) - 1) * (1 if ve is None else len(ve) - 1)
    p = _problem(dx.data_ptr(), du.data_ptr(), N, D, W, geom, op, de, ve, prec)
    out = []
    stream = torch.cuda.current_stream().cuda_stream
    for noalloc in (False, True):
        s = torch.zeros(ncell, dtype=torch.float64, device=dev)
        c = torch.zeros(ncell, dtype=torch.int64, device=dev)
        if noalloc:
            ctx.hist_device_noalloc([p], s.data_ptr(), c.data_ptr(), stream)
        else:
            ctx.hist_device([p], s.data_ptr(), c.data_ptr(), stream)
        torch.cuda.synchronize()
        out.append((s.cpu().numpy(), c.cpu().numpy().astype(np.uint64)))
    return out


def test_noalloc_matches_device_entry(ctx):
    torch = pytest.importorskip("torch")
    # first call: the spherical joint histogram in the fixed-point precision
    x, u, de, ve = synthetic.config4(N=20000)
    (s1, c1), (s2, c2) = _both(ctx, torch, x, u, 1, (0, 1, 0), de, ve, 2)
    assert np.array_equal(c1, c2) and c1.sum() > 0
    np.testing.assert_allclose(s2, s1, rtol=1e-12, atol=1e-12 * np.abs(s1).max())
    # second call, smaller: flat 2D 1D du_L^2 in the strict mixed precision, and the oracle
    x, u, de = synthetic.config1(N=3000)
    (s1, c1), (s2, c2) = _both(ctx, torch, x, u, 0, (0, 2, 0), de, None, 1)
    so, co, ao = _oracle.hist(x, u, de, op=(0, 2, 0), nthreads=8)
    assert np.array_equal(c2, co) and np.array_equal(c1, c2)
    assert np.all(np.abs(s2 - so) <= 1e-6 * np.maximum(ao, 1e-300) + 1e-300)
    np.testing.assert_allclose(s2, s1, rtol=1e-12, atol=1e-12 * np.abs(s1).max())


def test_noalloc_refuses_oversize(ctx):
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    N = 1_000_000  # sphere records of a million points exceed half of the 128 MiB scratch
    dx = torch.zeros(2 * N, dtype=torch.float64, device=dev)
    du = torch.zeros(2 * N, dtype=torch.float64, device=dev)
    de = np.geomspace(1e-3, np.pi, 9)
    s = torch.zeros(8, dtype=torch.float64, device=dev)
    c = torch.zeros(8, dtype=torch.int64, device=dev)
    p = _problem(dx.data_ptr(), du.data_ptr(), N, 2, 2, 1, (0, 2, 0), de, None, 1)
    with pytest.raises(Exception, match="allocation-free entry"):
        ctx.hist_device_noalloc([p], s.data_ptr(), c.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert int(c.sum().item()) == 0
