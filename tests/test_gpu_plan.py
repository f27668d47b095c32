"""GPU: plans and hipGraph replay.  Each replay runs after the slices' data were rewritten in
place and after unrelated eager GPU work, and is checked against an eager call on the same data
(and the first against the oracle)."""
import numpy as np
import pytest

import _oracle
from sfamd import _native

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("joint", [False, True])
def test_plan_graph_replays_match_eager(ctx, prec, joint):
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(11)
    sizes = [700, 64, 1800, 1]
    e = np.geomspace(1e-3, 1.5, 25)
    ve = np.linspace(-3, 3, 19) if joint else None
    nd, nv = len(e) - 1, (len(ve) - 1 if joint else 0)
    ncell = nd * max(nv, 1)
    xs = [torch.zeros((n, 2), dtype=torch.float64, device=dev) for n in sizes]
    us = [torch.zeros((n, 2), dtype=torch.float64, device=dev) for n in sizes]

    def problems():
        out = []
        for x, u in zip(xs, us):
            p = _native.SfProblem()
            p.D, p.W, p.N = 2, 2, x.shape[0]
            p.x, p.u = x.data_ptr(), u.data_ptr()
            p.geometry, p.radius, p.op, p.nl, p.nt = 0, 1.0, 0, 1, 0
            p.n_dist, p.dist_edges = nd, e.ctypes.data
            p.n_val, p.val_edges = (nv, ve.ctypes.data) if joint else (0, None)
            p.precision, p.count_bits = prec, 64
            out.append(p)
        return out

    plan = _native.Plan(ctx, problems())
    stream = torch.cuda.current_stream().cuda_stream
    g_s = torch.zeros(len(sizes) * ncell, dtype=torch.float64, device=dev)
    g_c = torch.zeros(len(sizes) * ncell, dtype=torch.int64, device=dev)
    e_s = torch.zeros_like(g_s)
    e_c = torch.zeros_like(g_c)
    for rep in range(3):
        hx = [rng.uniform(0, 1, (n, 2)) for n in sizes]
        hu = [rng.normal(size=(n, 2)) for n in sizes]
        for x, u, a, b in zip(xs, us, hx, hu):
            x.copy_(torch.from_numpy(a))
            u.copy_(torch.from_numpy(b))
        junk = torch.randn(1 << 20, device=dev).sum()  # unrelated eager work between replays
        g_s.zero_()
        g_c.zero_()
        plan.execute(g_s.data_ptr(), g_c.data_ptr(), stream, graph=True)
        assert not plan.sync()
        e_s.zero_()
        e_c.zero_()
        ctx.hist_device(problems(), e_s.data_ptr(), e_c.data_ptr(), stream)
        torch.cuda.synchronize()
        assert np.array_equal(g_c.cpu().numpy(), e_c.cpu().numpy()), rep
        np.testing.assert_allclose(g_s.cpu().numpy(), e_s.cpu().numpy(), rtol=1e-12, atol=1e-12)
        if rep == 0:
            for k, n in enumerate(sizes):
                s, c, a = _oracle.hist(hx[k].T, hu[k].T, e, op=(0, 1, 0), vedges=ve)
                gc = g_c.cpu().numpy()[k * ncell:(k + 1) * ncell].astype(np.uint64)
                assert np.array_equal(gc, c.ravel(order="F")), (k, n)
        float(junk)
    plan.close()


def test_autotune_then_call_matches_fp64(ctx):
    """sf_autotune picks a launch shape for a problem class and later calls with a matching key use
    it; whatever it picks, counts stay equal to the binary64 kernel's."""
    torch = pytest.importorskip("torch")
    from sfamd import synthetic
    x, u, de, ve = synthetic.config4(N=20000)
    dev = torch.device("cuda:0")
    dx = torch.from_numpy(np.ascontiguousarray(x.T)).to(dev)
    du = torch.from_numpy(np.ascontiguousarray(u.T)).to(dev)
    nd, nv = len(de) - 1, len(ve) - 1

    def prob(prec):
        p = _native.SfProblem()
        p.D, p.W, p.N = 2, 2, x.shape[1]
        p.x, p.u = dx.data_ptr(), du.data_ptr()
        p.geometry, p.radius = 1, 1.0
        p.op, p.nl, p.nt = 0, 1, 0
        p.n_dist, p.dist_edges = nd, de.ctypes.data
        p.n_val, p.val_edges = nv, ve.ctypes.data
        p.precision, p.count_bits = prec, 64
        return p

    block, jchunk, bpc, variant, copies, layout = ctx.autotune([prob(1)])
    assert block in (256, 512, 768, 1024) and variant in (0, 1, 2) and jchunk >= 0
    assert copies in (0, 1, 2) and layout in (0, 1, 2)
    outs = {}
    for prec in (0, 1):
        s = torch.zeros(nd * nv, dtype=torch.float64, device=dev)
        c = torch.zeros(nd * nv, dtype=torch.int64, device=dev)
        ctx.hist_device([prob(prec)], s.data_ptr(), c.data_ptr(), 0)
        torch.cuda.synchronize()
        outs[prec] = (s.cpu().numpy(), c.cpu().numpy())
    assert ctx.stats()["block"] == block
    assert np.array_equal(outs[0][1], outs[1][1])
    ok = outs[0][1] > 0
    assert np.all(np.abs(outs[1][0] - outs[0][0])[ok] <= 1e-5 * outs[0][1][ok] * 4.0)
