"""Spatial sort and tile culling (binary32 kernels, one-shot calls) against the oracle.

When the last distance edge is finite and small against the domain, the library reorders the
binary32 records along a Morton curve and drops work items whose tiles are farther apart than
that edge.  Results must not change: counts bit-identical to the oracle (SerialBackend order),
sums within the mixed-kernel bound, and the FP64 re-bin of queued pairs must see the caller's
pair order.
"""
import numpy as np
import pytest

import _oracle
from sfamd import synthetic
from test_gpu_parity import MAG, TOL, check, gpu_hist

pytestmark = pytest.mark.gpu


def flat_case(N, D, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1, (D, N))
    u = synthetic._flat_field(rng, x, 32, 2 * np.pi, 2 * np.pi * 32, -5 / 6)
    return x, u


@pytest.mark.parametrize("prec", ["mixed", "fixed"])
@pytest.mark.parametrize("D", [2, 3])
def test_cull_flat_matches_oracle(ctx, prec, D):
    x, u = flat_case(20000, D, 70 + D)
    edges = np.geomspace(1e-3, 0.05, 20)
    for op, ve in [((0, 2, 0), None), ((0, 1, 0), np.linspace(-3, 3, 25)), ((0, 0, 3), None)]:
        if prec == "fixed" and ve is None:
            continue
        os_, oc, oa, og = _oracle.hist(x, u, edges, op=op, vedges=ve, nthreads=8, with_mag=True)
        gs, gc = gpu_hist(ctx, x, u, edges, op, vedges=ve, prec=prec)
        items_cull = ctx.stats()["items"]
        label = f"flat D={D} op={op} {prec}"
        if prec == "mixed":
            check(gs, gc, os_, oc, oa, TOL["mixed"], label, og=og, mag=MAG["mixed"])
        else:
            # fixed-point joint cells: 2^-23 value-bin widths per pair on top (factor 1.5 spare)
            assert np.array_equal(gc, oc), label
            w = ve[1] - ve[0]
            bound = TOL["mixed"] * oa + MAG["mixed"] * og + 1.5 * 2.0 ** -23 * w * oc.astype(float)
            assert np.all(np.abs(gs - os_) <= bound), label
        # the same call with a catch-all last edge cannot be culled
        gpu_hist(ctx, x, u, np.append(edges, np.inf), op, vedges=ve, prec=prec)
        assert items_cull < ctx.stats()["items"], "tile culling did not drop any work item"


def test_cull_sphere_and_single_pass(ctx):
    rng = np.random.default_rng(5)
    N = 20000
    xd, _ = synthetic.sphere_points(rng, N, n_tracks=16, per_track=256)
    u = rng.normal(size=(2, N))
    edges = np.geomspace(1e-4, 0.05, 24)
    os_, oc, oa, og = _oracle.hist(xd, u, edges, op=(0, 2, 0), geometry="deg", nthreads=8, with_mag=True)
    gs, gc = gpu_hist(ctx, xd, u, edges, (0, 2, 0), geometry="deg", prec="mixed")
    check(gs, gc, os_, oc, oa, TOL["mixed"], "sphere L2", og=og, mag=MAG["mixed"])
    from test_gpu_single_pass import gpu_sp, row_scales, check_sp
    x2, u2 = flat_case(20000, 2, 9)
    e2 = np.geomspace(1e-3, 0.05, 16)
    ss, sc = _oracle.single_pass(x2, u2, e2, nthreads=8)
    gs2, gc2 = gpu_sp(ctx, x2, u2, e2, "flat", "mixed")
    oa2, og2 = row_scales(x2, u2, e2, "flat")
    check_sp(gs2, gc2, ss, sc, oa2, og2, "mixed", "single pass culled")


def test_cull_rebin_queue_uses_caller_order(ctx):
    """Points on a lattice put many pairs exactly on edges (all re-binned in binary64); with the
    records reordered, the queue must still hand the binary64 path each pair as (i < j) in the
    caller's order, so even values that depend on evaluation order land in the oracle's bins."""
    g = np.arange(120) / 119.0
    X, Y = np.meshgrid(g, g)
    x = np.stack([X.ravel(), Y.ravel()])
    rng = np.random.default_rng(3)
    u = rng.normal(size=x.shape)
    edges = np.array([0.5, 1.0, 2.0, 3.0]) / 119.0
    for op, ve in [((0, 1, 0), np.linspace(-4, 4, 17)), ((0, 0, 1), np.linspace(-4, 4, 17))]:
        os_, oc, oa = _oracle.hist(x, u, edges, op=op, vedges=ve, nthreads=8)
        gs, gc = gpu_hist(ctx, x, u, edges, op, vedges=ve, prec="mixed")
        assert np.array_equal(gc, oc), f"op={op}"
        assert ctx.stats()["rebinned"] > 0
