"""GPU parity at benchmark scale: the library (through its C ABI) against the CPU oracle itself,
not against another GPU kernel.

Cases:
  * the flat 2D configuration at its full size (N = 65536, 2.1e9 pairs, 64 log distance bins),
    du_L^2 and du_T^3, binary64 and binary32 kernels;
  * the spherical configuration's field at N = 32768 (half uniform, half on great-circle tracks
    at 1e-3 rad spacing with 1e-4 rad jitter, where many pairs lie near distance edges): the joint
    (distance, du_L) histogram in all three precisions and du_L^2 in 1D;
  * the flat 3D configuration at N = 32768: joint (distance, du_L) 48 x 128, all three precisions.

Bar (test/test_gpu_parity.jl:37-38): counts bit-identical; sums within 1e-12 of sum |value| per
bin (binary64 kernel); the binary32 kernels add their stated per-pair bounds (the rounding of the
pair arithmetic, 24 u32 ||du||^order, and for fixed-point cells 2^-23 value-bin widths).
"""
import os

import numpy as np
import pytest

import _oracle
from sfamd import synthetic
from test_gpu_parity import MAG, TOL, gpu_hist

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def _check(label, gs, gc, os_, oc, oa, og, prec, w=None):
    assert np.array_equal(gc, oc), f"{label}: counts differ at {np.argwhere(gc != oc)[:5].tolist()}"
    ok = ~np.isnan(os_)
    assert np.array_equal(np.isnan(gs), ~ok), f"{label}: NaN pattern differs"
    err = np.abs(gs - os_)[ok]
    base = "fp64" if prec == "fp64" else "mixed"
    bound = TOL[base] * np.maximum(oa, 1e-300) + MAG[base] * og
    if prec == "fixed":
        # fixed-point cells hold only values of at least a quarter bin width (smaller ones go to
        # binary64 cells), so their rounding of 0.75 2^-22 bin widths is at most 3 2^-22 |v|
        bound = bound + 3.0 * 2.0 ** -22 * np.maximum(oa, 1e-300)
    bound = bound[ok]
    assert np.all(err <= bound), f"{label}: max excess over the bound {np.max(err - bound):.3g}"
    return float(np.max(err / np.maximum(oa[ok], 1e-300))) if err.size else 0.0


@pytest.fixture(scope="module")
def flat2_full():
    return synthetic.config2()


@pytest.mark.parametrize("op", [(0, 2, 0), (0, 0, 3)])
def test_flat2_full_size_against_oracle(ctx, flat2_full, op):
    x, u, de = flat2_full
    os_, oc, oa, og = _oracle.hist(x, u, de, op=op, nthreads=THREADS, with_mag=True)
    assert int(oc.sum()) > 0
    for prec in ("fp64", "mixed"):
        gs, gc = gpu_hist(ctx, x, u, de, op, prec=prec)
        _check(f"flat2 N=65536 op={op} {prec}", gs, gc, os_, oc, oa, og, prec)


@pytest.fixture(scope="module")
def sphere_32k():
    return synthetic.config4(N=32768)


def test_sphere_tracks_joint_against_oracle(ctx, sphere_32k):
    x, u, de, ve = sphere_32k
    os_, oc, oa, og = _oracle.hist(x, u, de, op=(0, 1, 0), geometry="deg", vedges=ve, nthreads=THREADS,
                                   with_mag=True)
    assert int(oc.sum()) > 0.5 * (x.shape[1] * (x.shape[1] - 1) // 2)
    for prec in ("fp64", "mixed", "fixed"):
        gs, gc = gpu_hist(ctx, x, u, de, (0, 1, 0), geometry="deg", vedges=ve, prec=prec)
        _check(f"sphere N=32768 joint du_L {prec}", gs, gc, os_, oc, oa, og, prec, w=ve[1] - ve[0])


def test_sphere_tracks_1d_against_oracle(ctx, sphere_32k):
    x, u, de, _ = sphere_32k
    os_, oc, oa, og = _oracle.hist(x, u, de, op=(0, 2, 0), geometry="deg", nthreads=THREADS, with_mag=True)
    for prec in ("fp64", "mixed"):
        gs, gc = gpu_hist(ctx, x, u, de, (0, 2, 0), geometry="deg", prec=prec)
        _check(f"sphere N=32768 du_L^2 {prec}", gs, gc, os_, oc, oa, og, prec)


def test_flat3_joint_against_oracle(ctx):
    x, u, de, ve = synthetic.config3(N=32768)
    os_, oc, oa, og = _oracle.hist(x, u, de, op=(0, 1, 0), vedges=ve, nthreads=THREADS, with_mag=True)
    for prec in ("fp64", "mixed", "fixed"):
        gs, gc = gpu_hist(ctx, x, u, de, (0, 1, 0), vedges=ve, prec=prec)
        _check(f"flat3 N=32768 joint du_L {prec}", gs, gc, os_, oc, oa, og, prec, w=ve[1] - ve[0])
