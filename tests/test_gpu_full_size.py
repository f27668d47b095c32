"""The benchmark configurations at their full sizes, through size-independent properties (the CPU
oracle would take minutes at these sizes; it pins the binary64 kernel at smaller ones, and the
binary64 kernel then serves as the reference here):

* cfg 1 (flat 2D, N = 4096, 8.4e6 pairs) is small enough for the oracle itself;
* cfg 2 (flat 2D, N = 65536), all six projected operators: binary32 counts equal the binary64
  kernel's bit for bit; sums within 1e-6 of sum ||du||^order per bin (the magnitude sums come
  from the binary64 kernel's |du|, |du|^2, |du|^3 operators);
* cfg 3 (flat 3D, N = 131072, joint du_L 48 x 128, UInt64 counts): counts bit-identical in both
  binary32 precisions; sums within 1e-6 of count x max |edge| of the cell's value bin (plus the
  fixed-point cells' 2^-23 bin widths per pair);
* cfg 5 (256 ragged slices, 3.0e10 pairs per operator): per slice, counts of the batched
  binary32 call equal the batched binary64 call's.
The sphere configuration (cfg 4, N = 262144) has its own full-size test in test_gpu_api."""
import numpy as np
import pytest

import _oracle
import sfamd
from sfamd import synthetic
from test_gpu_parity import MAG, TOL, check, gpu_hist

pytestmark = pytest.mark.gpu

OPS6 = {"L1": (0, 1, 0), "T1": (0, 0, 1), "L2": (0, 2, 0), "T2": (0, 0, 2), "L3": (0, 3, 0), "T3": (0, 0, 3)}


def test_cfg1_full_size_against_oracle(ctx):
    x, u, e = synthetic.config1()
    os_, oc, oa, og = _oracle.hist(x, u, e, op=(0, 2, 0), nthreads=8, with_mag=True)
    for prec in ("fp64", "mixed"):
        gs, gc = gpu_hist(ctx, x, u, e, (0, 2, 0), prec=prec)
        check(gs, gc, os_, oc, oa, TOL[prec], f"cfg1 {prec}", og=og, mag=MAG[prec])


@pytest.fixture(scope="module")
def cfg2():
    return synthetic.config2()


def test_cfg2_full_size_six_operators(ctx, cfg2):
    x, u, e = cfg2
    # sum ||du||^k per bin, k = 1, 2, 3 (binary64 kernel): the scale of each order's sums
    mag = {1: gpu_hist(ctx, x, u, e, (5, 1, 0), prec="fp64")[0],
           2: gpu_hist(ctx, x, u, e, (1, 0, 0), prec="fp64")[0],
           3: gpu_hist(ctx, x, u, e, (5, 3, 0), prec="fp64")[0]}
    for name, op in OPS6.items():
        order = max(op[1], op[2])
        s64, c64 = gpu_hist(ctx, x, u, e, op, prec="fp64")
        s32, c32 = gpu_hist(ctx, x, u, e, op, prec="mixed")
        assert c64.sum() > 0
        assert np.array_equal(c32, c64), f"cfg2 {name}: counts differ at {np.argwhere(c32 != c64)[:5].tolist()}"
        err = np.abs(s32 - s64)
        bound = 1e-6 * mag[order] + 1e-300
        assert np.all(err <= bound), f"cfg2 {name}: max err / bound {np.max(err / bound):.3g}"


def test_cfg3_full_size_joint(ctx):
    x, u, e, ve = synthetic.config3()
    s64, c64 = gpu_hist(ctx, x, u, e, (0, 1, 0), vedges=ve, prec="fp64")
    assert c64.sum() > 0
    vmax = np.maximum(np.abs(ve[:-1]), np.abs(ve[1:]))[None, :]
    w = ve[1] - ve[0]
    for prec in ("mixed", "fixed"):
        s32, c32 = gpu_hist(ctx, x, u, e, (0, 1, 0), vedges=ve, prec=prec)
        assert np.array_equal(c32, c64), f"cfg3 {prec}: counts differ"
        bound = 1e-6 * c64 * vmax + 1e-300
        if prec == "fixed":
            bound = bound + 1.5 * 2.0 ** -23 * w * c64
        assert np.all(np.abs(s32 - s64) <= bound), f"cfg3 {prec}"


def test_cfg5_full_size_batch_counts():
    xs, us, e, ves = synthetic.config5()
    assert len(xs) == 256
    for sf, vb in ((sfamd.L1SF, ves), (sfamd.L2SF, None)):
        r64 = sfamd.calculate_structure_function_batch(sf, xs, us, e, vb,
                                                       backend=sfamd.MI355XBackend(precision="fp64"))
        r32 = sfamd.calculate_structure_function_batch(sf, xs, us, e, vb,
                                                       backend=sfamd.MI355XBackend(precision="mixed_fixed"))
        for k, (a, b) in enumerate(zip(r64, r32)):
            assert np.array_equal(a.counts, b.counts), f"cfg5 {type(sf).__name__} slice {k}"
