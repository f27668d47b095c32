"""The deterministic binary64 functions of csrc/sfmath.h (run on the host through the oracle
library, which includes the same header) against mpmath at 50 digits: error in ulps of the
correctly rounded result, over multiples of 90 degrees, signed zeros, tiny and large arguments and
random sweeps.  Also checks that every constant in the header is the value its generator
(csrc/gen_sfmath_constants.py) computes."""
import importlib.util
import math
import os
import re

import mpmath as mp
import numpy as np
import pytest

import _oracle

mp.mp.dps = 50
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "structurefunctions.jl_amd", "csrc")


def ulp_err(got, exact):
    """|got - exact| in units of the binary64 ulp at the exact value."""
    exact = mp.mpf(exact)
    if exact == 0:
        return 0.0 if got == 0.0 else math.inf
    ref = float(exact)
    u = math.ulp(abs(ref)) if ref != 0.0 else math.ulp(0.0)
    return float(abs(mp.mpf(got) - exact) / u)


def test_constants_match_generator():
    spec = importlib.util.spec_from_file_location("gen", os.path.join(CSRC, "gen_sfmath_constants.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    c = gen.constants()
    text = open(os.path.join(CSRC, "sfmath.h")).read()
    for name in ("SF_PI180_HI", "SF_PI180_LO", "SF_PIO2_HI", "SF_PIO2_LO", "SF_PI_HI", "SF_PI_LO",
                 "SF_TWO_OVER_PI", "SF_PIO2_P1", "SF_PIO2_P2", "SF_PIO2_P3"):
        m = re.search(r"#define %s ([-0-9.e+]+)" % name, text)
        assert m and float(m.group(1)) == c[name], name
    for arr in ("atan_hi", "atan_lo"):
        body = re.search(r"const double %s\[9\] = \{([^}]*)\}" % arr, text).group(1)
        vals = [float(v) for v in body.replace("\n", " ").split(",")]
        assert vals == c[arr], arr
    def horner(fn, var):
        body = text[text.index("SF_HD double %s(" % fn):]
        body = body[:body.index("\n}\n")]
        first = float(re.search(r"double %s = ([-0-9.e+]+);" % var, body).group(1))
        rest = [float(v) for v in re.findall(r"%s = %s \* z2? \+ ([-0-9.e+]+);" % (var, var), body)]
        return [first] + rest
    assert horner("sf_sin_kernel", "s") == c["sin"]
    assert horner("sf_cos_kernel", "c") == c["cos"]
    assert horner("sf_atan01", "p") == c["atan"]


def _sincosd_args():
    rng = np.random.default_rng(1)
    a = [0.0, -0.0, 1e-300, -1e-300, 1e-20, 5e-324, 30.0, 45.0, 60.0, 89.99999999999999]
    a += [90.0 * k for k in range(-12, 13)]
    a += list(rng.uniform(-720, 720, 4000)) + list(rng.uniform(-1e5, 1e5, 1000))
    a += list(np.nextafter(np.array([90.0 * k for k in range(1, 9)]), 0)) + [1e5, -1e5, 123456.789]
    return a


def test_sincosd_vs_mpmath():
    worst = 0.0
    for x in _sincosd_args():
        s, c = _oracle.sincosd(x)
        xr = mp.mpf(x) * mp.pi / 180
        es, ec = mp.sin(xr), mp.cos(xr)
        if x % 90.0 == 0.0:
            # multiples of 90 degrees are exact: +-1 and zeros with Julia's signs
            k = int(round(x / 90.0)) % 4
            assert (s, c) == ((0.0, 1.0), (1.0, 0.0), (0.0, -1.0), (-1.0, 0.0))[k], x
            assert math.copysign(1.0, c) > 0 or c != 0.0  # cosd never returns -0.0
            if s == 0.0:
                assert math.copysign(1.0, s) == math.copysign(1.0, x), x  # sind(+-x) = +-sind(x)
            continue
        worst = max(worst, ulp_err(s, es), ulp_err(c, ec))
    assert worst < 1.0, worst


def test_sincos_radians_vs_mpmath():
    rng = np.random.default_rng(2)
    args = [0.0, -0.0, 1e-300, 1e-20, -1e-8, math.pi / 4, math.pi / 2, math.pi, 2 * math.pi]
    args += list(rng.uniform(-10, 10, 4000)) + list(rng.uniform(-1e5, 1e5, 1000)) + [1e5, -1.6e6]
    worst = 0.0
    for x in args:
        s, c = _oracle.sincos(x)
        if x == 0.0:
            assert s == x and math.copysign(1.0, s) == math.copysign(1.0, x) and c == 1.0
            continue
        worst = max(worst, ulp_err(s, mp.sin(mp.mpf(x))), ulp_err(c, mp.cos(mp.mpf(x))))
    assert worst < 1.0, worst


def test_atan2_vs_mpmath():
    rng = np.random.default_rng(3)
    n = 5000
    y = np.abs(rng.normal(size=n)) * 10.0 ** rng.uniform(-300, 300, n)
    x = np.abs(rng.normal(size=n)) * 10.0 ** rng.uniform(-300, 300, n)
    y2 = rng.uniform(0, 2, n)
    x2 = rng.uniform(0, 2, n)
    ys = np.concatenate([y, y2, -y[:500], y[:500]])
    xs = np.concatenate([x, x2, x[:500], -x[:500]])
    worst = 0.0
    for a, b in zip(ys, xs):
        worst = max(worst, ulp_err(_oracle.atan2(float(a), float(b)), mp.atan2(mp.mpf(a), mp.mpf(b))))
    # the central angle's arguments: |d| <= 2 and |p + q| <= 2, ratios near 0, 1 and large
    for t in [1e-300, 1e-10, 0.124, 0.125, 0.126, 0.5, 1.0 - 2 ** -52, 1.0, 1.0 + 2 ** -52, 8.0, 1e10]:
        worst = max(worst, ulp_err(_oracle.atan2(t, 1.0), mp.atan2(t, 1)))
    assert worst <= 1.25, worst


@pytest.mark.parametrize("y,x,want", [
    (0.0, 1.0, 0.0), (-0.0, 1.0, -0.0), (0.0, -1.0, math.pi), (-0.0, -1.0, -math.pi),
    (1.0, 0.0, math.pi / 2), (-1.0, 0.0, -math.pi / 2), (0.0, 0.0, 0.0), (0.0, -0.0, math.pi),
    (math.inf, math.inf, math.pi / 4), (math.inf, -math.inf, 3 * math.pi / 4),
    (1.0, math.inf, 0.0), (1.0, -math.inf, math.pi), (math.inf, 1.0, math.pi / 2)])
def test_atan2_special_values(y, x, want):
    """IEEE atan2 special cases (Julia's atan(y, x) follows them)."""
    got = _oracle.atan2(y, x)
    assert got == want and math.copysign(1.0, got) == math.copysign(1.0, want)
    assert math.isnan(_oracle.atan2(math.nan, 1.0)) and math.isnan(_oracle.atan2(1.0, math.nan))
