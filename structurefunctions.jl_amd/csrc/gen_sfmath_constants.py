"""Constants of sfmath.h, computed with mpmath at 60 significant digits.

Run `python gen_sfmath_constants.py` to print them; tests/test_sfmath.py checks that every constant
in sfmath.h equals the value printed here, bit for bit.

  * SF_PI180_*, SF_PIO2_*, SF_PI_*: pi/180, pi/2, pi as a double-double (hi = nearest binary64, lo =
    nearest binary64 of the remainder);
  * SF_TWO_OVER_PI: nearest binary64 of 2/pi;
  * SF_PIO2_P1, P2, P3: Cody-Waite split of pi/2, P1 and P2 rounded to 33 significant bits (so
    k * P1 and k * P2 are exact for |k| < 2^20), P3 the nearest binary64 of the rest;
  * atan_hi[k], atan_lo[k]: atan(k/8), k = 0..8, as a double-double;
  * sin / cos / atan polynomial coefficients: the Taylor coefficients (-1)^n / (2n+1)!,
    (-1)^n / (2n)! and (-1)^n / (2n+1), rounded to nearest.
"""
import mpmath as mp

mp.mp.dps = 60


def d(x):
    """Nearest binary64 of an mpmath number."""
    return float(mp.mpf(x))


def dd(x):
    hi = d(x)
    return hi, d(mp.mpf(x) - mp.mpf(hi))


def round_bits(x, bits):
    """x rounded to `bits` significant bits (nearest)."""
    m, e = mp.frexp(mp.mpf(x))
    return float(mp.ldexp(mp.nint(mp.ldexp(m, bits)), e - bits))


def constants():
    c = {}
    c["SF_PI180_HI"], c["SF_PI180_LO"] = dd(mp.pi / 180)
    c["SF_PIO2_HI"], c["SF_PIO2_LO"] = dd(mp.pi / 2)
    c["SF_PI_HI"], c["SF_PI_LO"] = dd(mp.pi)
    c["SF_TWO_OVER_PI"] = d(2 / mp.pi)
    p1 = round_bits(mp.pi / 2, 33)
    p2 = round_bits(mp.pi / 2 - p1, 33)
    c["SF_PIO2_P1"], c["SF_PIO2_P2"] = p1, p2
    c["SF_PIO2_P3"] = d(mp.pi / 2 - p1 - p2)
    c["atan_hi"] = [dd(mp.atan(mp.mpf(k) / 8))[0] for k in range(9)]
    c["atan_lo"] = [dd(mp.atan(mp.mpf(k) / 8))[1] for k in range(9)]
    # Horner order as in sfmath.h: highest power first
    c["sin"] = [d((-1) ** n / mp.factorial(2 * n + 1)) for n in range(8, 0, -1)]
    c["cos"] = [d((-1) ** n / mp.factorial(2 * n)) for n in range(9, 1, -1)]
    c["atan"] = [d(mp.mpf((-1) ** n) / (2 * n + 1)) for n in range(7, 0, -1)]
    return c


if __name__ == "__main__":
    for k, v in constants().items():
        print(k, [repr(x) for x in v] if isinstance(v, list) else repr(v))
