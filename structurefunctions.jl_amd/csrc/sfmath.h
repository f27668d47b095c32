// sfmath.h -- deterministic FP64 elementary functions shared by host and device.
//
// The spherical pair path bins pairs by the central angle
//     sigma = 2 * atan(sqrt(d.d), sqrt(s.s))
// (src/HelperFunctions.jl:196) of unit positions built from sincosd(lon), sincosd(lat)
// (src/HelperFunctions.jl:133-153; sincos for radians, :406-409).  For integer counts to be
// bit-identical between the CPU oracle and the gfx950 kernels, both sides must evaluate these
// functions with the same sequence of IEEE-754 binary64 operations.  Library sin/cos/atan2 differ
// between glibc and the ROCm device library, so this header carries one implementation of each,
// written with only +, -, *, /, sqrt, fma, rint, floor, fmod and sign manipulation -- operations
// that are correctly rounded (or exact) on both x86-64 and gfx950.
//
// Methods (textbook, coefficients are exact Taylor coefficients rounded to binary64):
//   * sin/cos on |x| <= pi/4: Taylor polynomials to x^17 / x^18 with a first-order tail
//     correction for a double-double argument; cos uses the compensated 1 - x^2/2 split.
//   * sincosd: exact reduction in degrees (fmod by 360, then by quadrants of 90), conversion of
//     the reduced angle to radians as a double-double, quadrant selection.  Zero results follow
//     Julia's sind/cosd sign conventions (cosd never returns -0.0; sind(-x) == -sind(x)).
//   * sincos (radians): Cody-Waite reduction by pi/2 with a three-part constant.
//   * atan2: reduction to t in [0, 1], table atan(k/8) in double-double plus an odd Taylor
//     polynomial in z = (t - k/8) / (1 + t k/8), |z| <= 1/16.
// Accuracy is about 1 ulp.  The functions are NOT bit-identical to Julia's Base (openlibm-derived)
// implementations; they are identical between this repository's host and device builds, which is
// what bit-exact counts require.  Constants were computed with mpmath at 60 digits
// (gen_sfmath_constants.py in this directory).
//
// Every translation unit that includes this header must be compiled with FMA contraction off
// (gcc/hipcc: -ffp-contract=off); the functions below also request it for clang.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define SF_HD __host__ __device__ inline
#else
#define SF_HD static inline
#endif

#if defined(__HIPCC__) || defined(__clang__)
#define SF_NO_CONTRACT _Pragma("clang fp contract(off)")
#else
#define SF_NO_CONTRACT
#endif

#if defined(__HIP_DEVICE_COMPILE__)
#define SF_FMA(a, b, c) __builtin_fma((a), (b), (c))
#define SF_SQRT(a) __builtin_sqrt(a)
#define SF_RINT(a) __builtin_rint(a)
#define SF_FLOOR(a) __builtin_floor(a)
#define SF_FMOD(a, b) __builtin_fmod((a), (b))
#define SF_FABS(a) __builtin_fabs(a)
#define SF_COPYSIGN(a, b) __builtin_copysign((a), (b))
#else
#include <math.h>
#define SF_FMA(a, b, c) fma((a), (b), (c))
#define SF_SQRT(a) sqrt(a)
#define SF_RINT(a) rint(a)
#define SF_FLOOR(a) floor(a)
#define SF_FMOD(a, b) fmod((a), (b))
#define SF_FABS(a) fabs(a)
#define SF_COPYSIGN(a, b) copysign((a), (b))
#endif

// ---------------------------------------------------------------------------------------------
// constants
// ---------------------------------------------------------------------------------------------
#define SF_PI180_HI 0.017453292519943295
#define SF_PI180_LO 2.9486522708701687e-19
#define SF_PIO2_HI 1.5707963267948966
#define SF_PIO2_LO 6.123233995736766e-17
#define SF_PI_HI 3.141592653589793
#define SF_PI_LO 1.2246467991473532e-16
#define SF_TWO_OVER_PI 0.6366197723675814
// pi/2 = P1 + P2 + P3, P1 and P2 with 33 significant bits (k*P1, k*P2 exact for |k| < 2^20)
#define SF_PIO2_P1 1.5707963267341256
#define SF_PIO2_P2 6.077100506303966e-11
#define SF_PIO2_P3 2.0222662487959506e-21

SF_HD bool sf_isnan(double x) { return x != x; }
SF_HD bool sf_signbit(double x) { return SF_COPYSIGN(1.0, x) < 0.0; }
SF_HD bool sf_isinf(double x) { return SF_FABS(x) == __builtin_inf(); }

// error-free transformation a + b = s + e (Knuth two-sum)
SF_HD void sf_two_sum(double a, double b, double *s, double *e) {
  SF_NO_CONTRACT
  const double t = a + b;
  const double bb = t - a;
  *e = (a - (t - bb)) + (b - bb);
  *s = t;
}

// sin(x + y) for |x| <= pi/4 + small, |y| <= ulp(x)
SF_HD double sf_sin_kernel(double x, double y) {
  SF_NO_CONTRACT
  const double z = x * x;
  double s = 2.8114572543455206e-15;
  s = s * z + -7.647163731819816e-13;
  s = s * z + 1.6059043836821613e-10;
  s = s * z + -2.505210838544172e-08;
  s = s * z + 2.7557319223985893e-06;
  s = s * z + -0.0001984126984126984;
  s = s * z + 0.008333333333333333;
  s = s * z + -0.16666666666666666;
  // x + (x z S(z) + y (1 - z/2))
  const double corr = x * z * s + y * (1.0 - 0.5 * z);
  return x + corr;
}

// cos(x + y) for |x| <= pi/4 + small, |y| <= ulp(x)
SF_HD double sf_cos_kernel(double x, double y) {
  SF_NO_CONTRACT
  const double z = x * x;
  double c = -1.5619206968586225e-16;
  c = c * z + 4.779477332387385e-14;
  c = c * z + -1.1470745597729725e-11;
  c = c * z + 2.08767569878681e-09;
  c = c * z + -2.755731922398589e-07;
  c = c * z + 2.48015873015873e-05;
  c = c * z + -0.001388888888888889;
  c = c * z + 0.041666666666666664;
  const double hz = 0.5 * z;
  const double w = 1.0 - hz;                // exact rounding error recovered below (Sterbenz)
  const double err = (1.0 - w) - hz;
  const double r = z * z * c - x * y;
  return w + (err + r);
}

// sin and cos of an angle already reduced to [-45, 45] degrees (exactly), quadrant q in 0..3
SF_HD void sf_sincos_reduced_deg(double rr, int q, double *s, double *c) {
  SF_NO_CONTRACT
  double sn, cs;
  if (rr == 0.0) {
    sn = rr;
    cs = 1.0;
  } else {
    const double hi = rr * SF_PI180_HI;
    const double lo = SF_FMA(rr, SF_PI180_HI, -hi) + rr * SF_PI180_LO;
    const double th = hi + lo;
    const double tl = lo - (th - hi);
    sn = sf_sin_kernel(th, tl);
    cs = sf_cos_kernel(th, tl);
  }
  switch (q & 3) {
    case 0: *s = sn; *c = cs; break;
    case 1: *s = cs; *c = -sn; break;
    case 2: *s = -sn; *c = -cs; break;
    default: *s = -cs; *c = sn; break;
  }
}

// sind and cosd of a non-negative finite angle in degrees
SF_HD void sf_sincosd_nonneg(double a, double *s, double *c) {
  SF_NO_CONTRACT
  const double r = SF_FMOD(a, 360.0);      // exact, in [0, 360)
  const double n = SF_RINT(r / 90.0);      // 0..4
  const double rr = r - 90.0 * n;          // exact, |rr| <= 45
  sf_sincos_reduced_deg(rr, (int)n, s, c);
  // exact zeros are +0.0 for a >= 0
  *s = *s + 0.0;
  *c = *c + 0.0;
}

// (sind(x), cosd(x)) with Julia's sign conventions for exact zeros.  NaN/Inf give NaN.
SF_HD void sf_sincosd(double x, double *s, double *c) {
  SF_NO_CONTRACT
  if (sf_isnan(x) || sf_isinf(x)) {
    *s = __builtin_nan("");
    *c = __builtin_nan("");
    return;
  }
  double sp, cp;
  sf_sincosd_nonneg(SF_FABS(x), &sp, &cp);
  *c = cp;                                  // cosd is even
  if (x == 0.0) {
    *s = x;                                 // sind(+-0) = +-0
  } else {
    *s = sf_signbit(x) ? -sp : sp;          // sind is odd
  }
}

// (sin(x), cos(x)) in radians; Cody-Waite reduction valid for |x| < 2^20 * pi/2.
SF_HD void sf_sincos(double x, double *s, double *c) {
  SF_NO_CONTRACT
  if (sf_isnan(x) || sf_isinf(x)) {
    *s = __builtin_nan("");
    *c = __builtin_nan("");
    return;
  }
  if (x == 0.0) {
    *s = x;
    *c = 1.0;
    return;
  }
  const double k = SF_RINT(x * SF_TWO_OVER_PI);
  const double t1 = x - k * SF_PIO2_P1;        // k*P1 exact; subtraction exact (Sterbenz)
  const double w2 = k * SF_PIO2_P2;            // exact
  double t2, e2;
  sf_two_sum(t1, -w2, &t2, &e2);
  const double w3 = k * SF_PIO2_P3;
  double hi, lo;
  sf_two_sum(t2, e2 - w3, &hi, &lo);
  const double ki = k - 4.0 * SF_FLOOR(k * 0.25);  // k mod 4, exact
  const int q = (int)ki;
  const double sn = sf_sin_kernel(hi, lo);
  const double cs = sf_cos_kernel(hi, lo);
  switch (q & 3) {
    case 0: *s = sn; *c = cs; break;
    case 1: *s = cs; *c = -sn; break;
    case 2: *s = -sn; *c = -cs; break;
    default: *s = -cs; *c = sn; break;
  }
}

// atan(t) for t in [0, 1] as hi + lo (lo tiny); returns hi + lo rounded, and writes the parts.
SF_HD double sf_atan01(double t, double *lo_out) {
  SF_NO_CONTRACT
  const double atan_hi[9] = {0.0,
                             0.12435499454676144,
                             0.24497866312686414,
                             0.35877067027057225,
                             0.4636476090008061,
                             0.5585993153435624,
                             0.6435011087932844,
                             0.7188299996216245,
                             0.7853981633974483};
  const double atan_lo[9] = {0.0,
                             -3.1253241424539383e-18,
                             1.0698755618734451e-17,
                             -2.4623815582638635e-17,
                             2.2698777452961687e-17,
                             -5.4556305485916264e-18,
                             1.5834785051444286e-17,
                             -2.1478388444456983e-17,
                             3.061616997868383e-17};
  const int k = (int)SF_FLOOR(t * 8.0 + 0.5);  // 0..8
  const double ck = (double)k * 0.125;
  double z;
  if (k == 0) {
    z = t;
  } else {
    z = (t - ck) / (1.0 + t * ck);             // t - ck exact (Sterbenz)
  }
  const double z2 = z * z;
  double p = -0.06666666666666667;
  p = p * z2 + 0.07692307692307693;
  p = p * z2 + -0.09090909090909091;
  p = p * z2 + 0.1111111111111111;
  p = p * z2 + -0.14285714285714285;
  p = p * z2 + 0.2;
  p = p * z2 + -0.3333333333333333;
  const double tail = z * z2 * p;              // atan(z) = z + tail
  const double hi = atan_hi[k];
  const double s1 = hi + z;
  const double e1 = (k == 0) ? 0.0 : ((hi - s1) + z);  // |hi| >= |z| when k > 0
  const double lo = e1 + (atan_lo[k] + tail);
  const double r = s1 + lo;
  *lo_out = lo - (r - s1);
  return r;
}

// atan2(y, x) with IEEE/Julia special-value semantics.
SF_HD double sf_atan2(double y, double x) {
  SF_NO_CONTRACT
  if (sf_isnan(x) || sf_isnan(y)) return x + y;
  const bool yneg = sf_signbit(y);
  const bool xneg = sf_signbit(x);
  const double ay = SF_FABS(y), ax = SF_FABS(x);
  double r;
  if (ay == 0.0) {
    r = xneg ? SF_PI_HI : 0.0;
  } else if (ax == 0.0) {
    r = SF_PIO2_HI;
  } else if (sf_isinf(ax) && sf_isinf(ay)) {
    r = xneg ? 3.0 * (0.5 * SF_PIO2_HI) : 0.5 * SF_PIO2_HI;
  } else if (sf_isinf(ax)) {
    r = xneg ? SF_PI_HI : 0.0;
  } else if (sf_isinf(ay)) {
    r = SF_PIO2_HI;
  } else {
    double lo;
    double hi;
    if (ay <= ax) {
      hi = sf_atan01(ay / ax, &lo);
    } else {
      // pi/2 - atan(ax/ay)
      const double a = sf_atan01(ax / ay, &lo);
      const double h = SF_PIO2_HI - a;                 // |pi/2| >= |a|: fast two-sum
      const double e = (SF_PIO2_HI - h) - a;
      lo = (SF_PIO2_LO - lo) + e;
      hi = h;
    }
    if (xneg) {
      // pi - angle
      const double h = SF_PI_HI - hi;
      const double e = (SF_PI_HI - h) - hi;
      lo = (SF_PI_LO - lo) + e;
      hi = h;
    }
    r = hi + lo;
  }
  return yneg ? -r : r;
}
