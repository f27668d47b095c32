// sf_kernels.hip -- gfx950 kernels for the O(N^2) pair histogram.
//
// Kernels
//   prep64<G,D>      per-point binary64 records (sphere: sincosd once per point, not per pair)
//   prep32<G,D>      per-point binary32 records for the mixed kernel
//   hist64<G,D,J>    persistent upper-triangle tile kernel, binary64 pair arithmetic written in
//                    the reference's operation order: counts bit-identical to SerialBackend
//                    (bins certified from the exact key and value, the reference's rules where a
//                    key or value is within the margin of a decision point)
//   hist32<G,D,J>    same schedule, binary32 pair arithmetic; a pair whose distance key or value
//                    lies within the binary32 error bound of a bin edge (or is degenerate) is
//                    appended to a queue instead of being binned
//   recheck<G,D,J>   bins every queued pair with the binary64 arithmetic of hist64
//   merge            adds the internal accumulators into the caller's outputs
//
// Schedule (both hist kernels): a work item is (slice, i-tile of BLOCK rows, column range
// [j0, j1)).  Each thread owns one row i and keeps its record in registers; the column records
// are read at wave-uniform addresses, by scalar loads into SGPRs or (some binary32 variants)
// staged BLOCK at a time in LDS, so every wave reads them as broadcasts.  Pairs are (i, j) with
// j > i.  Workgroups are persistent and pop items from a global counter; each keeps
// the current slice's histogram in LDS (binary64 sums, 32-bit counts) and flushes it to the
// accumulators with global atomics when it changes slice, before its 32-bit counts could wrap,
// and when it runs out of items.
//
// All files of this library are compiled with -ffp-contract=off: the binary64 path must perform
// exactly the operations of the reference (src/HelperFunctions.jl, src/StructureFunctionTypes.jl,
// src/BinEdges.jl) with no fused multiply-adds.  The binary32 path asks for fma explicitly.
#include <hip/hip_runtime.h>

#include "sf_internal.h"
#include "sfmath.h"

namespace sfk {

#define DEV __device__ __forceinline__

// =============================================================================================
// binary64 exact per-pair arithmetic (mirrors the oracle, which mirrors the reference)
// =============================================================================================

// StaticArrays dot: left fold from the first product
template <int D>
DEV double sdot(const double *a, const double *b) {
  double s = a[0] * b[0];
#pragma unroll
  for (int k = 1; k < D; ++k) s = s + a[k] * b[k];
  return s;
}

// _sum_abs2 / norm2: 0 + a1^2 + ...
template <int D>
DEV double sum_abs2(const double *a) {
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < D; ++k) s += a[k] * a[k];
  return s;
}

DEV double ipow(double x, int n) {
  if (n == 1) return x;
  if (n == 2) return x * x;
  return x * x * x;
}

// mdu_t = dot(du, nhat(rhat))   (HelperFunctions.jl:450-471, 640-643)
template <int D>
DEV double mdu_t64(const double *du, const double *rh) {
  if (D == 2) {
    const double n[2] = {-rh[1], rh[0]};
    return sdot<2>(du, n);
  } else {
    const double c[3] = {0.0 * rh[2] - 1.0 * rh[1], 1.0 * rh[0] - 0.0 * rh[2],
                         0.0 * rh[1] - 0.0 * rh[0]};
    const double nrm = __builtin_sqrt(sdot<3>(c, c));
    const double n[3] = {c[0] / nrm, c[1] / nrm, c[2] / nrm};
    return sdot<3>(du, n);
  }
}

// operator functor sf(du, rhat)   (StructureFunctionTypes.jl:58-151)
template <int D>
DEV double op_apply64(int op, int nl, int nt, const double *du, const double *rh) {
  switch (op) {
    case 0: {  // ProjectedStructureFunctionType{nl, nt}
      double ex = 1.0;
      if (nl != 0) ex = ex * ipow(sdot<D>(du, rh), nl);
      if (nt != 0) {
        if (nt == 2) {
          const double L = sdot<D>(du, rh);
          ex = ex * (sum_abs2<D>(du) - L * L);
        } else {
          ex = ex * ipow(mdu_t64<D>(du, rh), nt);
        }
      }
      return ex;
    }
    case 1:  // SecondOrder
      return sum_abs2<D>(du);
    case 2:  // ThirdOrder: du_L * ||du||^2
      return sdot<D>(du, rh) * sum_abs2<D>(du);
    case 3: {  // T2Component
      const double L = sdot<D>(du, rh);
      return (sum_abs2<D>(du) - L * L) / (double)(D - 1);
    }
    case 4: {  // L1T2Component
      const double L = sdot<D>(du, rh);
      return L * ((sum_abs2<D>(du) - L * L) / (double)(D - 1));
    }
    default: {  // FullVector{nl}
      if (nl == 2) return sum_abs2<D>(du);
      return ipow(__builtin_sqrt(sum_abs2<D>(du)), nl);
    }
  }
}

// _sf_raw(sf, du, dx, r2) on the flat path   (StructureFunctionTypes.jl:227-244)
template <int D>
DEV double sf_raw64(int op, int nl, int nt, const double *du, const double *dx, double r2) {
  if (op == 1) return sum_abs2<D>(du);
  if (op == 0 && nl == 2 && nt == 0) {
    const double p = sdot<D>(du, dx);
    return p * p / r2;
  }
  if (op == 0 && nl == 0 && nt == 2) {
    const double p = sdot<D>(du, dx);
    return sum_abs2<D>(du) - p * p / r2;
  }
  if (op == 3) {
    const double p = sdot<D>(du, dx);
    return (sum_abs2<D>(du) - p * p / r2) / (double)(D - 1);
  }
  const double s = __builtin_sqrt(r2);
  double rh[D];
#pragma unroll
  for (int k = 0; k < D; ++k) rh[k] = dx[k] / s;
  return op_apply64<D>(op, nl, nt, du, rh);
}

// Spherical pair frame (HelperFunctions.jl:186-200, 265-276) from binary64 records
// (P0 P1 P2 E0 E1 N0 N1 N2 u...): distance, ok, and the pieces pair_delta needs.
struct SphereFrame {
  double d[3];
  double dp, s2, distance;
  bool ok;
};

// Everything of the frame but the central angle: d = q - p, p.d, sin^2 = |p x d|^2, ok.
DEV SphereFrame sphere_frame64_noangle(const double *p, const double *q) {
  SphereFrame f;
  f.d[0] = q[0] - p[0]; f.d[1] = q[1] - p[1]; f.d[2] = q[2] - p[2];
  f.dp = sdot<3>(f.d, p);
  const double *d = f.d;
  const double w[3] = {p[1] * d[2] - p[2] * d[1], p[2] * d[0] - p[0] * d[2],
                       p[0] * d[1] - p[1] * d[0]};
  f.s2 = sdot<3>(w, w);
  f.ok = f.s2 > 2.2250738585072014e-308;  // floatmin(Float64)
  f.distance = 0.0;
  return f;
}

// distance = radius * 2 atan2(|d|, |p + q|), the reference's central angle
DEV double sphere_distance64(const double *p, const double *q, const double *d, double radius) {
  const double s[3] = {p[0] + q[0], p[1] + q[1], p[2] + q[2]};
  const double sigma = 2.0 * sf_atan2(__builtin_sqrt(sdot<3>(d, d)), __builtin_sqrt(sdot<3>(s, s)));
  return radius * sigma;
}

DEV SphereFrame sphere_frame64(const double *p, const double *q, double radius) {
  SphereFrame f = sphere_frame64_noangle(p, q);
  f.distance = sphere_distance64(p, q, f.d, radius);
  return f;
}

// pair_delta(::SphericalGeometry{D}) (HelperFunctions.jl:324-348) given the frame
template <int D>
DEV void sphere_delta64(const SphereFrame &f, const double *A, const double *B, double *du) {
  const double *p = A, *q = B, *d = f.d;
  const double dp = f.dp;
  const double t_A[3] = {d[0] - dp * p[0], d[1] - dp * p[1], d[2] - dp * p[2]};
  const double t_B[3] = {dp * q[0] + d[0], dp * q[1] + d[1], dp * q[2] + d[2]};
  const double inv_s = f.ok ? 1.0 / __builtin_sqrt(f.s2) : 0.0;
  const double ta[3] = {t_A[0] * inv_s, t_A[1] * inv_s, t_A[2] * inv_s};
  const double tb[3] = {t_B[0] * inv_s, t_B[1] * inv_s, t_B[2] * inv_s};
  const double EA[3] = {A[3], A[4], 0.0}, NA[3] = {A[5], A[6], A[7]};
  const double EB[3] = {B[3], B[4], 0.0}, NB[3] = {B[5], B[6], B[7]};
  const double tA[2] = {sdot<3>(ta, EA), sdot<3>(ta, NA)};
  const double tB[2] = {sdot<3>(tb, EB), sdot<3>(tb, NB)};
  const double nA[2] = {-tA[1], tA[0]};
  const double nB[2] = {-tB[1], tB[0]};
  const double *uA = A + 8, *uB = B + 8;
  du[0] = sdot<2>(uB, tB) - sdot<2>(uA, tA);
  du[1] = sdot<2>(uB, nB) - sdot<2>(uA, nA);
  if (D == 3) du[2] = uB[2] - uA[2];
}

// ---------------------------------------------------------------------------------------------
// exact bin lookup
// ---------------------------------------------------------------------------------------------
DEV int cell64(double key, const Slice &s) {
  const long long bits = __double_as_longlong(key);
  long long c = (bits >> (52 - s.M64)) - s.base64;
  c = c < 0 ? 0 : c;
  c = c >= s.ncell64 ? s.ncell64 - 1 : c;
  return (int)c;
}

// number of exact distance-key edges strictly below key (nE for NaN, which sorts last).  Keys <= 0
// (coincident points, r^2 = 0) are counted against the edges directly: a leading -Inf edge
// (InfPaddedBinEdges) lies below them.
DEV int dist_bin64(double key, const Cell *tab, const double *E, int nE, const Slice &s) {
  if (!(key > 0.0)) {
    if (key != key) return nE;
    // Julia isless: -0.0 sorts below +0.0 (a flat distance of +0.0 lies above a -0.0 edge)
    int b = 0;
    while (b < nE && (E[b] < key || (E[b] == key && __builtin_signbit(E[b]) && !__builtin_signbit(key)))) ++b;
    return b;
  }
  int b = tab[cell64(key, s)].bx;
  while (b < nE && E[b] < key) ++b;
  return b;
}

// digitize(x, ::LinearBinEdges) = searchsortedfirst(v, x) - 1 (BinEdges.jl:114-131), or inside
// InfPaddedBinEdges (:306-316): the histogram bin, valid 1 .. nb (0 and nb + 1 are out of range).
// muladd is fused.  A NaN query (the reference's floor(Int, NaN) throws) is out of range.
// Branch-free: every case is evaluated and the answer selected.  The reference's Int64 index
// clamp(floor(t) + 1, 1, n) is formed in binary64, which is exact wherever it is used (x in
// (f, l], so |t| <= n + 1 < 2^31; the host keeps n below 2^31); elsewhere NaN and infinities are
// clamped away (max / min return the number) before the conversion.
DEV int lin_rule_bin(const LinRule &r, double x, int nb) {
  const double c = -r.f * r.inv;  // a separate product, as in the reference
  const double t = __builtin_fma(x, r.inv, c);
  const double n = (double)r.n;
  const double fi = __builtin_fmin(__builtin_fmax(__builtin_floor(t) + 1.0, 1.0), n);
  const double u = __builtin_fma(fi - 1.0, r.step, r.f);
  int k = (int)fi + (u < x ? 1 : 0);
  k = (x <= r.f) ? 1 : k;
  k = (x > r.l) ? (int)r.n + 1 : k;
  k = r.padded ? k : k - 1;
  k = (r.padded && x <= -__builtin_inf()) ? 0 : k;
  return (x != x) ? nb + 1 : k;
}

// searchsortedfirst(value_edges, v) - 1 with Julia isless semantics
DEV int val_bin64(double v, const double *E, int nE, const Slice &s) {
  if (v != v) return nE;
  int g;
  if (s.v_uniform) {
    double t = (v - s.v0) * s.vinv;
    t = t < -1.0 ? -1.0 : t;
    t = t > (double)nE ? (double)nE : t;
    g = (int)__builtin_floor(t) + 1;
    g = g < 0 ? 0 : (g > nE ? nE : g);
    while (g > 0 && !(E[g - 1] < v)) --g;
    while (g < nE && E[g] < v) ++g;
  } else if (s.v_log) {
    // log-spaced edges from edge vk0 on: a guess from the binary32 log2 of v, then the walk,
    // which alone decides (the count of edges below v; v <= 0 starts from edge 0)
    float tg = __builtin_fmaf(__builtin_amdgcn_logf((float)v) - s.vl0f, s.vlinvf, (float)s.vk0 + 1.0f);
    tg = __builtin_fminf(__builtin_fmaxf(tg, 0.0f), (float)nE);
    g = (int)tg;
    while (g > 0 && !(E[g - 1] < v)) --g;
    while (g < nE && E[g] < v) ++g;
  } else {
    int lo = 0, hi = nE;
    while (lo < hi) {
      const int m = (lo + hi) >> 1;
      if (E[m] < v) lo = m + 1; else hi = m;
    }
    g = lo;
  }
  if (v == 0.0 && !sf_signbit(v)) g += s.v_negzero;
  return g;
}

// Full exact evaluation of pair (i, j) of slice s from binary64 records.  Returns the output cell
// (0-based within the slice) or -1 if the pair is not counted; writes the value.  OPK fixes the
// operator at compile time (1: {1,0}, 2: {2,0}) with the same operation sequence as the generic
// functor; 0 dispatches on P.op at run time; 3 is the single pass, which writes du_L to val[0] and
// |du|^2 to val[1] (flat: _pf_sp_simd_pairs!, serial_single_pass.jl:245-253, du_L =
// dot(du, dx) * inv(sqrt(r2)); sphere: _sp1d_pairs!, :333-338, du_L = dot(du, e1)); 4 is a plane
// of the single pass's 2D form: the same du_L and |du|^2, then invariant s.plane as value.
// Invariant t of the single pass's 2D form from du_L and |du|^2, in the reference's operation
// order (_sp2d_simd_pairs!, serial_single_pass.jl:906-911): |du|^2, du_L^2, |du|^2 - du_L^2,
// du_L |du|^2, du_L du_L^2, du_L (|du|^2 - du_L^2).  t is uniform over a slice.
template <typename T>
DEV T sp_invariant(int t, T duL, T dn2) {
  const T duL2 = duL * duL;
  const T duT2 = dn2 - duL2;
  switch (t) {
    case 0: return dn2;
    case 1: return duL2;
    case 2: return duT2;
    case 3: return duL * dn2;
    case 4: return duL * duL2;
    default: return duL * duT2;
  }
}

DEV int cell32(float key, const Slice &s);

// CERT (sphere): the distance bin is read from the chord^2 cell table of the binary32 kernel
// (lo, hi, blo: ambiguity interval and bin count per cell) with the binary64 chord^2 |q - p|^2,
// which is the reference's own |d|^2; the reference's central angle 2 atan2(|d|, |p + q|) is
// monotone in it, and the table's margins (2^-20 relative and more) exceed the error of both the
// binary32 rounding of the key and the reference's angle.  Only a pair whose chord^2 lies in a
// cell's ambiguity interval computes the angle and searches (tab is then the global binary64 key
// table).  Values are always the reference's operation sequence.
template <int G, int D, bool J, int OPK = 0, bool CERT = false>
DEV int exact_cell(const Params &P, const Slice &s, const double *A, const double *B,
                   const Cell *tab, const double *dE, const double *vE, double *val,
                   const float2 *lohi = nullptr, const int2 *blo_tab = nullptr, int *bins = nullptr) {
  const int nEd = P.n_dist + 1;
  int db;
  double v;
  if (G == kFlat) {
    double dx[D], du[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      dx[k] = B[k] - A[k];
      du[k] = B[D + k] - A[D + k];
    }
    const double r2 = sdot<D>(dx, dx);
    if (P.bl) {
      // batch-leading kernels (batch_leading.jl:105-330): pair_frame's r = sqrt(dx.dx) binned
      // against the edges themselves, rhat = dx / r (pair_direction), values from the operator
      // on (du, rhat), single-pass du_L = dot(du, rhat) (HelperFunctions.jl:248-330)
      const double r = __builtin_sqrt(r2);
      db = s.drule.on ? lin_rule_bin(s.drule, r, P.n_dist) : dist_bin64(r, tab, dE, nEd, s);
      if (db < 1 || db > P.n_dist) return -1;
      if (bins) bins[0] = db - 1;
      double rh[D];
#pragma unroll
      for (int k = 0; k < D; ++k) rh[k] = dx[k] / r;
      if constexpr (OPK == 3 || OPK == 4) {
        const double duL = sdot<D>(du, rh);
        const double dn2 = sdot<D>(du, du);
        if constexpr (OPK == 3) {
          val[0] = duL;
          val[1] = dn2;
          return db - 1;
        } else {
          v = sp_invariant(s.plane, duL, dn2);
        }
      } else if constexpr (OPK == 1) {
        v = 1.0 * sdot<D>(du, rh);
      } else if constexpr (OPK == 2) {
        const double L = sdot<D>(du, rh);
        v = 1.0 * (L * L);
      } else {
        v = op_apply64<D>(P.op, P.nl, P.nt, du, rh);
      }
      *val = v;
      if (!J) return db - 1;
      const int vb = s.vrule.on ? lin_rule_bin(s.vrule, v, P.n_val) : val_bin64(v, vE, P.n_val + 1, s);
      if (bins) bins[1] = vb - 1;
      if (vb < 1 || vb > P.n_val) return -1;
      return (db - 1) + P.n_dist * (vb - 1);
    }
    // squared-distance plan: binary search of r^2 in the squared edges, or (SquaredLinearPlan,
    // BinEdges.jl:506,516) the linear rule on sqrt(r^2)
    db = s.drule.on ? lin_rule_bin(s.drule, __builtin_sqrt(r2), P.n_dist) : dist_bin64(r2, tab, dE, nEd, s);
    if (db < 1 || db > P.n_dist) return -1;
    if (bins) bins[0] = db - 1;
    if constexpr (OPK == 3 || OPK == 4) {
      const double inv_r = 1.0 / __builtin_sqrt(r2);
      const double duL = sdot<D>(du, dx) * inv_r;
      const double dn2 = sdot<D>(du, du);
      if constexpr (OPK == 3) {
        val[0] = duL;
        val[1] = dn2;
        return db - 1;
      } else {
        v = sp_invariant(s.plane, duL, dn2);
      }
    } else if constexpr (OPK == 2) {
      const double p = sdot<D>(du, dx);
      v = p * p / r2;
    } else if constexpr (OPK == 1) {
      const double sq = __builtin_sqrt(r2);
      double rh[D];
#pragma unroll
      for (int k = 0; k < D; ++k) rh[k] = dx[k] / sq;
      v = 1.0 * sdot<D>(du, rh);
    } else {
      v = sf_raw64<D>(P.op, P.nl, P.nt, du, dx, r2);
    }
  } else {
    SphereFrame f;
    bool exact_dist = true;
    if constexpr (CERT) {
      f = sphere_frame64_noangle(A, B);
      const float key = (float)sdot<3>(f.d, f.d);
      const int c = cell32(key, s);
      const float2 lh = lohi[c];
      exact_dist = (key >= lh.x) & (key <= lh.y);
      db = (key > lh.y) ? blo_tab[c].x + 1 : blo_tab[c].x;
      if (exact_dist) f.distance = sphere_distance64(A, B, f.d, P.radius);
    } else {
      f = sphere_frame64(A, B, P.radius);
    }
    if (exact_dist)
      db = s.drule.on ? lin_rule_bin(s.drule, f.distance, P.n_dist) : dist_bin64(f.distance, tab, dE, nEd, s);
    if (!f.ok || db < 1 || db > P.n_dist) return -1;
    if (bins) bins[0] = db - 1;
    double du[3];
    sphere_delta64<D>(f, A, B, du);
    const double rh[3] = {1.0, 0.0, 0.0};
    if constexpr (OPK == 3) {
      val[0] = sdot<D>(du, rh);
      val[1] = sdot<D>(du, du);
      return db - 1;
    } else if constexpr (OPK == 4) {
      v = sp_invariant(s.plane, sdot<D>(du, rh), sdot<D>(du, du));
    } else if constexpr (OPK == 1) v = 1.0 * sdot<D>(du, rh);
    else if constexpr (OPK == 2) {
      const double L = sdot<D>(du, rh);
      v = 1.0 * (L * L);
    } else {
      v = op_apply64<D>(P.op, P.nl, P.nt, du, rh);
    }
  }
  *val = v;
  if (!J) return db - 1;
  const int vb = s.vrule.on ? lin_rule_bin(s.vrule, v, P.n_val) : val_bin64(v, vE, P.n_val + 1, s);
  if (bins) bins[1] = vb - 1;
  if (vb < 1 || vb > P.n_val) return -1;
  return (db - 1) + P.n_dist * (vb - 1);
}

// =============================================================================================
// prep kernels
// =============================================================================================
struct PrepSlice {
  const double *x, *u;
  int64_t N, pt_off;
  const int32_t *perm;  // binary32 record k is point perm[k] of the slice (NULL: see mul)
  int64_t mul;          // without perm: record k is point (k * mul) mod N (0: identity)
};

template <int G, int D>
__global__ void prep64(const PrepSlice *ps, int geom_code, double *rec64) {
  const PrepSlice s = ps[blockIdx.y];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= s.N) return;
  constexpr int S = Rec<G, D>::d64;
  double *r = rec64 + (s.pt_off + i) * S;
  if (G == kFlat) {
#pragma unroll
    for (int k = 0; k < D; ++k) {
      r[k] = s.x[D * i + k];
      r[D + k] = s.u[D * i + k];
    }
  } else {
    const double lon = s.x[2 * i], lat = s.x[2 * i + 1];
    double sl, cl, sp, cp;
    if (geom_code == 1) {
      sf_sincosd(lon, &sl, &cl);
      sf_sincosd(lat, &sp, &cp);
    } else {
      sf_sincos(lon, &sl, &cl);
      sf_sincos(lat, &sp, &cp);
    }
    r[0] = cp * cl; r[1] = cp * sl; r[2] = sp;     // unit position
    r[3] = -sl; r[4] = cl;                          // east (third component 0)
    r[5] = -sp * cl; r[6] = -sp * sl; r[7] = cp;   // north
    r[8] = s.u[D * i]; r[9] = s.u[D * i + 1];
    r[10] = (D == 3) ? s.u[D * i + 2] : 0.0;
    r[11] = 0.0;
  }
}

// binary32 value >= x for x >= 0 (norms used in error bounds)
DEV float round_up_f(double x) {
  float f = (float)x;
  if ((double)f < x) f = __int_as_float(__float_as_int(f) + 1);
  return f;
}

DEV void split2(double v, float *hi, float *lo) {
  const float h = (float)v;
  *hi = h;
  *lo = (float)(v - (double)h);
}

template <int G, int D>
__global__ void prep32(const PrepSlice *ps, const double *rec64, float *rec32) {
  const PrepSlice s = ps[blockIdx.y];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= s.N) return;
  constexpr int S = Rec<G, D>::d64;
  constexpr int F = Rec<G, D>::f32;
  const int64_t src = s.perm ? (int64_t)s.perm[i] : (s.mul ? (i * s.mul) % s.N : i);
  const double *r = rec64 + (s.pt_off + src) * S;
  float *o = rec32 + (s.pt_off + i) * F;
  for (int k = 0; k < F; ++k) o[k] = 0.0f;
  if (G == kFlat) {
    double n2 = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      split2(r[k], &o[k], &o[D + k]);
      split2(r[D + k], &o[2 * D + k], &o[3 * D + k]);
      n2 += r[D + k] * r[D + k];
    }
    o[4 * D] = round_up_f(__builtin_sqrt(n2) * (1.0 + 1e-12));
  } else {
#pragma unroll
    for (int k = 0; k < 3; ++k) split2(r[k], &o[k], &o[3 + k]);
    const double ue = r[8], un = r[9];
    const double E[3] = {r[3], r[4], 0.0}, Nn[3] = {r[5], r[6], r[7]};
#pragma unroll
    for (int k = 0; k < 3; ++k) split2(ue * E[k] + un * Nn[k], &o[6 + k], &o[9 + k]);
    const double n2 = ue * ue + un * un + (D == 3 ? r[10] * r[10] : 0.0);
    o[12] = round_up_f(__builtin_sqrt(n2) * (1.0 + 1e-12));
    split2(r[10], &o[13], &o[14]);
  }
  o[Rec<G, D>::id] = __int_as_float((int)src);
}

// =============================================================================================
// persistent histogram kernels
// =============================================================================================
struct Smem {
  double *hs;
  unsigned *hc;
  Cell *tab;      // binary64 kernel: cell table (bx)
  float2 *lohi;   // binary32 kernel: ambiguity interval per cell ...
  int2 *blo;      // ... and bin count below it (.x; .y holds it minus one, the 0-based bin of keys
                  // below the interval; same LDS region as tab, kCell32Max * 8 bytes
                  // after lohi at the same stride, so that one address serves both reads; one
                  // 16-byte read of a whole cell measured 2 % slower than these 8 + 4 bytes)
  double *dE;
  double *vE;
  float *vE32;
  char *stage;
  int ncell_h;
  int copies;     // 1D: lane-group copies; joint: copies by lane halves (1 or 2)
  bool joint;
  int nd, nslots;
  int jv, jd, jslots;  // joint slot of (d, v) in one copy: v jv + d jd; a copy holds jslots slots
  int coff;            // this lane's copy offset (joint; slots)
  int scr;             // this lane's scratch slot (target of updates that add nothing)
  int nmom;     // sums per slot, moment-major: moment k of slot h at hs[k * mstride + h]
  int mstride;  // nslots + kDummySlots; slot nslots + lane is lane's scratch slot
  bool fx;                  // fixed-point joint cells (hq aliases hs; no hc)
  unsigned long long *hq;
  double *zs;               // fixed-point joint: binary64 cells of small values (kZeroBins per
  unsigned *zc;             // distance bin) and their counts
  // fused 1D outputs: output k's sums at ms[k mstr + slot], shared counts mc[slot]; slot =
  // distance bin * mcop + lane group; mscr is the lane's scratch slot
  double *ms;
  unsigned *mc;
  int n1d, mcop, mstr, mscr;
  int qcap;                 // re-bin queue staged in LDS: entries in all
  int qw;                   // entries per wave: wave w owns q[w qw, (w + 1) qw) and fill qn[w]
  unsigned *qn;
  int2 *q;
};

// JT: the launch keeps a joint histogram (P.n_val > 0), known at compile time
template <bool JT>
DEV Smem carve(char *base, const Params &P, int stage_bytes) {
  const SmemLayout L = smem_layout(P.n_dist, P.n_val, P.ncell_max, stage_bytes, P.copies, P.nmom, P.fx,
                                   P.qcap, P.strip, P.jlayout, P.n1d, P.mcopies);
  Smem m;
  m.nmom = P.nmom;
  m.fx = P.fx != 0;
  m.ncell_h = P.n_dist * (P.n_val > 0 ? P.n_val : 1);
  m.copies = P.copies;
  m.joint = P.n_val > 0;
  m.nd = P.n_dist;
  const JointLayout JL = joint_layout(P.n_dist, P.n_val, P.jlayout);
  m.jv = JL.jv;
  m.jd = JL.jd;
  m.jslots = P.n_val > 0 ? JL.slots : 0;
  m.coff = (P.n_val > 0 && P.copies == 2) ? (int)(__lane_id() >> 5) * JL.slots : 0;
  m.nslots = P.n_val > 0 ? JL.slots * P.copies : P.n_dist * P.copies;
  // B-strip: field b's histogram at slots [b nslots, (b + 1) nslots), scratch slots after all
  m.mstride = m.nslots * (P.strip > 1 ? P.strip : 1) + kDummySlots;
  m.scr = m.nslots * (P.strip > 1 ? P.strip : 1) + (int)__lane_id();
  m.hs = (double *)(base + L.hs);
  m.hq = (unsigned long long *)(base + L.hs);
  m.hc = (unsigned *)(base + L.hc);
  m.zs = (double *)(base + L.zs);
  m.zc = (unsigned *)(base + L.zc);
  m.ms = (double *)(base + L.ms);
  m.mc = (unsigned *)(base + L.mc);
  m.n1d = P.n1d;
  m.mcop = P.mcopies > 0 ? P.mcopies : 1;
  m.mstr = P.n_dist * m.mcop + kDummySlots;
  m.mscr = P.n_dist * m.mcop + (int)__lane_id();
  m.tab = (Cell *)(base + L.tab);
  m.lohi = (float2 *)(base + L.tab);
  m.blo = (int2 *)(base + L.tab + 8 * (JT ? kCell32Max : P.ncell_max));
  m.dE = (double *)(base + L.dE);
  m.vE = (double *)(base + L.vE);
  m.vE32 = (float *)(base + L.vE32);
  m.stage = base + L.stage;
  m.qcap = P.qcap;
  m.qw = P.qw;
  m.qn = (unsigned *)(base + L.q);
  m.q = (int2 *)(base + L.q + kQueueHead);
  return m;
}

DEV void zero_hist(const Smem &m) {
  if (!m.fx)
    for (int c = threadIdx.x; c < m.mstride; c += blockDim.x) m.hc[c] = 0u;
  for (int c = threadIdx.x; c < m.mstride * m.nmom; c += blockDim.x) m.hs[c] = 0.0;
  if (m.fx)
    for (int c = threadIdx.x; c < kZeroBins * m.nd + kDummySlots; c += blockDim.x) {
      m.zs[c] = 0.0;
      m.zc[c] = 0u;
    }
  for (int c = threadIdx.x; c < m.mstr * m.n1d; c += blockDim.x) m.ms[c] = 0.0;
  if (m.n1d > 0)
    for (int c = threadIdx.x; c < m.mstr; c += blockDim.x) m.mc[c] = 0u;
}

// fused 1D outputs of slice slice_id: every output gets the shared counts and its own sums
DEV void flush_multi(const Smem &m, const Params &P, int slice_id) {
  for (int c = threadIdx.x; c < m.nd; c += blockDim.x) {
    unsigned long long n = 0;
    for (int k = 0; k < m.mcop; ++k) {
      n += m.mc[c * m.mcop + k];
      m.mc[c * m.mcop + k] = 0u;
    }
    for (int o = 0; o < m.n1d; ++o) {
      double v = 0.0;
      for (int k = 0; k < m.mcop; ++k) {
        const int h = o * m.mstr + c * m.mcop + k;
        v += m.ms[h];
        m.ms[h] = 0.0;
      }
      if (n) {
        const int64_t off = ((int64_t)o * P.n_slices + slice_id) * m.nd + c;
        atomicAdd(&P.mcounts[off], n);
        atomicAdd(&P.msums[off], v);
      }
    }
  }
}

// LDS slot of output cell c = db0 + n_dist * vb0 (0-based bins) for this lane
DEV int hslot(const Smem &m, int db0, int vb0) {
  if (m.joint) return vb0 * m.jv + db0 * m.jd + m.coff;
  return db0 * m.copies + (__lane_id() & (m.copies - 1));
}

// one pair into LDS slot h: the value, or (single pass) the four moments of (du_L, |du|^2) in the
// reference's order: |du|^2, du_L^2, du_L |du|^2, du_L du_L^2 (serial_single_pass.jl:257-264)
template <typename T>
DEV void hist_add(const Smem &m, int h, T v) {
  atomicAdd(&m.hs[h], (double)v);
  atomicAdd(&m.hc[h], 1u);
}
template <typename T>
DEV void hist_add_sp(const Smem &m, int h, T duL, T dn2) {
  const T duL2 = duL * duL;
  atomicAdd(&m.hs[h], (double)dn2);
  atomicAdd(&m.hs[m.mstride + h], (double)duL2);
  atomicAdd(&m.hs[2 * m.mstride + h], (double)(duL * dn2));
  atomicAdd(&m.hs[3 * m.mstride + h], (double)(duL * duL2));
  atomicAdd(&m.hc[h], 1u);
}

DEV void flush_hist(const Smem &m, const Params &P, const Slice &s) {
  // output cell c of the slice is out_off + c, or every sixth cell for a 2D single-pass plane
  const int64_t ostride = sp2d(P.op, P.n_val) ? kSpRows : 1;
  if (m.fx) {
    // fixed-point joint cells: n = word >> 44, sum v = (sum q + n (2^23 - c_g)) / vsf
    for (int c = threadIdx.x; c < m.ncell_h; c += blockDim.x) {
      const int db0 = c % m.nd, vb0 = c / m.nd;
      const int h0 = vb0 * m.jv + db0 * m.jd;
      unsigned long long n = 0, sq = 0;
      for (int k = 0; k < m.copies; ++k) {
        const int h = h0 + k * m.jslots;
        const unsigned long long wd = m.hq[h];
        m.hq[h] = 0ull;
        // each update added kFxUnit = 2^44 + kFxBase plus its q < 2^23; fewer than 2^20 updates
        // keep sum q below kFxUnit, so the quotient is the count
        const unsigned long long nk = wd / kFxUnit;
        n += nk;
        sq += wd - nk * kFxUnit;
      }
      if (n) {
        // c_g exactly as the pair loop formed it; every term below is exact in binary64 but the
        // division
        const float cg = __builtin_fmaf(-(float)vb0, s.fxc1, s.fxc0);
        const double v = ((double)sq + (double)n * (8388608.0 - (double)cg)) / (double)s.vsf;
        const int64_t o = s.out_off + ostride * c;
        atomicAdd(&P.counts[o], n);
        atomicAdd(&P.sums[o], v);
      }
    }
    for (int c = threadIdx.x; c < kZeroBins * m.nd; c += blockDim.x) {
      const unsigned n = m.zc[c];
      const double v = m.zs[c];
      m.zc[c] = 0u;
      m.zs[c] = 0.0;
      const int db0 = c / kZeroBins, vb0 = s.gz + c % kZeroBins;
      if (n && vb0 >= 0 && vb0 < P.n_val) {
        const int64_t o = s.out_off + ostride * (db0 + m.nd * vb0);
        atomicAdd(&P.counts[o], (unsigned long long)n);
        atomicAdd(&P.sums[o], v);
      }
    }
    return;
  }
  if (m.nmom > 1) {
    // single pass: distance bins only, per-lane copies, rows S2 L2 S3 L3 of a (6, n_dist) output
    for (int c = threadIdx.x; c < m.nd; c += blockDim.x) {
      unsigned long long n = 0;
      double v[4] = {0.0, 0.0, 0.0, 0.0};
      for (int k = 0; k < m.copies; ++k) {
        const int h = c * m.copies + k;
        n += m.hc[h];
        m.hc[h] = 0u;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          v[q] += m.hs[q * m.mstride + h];
          m.hs[q * m.mstride + h] = 0.0;
        }
      }
      if (n) {
        const int64_t o = s.out_off + (int64_t)kSpRows * c;
        atomicAdd(&P.counts[o], n);
#pragma unroll
        for (int q = 0; q < 4; ++q) atomicAdd(&P.sums[o + sp_row(q)], v[q]);
      }
    }
    return;
  }
  for (int c = threadIdx.x; c < m.ncell_h; c += blockDim.x) {
    unsigned long long n = 0;
    double v = 0.0;
    if (m.joint) {
      const int db0 = c % m.nd, vb0 = c / m.nd;
      const int h0 = vb0 * m.jv + db0 * m.jd;
      for (int k = 0; k < m.copies; ++k) {
        const int h = h0 + k * m.jslots;
        n += m.hc[h];
        v += m.hs[h];
        m.hc[h] = 0u;
        m.hs[h] = 0.0;
      }
    } else {
      for (int k = 0; k < m.copies; ++k) {
        n += m.hc[c * m.copies + k];
        v += m.hs[c * m.copies + k];
        m.hc[c * m.copies + k] = 0u;
        m.hs[c * m.copies + k] = 0.0;
      }
    }
    if (n) {
      atomicAdd(&P.counts[s.out_off + ostride * c], n);
      atomicAdd(&P.sums[s.out_off + ostride * c], v);
    }
  }
}

DEV void load_tables(const Smem &m, const Params &P, const Slice &s, bool fast32) {
  const Cell *src = P.cells + (fast32 ? s.tab32_off : s.tab_off);
  const int nc = fast32 ? s.ncell32 : s.ncell64;
  if (fast32) {
    for (int c = threadIdx.x; c < nc; c += blockDim.x) {
      const Cell e = src[c];
      m.lohi[c] = make_float2(e.lo, e.hi);
      m.blo[c] = make_int2(e.blo, e.blo - 1);
    }
  } else {
    for (int c = threadIdx.x; c < nc; c += blockDim.x) m.tab[c] = src[c];
  }
  for (int k = threadIdx.x; k <= P.n_dist; k += blockDim.x) m.dE[k] = P.dedges[s.dedge_off + k];
  if (P.n_val > 0) {
    for (int k = threadIdx.x; k <= P.n_val; k += blockDim.x) {
      m.vE[k] = P.vedges[s.vedge_off + k];
      m.vE32[k] = P.vedges32[s.vedge_off + k];
    }
  }
}

// ---------------------------------------------------------------------------------------------
// hist64: binary64 exact pair arithmetic for every pair
// ---------------------------------------------------------------------------------------------
template <int G, int D, bool J, int OPK, bool MASK>
DEV void pair64(const Params &P, const Slice &s, const Smem &m, const double *ri, int i, int j,
                const double *rj) {
  if (MASK && j <= i) return;
  double v[2] = {0.0, 0.0};
  int bins[2] = {0, 0};
  int c;
  if constexpr (G == kSphere)
    c = exact_cell<G, D, J, OPK, true>(P, s, ri, rj, P.cells + s.tab_off, m.dE, m.vE, v, m.lohi, m.blo, bins);
  else
    c = exact_cell<G, D, J, OPK>(P, s, ri, rj, m.tab, m.dE, m.vE, v, nullptr, nullptr, bins);
  // the slot from the bins (no division of c), and a pair that is not counted goes to the lane's
  // scratch slot (never flushed): the update is unconditional, the loop body has no branch here
  int h = J ? bins[1] * m.jv + bins[0] * m.jd + m.coff : bins[0] * m.copies + (int)(__lane_id() & (m.copies - 1));
  h = c >= 0 ? h : m.scr;
  if constexpr (OPK == 3) hist_add_sp(m, h, v[0], v[1]);
  else hist_add(m, h, v[0]);
}

template <int S>
DEV void load_rec64(const double *__restrict__ p, double *r) {
#pragma unroll
  for (int k = 0; k < S; ++k) r[k] = p[k];
}

// Columns j0 .. j1 - 1 of one row: the column records are read straight from global memory at
// wave-uniform addresses (scalar loads into SGPRs: no LDS stage, no barrier per chunk), two
// record buffers alternating so that the next column's record is requested before the current
// pair is evaluated.  The request past the last column reads at most one record beyond the
// slice: the next slice's first record, or the record array's tail pad.
template <int G, int D, bool J, int OPK, bool MASK>
DEV void run_cols64(const Params &P, const Slice &s, const Smem &m, const double *ri, int i, int j0, int j1,
                    const double *__restrict__ rec) {
  constexpr int S = Rec<G, D>::d64;
  const double *col = rec + (int64_t)s.pt_off * S;
  double ra[S], rb[S];
  load_rec64<S>(col + (int64_t)j0 * S, ra);
  int j = j0;
  for (; j + 1 < j1; j += 2) {
    load_rec64<S>(col + (int64_t)(j + 1) * S, rb);
    pair64<G, D, J, OPK, MASK>(P, s, m, ri, i, j, ra);
    load_rec64<S>(col + (int64_t)(j + 2) * S, ra);
    pair64<G, D, J, OPK, MASK>(P, s, m, ri, i, j + 1, rb);
  }
  if (j < j1) pair64<G, D, J, OPK, MASK>(P, s, m, ri, i, j, ra);
}

// Row of the work item's tile that this thread owns.  Joint histograms: lane l of wave w takes
// row l (B / 64) + w, so that a wave's 64 rows lie B / 64 points apart rather than side by side.
// Neighbouring points of a smooth field have nearly the same distance to a column and nearly the
// same velocity, so a wave of adjacent rows sends many lanes to one joint slot in the same
// instruction, and LDS atomics to one address serialise (binary64 spherical joint kernel: 192 ->
// 162 ms at N = 262144).  1D histograms keep adjacent rows: their lane-group copies already keep
// the atomics apart, and adjacent rows read the same cells of the key table (a broadcast), where
// spread rows read scattered ones (4-5 % slower on the sphere).  Every tile still covers rows
// i0 .. i0 + B - 1.
template <bool J>
DEV int tile_row() {
  const int B = blockDim.x;
  const int t = threadIdx.x;
  return (!J || (B & 63)) ? t : (t & 63) * (B >> 6) + (t >> 6);
}

// Sphere, and flat point fields (not the batch-leading pairs): the same pairs as pair64 /
// exact_cell, with the reference's operation sequence for every value, arranged like the binary32
// kernel's pair.  The distance bin is read from the binary32 key table (chord^2 or r^2: its
// margins, 2^-20 relative and more, exceed the binary32 rounding of the binary64 key); the
// previous pair's histogram update is deferred until this pair's table read is in flight (LDS
// operations complete in order: the wait for the read would otherwise wait for the atomics), the
// next column's record is requested only after that wait (a scalar load in flight would be
// waited for with it), and the body has no branch but the rare one that bins a pair whose key is
// in a cell's ambiguity interval (sphere: with the central angle) or whose value is within its
// margin of a value decision point, by the exact rules.  Flat keys that are not above 2^-100
// (coincident points, NaN) also take the exact rules.  A pair that is not counted updates the
// lane's scratch slot.
struct Pending64 {
  int h;
  double v, w;  // value (single pass: du_L) and |du|^2 (single pass)
};

template <int G, int D, bool J, int OPK, bool MASK, bool VU>
DEV void pair64p(const Params &P, const Slice &s, const Smem &m, const double *A, int i, int j,
                 const double *B, double *rn, const double *next, Pending64 &pa) {
  constexpr int S = Rec<G, D>::d64;
  SphereFrame f;
  double dx[D], du[3], r2 = 0.0;
  float key;
  if constexpr (G == kSphere) {
    f = sphere_frame64_noangle(A, B);
    key = (float)sdot<3>(f.d, f.d);
  } else {
#pragma unroll
    for (int k = 0; k < D; ++k) {
      dx[k] = B[k] - A[k];
      du[k] = B[D + k] - A[D + k];
    }
    r2 = sdot<D>(dx, dx);
    key = (float)r2;
  }
  const int c = cell32(key, s);
  const float2 lh = m.lohi[c];
  const int blo = m.blo[c].x;
  __builtin_amdgcn_sched_barrier(0);
  if constexpr (OPK == 3) hist_add_sp(m, pa.h, pa.v, pa.w);
  else hist_add(m, pa.h, pa.v);
  asm volatile("" ::"v"(lh.x), "v"(lh.y), "v"(blo) : "memory");
  // outside [lo, hi] a key is above the edge exactly when key >= lo
  const bool ge = key >= lh.x;
  bool exact_dist = ge & (key <= lh.y);
  if constexpr (G == kFlat) exact_dist = exact_dist | !(key > 0x1p-100f);
  int db = ge ? blo + 1 : blo;
  load_rec64<S>(next, rn);
  __builtin_amdgcn_sched_barrier(0);
  double v, w = 0.0;
  if constexpr (G == kFlat) {
    // exact_cell's flat point-field values (sf_raw64 and its specialisations)
    if constexpr (OPK == 3 || OPK == 4) {
      const double inv_r = 1.0 / __builtin_sqrt(r2);
      const double duL = sdot<D>(du, dx) * inv_r;
      const double dn2 = sdot<D>(du, du);
      if constexpr (OPK == 3) {
        v = duL;
        w = dn2;
      } else {
        v = sp_invariant(s.plane, duL, dn2);
      }
    } else if constexpr (OPK == 2) {
      const double p = sdot<D>(du, dx);
      v = p * p / r2;
    } else if constexpr (OPK == 1) {
      const double sq = __builtin_sqrt(r2);
      double rh[D];
#pragma unroll
      for (int k = 0; k < D; ++k) rh[k] = dx[k] / sq;
      v = 1.0 * sdot<D>(du, rh);
    } else {
      v = sf_raw64<D>(P.op, P.nl, P.nt, du, dx, r2);
    }
  } else {
    sphere_delta64<D>(f, A, B, du);
    const double rh[3] = {1.0, 0.0, 0.0};
    if constexpr (OPK == 3) {
      v = sdot<D>(du, rh);
      w = sdot<D>(du, du);
    } else if constexpr (OPK == 4) {
      v = sp_invariant(s.plane, sdot<D>(du, rh), sdot<D>(du, du));
    } else if constexpr (OPK == 1) {
      v = 1.0 * sdot<D>(du, rh);
    } else if constexpr (OPK == 2) {
      const double L = sdot<D>(du, rh);
      v = 1.0 * (L * L);
    } else {
      v = op_apply64<D>(P.op, P.nl, P.nt, du, rh);
    }
  }
  // value bin (1-based).  Uniform value edges (a LinearBinEdges rule or an equally spaced edge
  // vector): read from t = (v - v0) / w in binary32, certified like the binary32 kernel's (the
  // host slack covers the edges' distance from v0 + k w, and the rule's decision points' distance
  // from the edges), with twice its relative term for the rounding of v - v0 to binary32.  Only a
  // value within that margin of a decision point (or non-finite) takes the exact rule or search.
  int vb = 1;
  bool amb_v = false;
  if constexpr (J) {
    if constexpr (VU) {  // s.v_uniform, chosen once per column range
      const float t = (float)(v - s.v0) * s.vinvf;
      const float eps = __builtin_fmaf(9.6e-7f, __builtin_fabsf(t), s.vslack);
      const float gh = __builtin_floorf(t + eps);
      amb_v = !(t - eps >= gh);
      vb = amb_v ? 0 : (int)gh + 1;  // |t| < 2^21 when certain: the conversion is exact
    } else {
      vb = s.vrule.on ? lin_rule_bin(s.vrule, v, P.n_val) : val_bin64(v, m.vE, P.n_val + 1, s);
    }
  }
  if (exact_dist | amb_v) {
    if (exact_dist) {
      if constexpr (G == kSphere) {
        f.distance = sphere_distance64(A, B, f.d, P.radius);
        db = s.drule.on ? lin_rule_bin(s.drule, f.distance, P.n_dist)
                        : dist_bin64(f.distance, P.cells + s.tab_off, m.dE, P.n_dist + 1, s);
      } else {
        // squared-distance plan, or (SquaredLinearPlan) the linear rule on sqrt(r^2)
        db = s.drule.on ? lin_rule_bin(s.drule, __builtin_sqrt(r2), P.n_dist)
                        : dist_bin64(r2, P.cells + s.tab_off, m.dE, P.n_dist + 1, s);
      }
    }
    if (J && amb_v) vb = s.vrule.on ? lin_rule_bin(s.vrule, v, P.n_val) : val_bin64(v, m.vE, P.n_val + 1, s);
  }
  const bool ok = (G == kSphere) ? f.ok : true;
  const bool add = (!MASK || j > i) & ok & ((unsigned)(db - 1) < (unsigned)P.n_dist) &
                   (!J || (unsigned)(vb - 1) < (unsigned)P.n_val);
  const int h = J ? (vb - 1) * m.jv + (db - 1) * m.jd + m.coff
                  : (db - 1) * m.copies + (int)(__lane_id() & (m.copies - 1));
  pa.h = add ? h : m.scr;
  pa.v = v;
  pa.w = w;
}

// Columns j0 .. j1 - 1 of one row with pair64p; two record buffers alternate (the request past
// the last column reads at most one record beyond the slice, as in run_cols64).
template <int G, int D, bool J, int OPK, bool MASK, bool VU>
DEV void run_cols64pv(const Params &P, const Slice &s, const Smem &m, const double *ri, int i, int j0, int j1,
                      const double *__restrict__ rec) {
  constexpr int S = Rec<G, D>::d64;
  const double *col = rec + (int64_t)s.pt_off * S;
  double ra[S], rb[S];
  load_rec64<S>(col + (int64_t)j0 * S, ra);
  Pending64 pa{m.scr, 0.0, 0.0};
  int j = j0;
  for (; j + 1 < j1; j += 2) {
    pair64p<G, D, J, OPK, MASK, VU>(P, s, m, ri, i, j, ra, rb, col + (int64_t)(j + 1) * S, pa);
    pair64p<G, D, J, OPK, MASK, VU>(P, s, m, ri, i, j + 1, rb, ra, col + (int64_t)(j + 2) * S, pa);
  }
  if (j < j1) pair64p<G, D, J, OPK, MASK, VU>(P, s, m, ri, i, j, ra, rb, col + (int64_t)(j + 1) * S, pa);
  if constexpr (OPK == 3) hist_add_sp(m, pa.h, pa.v, pa.w);
  else hist_add(m, pa.h, pa.v);
}

// the value-bin lookup (uniform or exact) is chosen once per column range, not per pair
template <int G, int D, bool J, int OPK, bool MASK>
DEV void run_cols64p(const Params &P, const Slice &s, const Smem &m, const double *ri, int i, int j0, int j1,
                     const double *__restrict__ rec) {
  if (J && s.v_uniform) run_cols64pv<G, D, J, OPK, MASK, true>(P, s, m, ri, i, j0, j1, rec);
  else run_cols64pv<G, D, J, OPK, MASK, false>(P, s, m, ri, i, j0, j1, rec);
}

template <int G, int D, bool J, int OPK>
__global__ void __launch_bounds__(1024) hist64(Params P, const double *__restrict__ rec) {
  extern __shared__ __align__(16) char smem_raw[];
  constexpr int S = Rec<G, D>::d64;
  const Smem m = carve<J>(smem_raw, P, P.stage_bytes);
  __shared__ int s_item;
  const int B = blockDim.x;
  int cur = -1;
  Slice s{};
  unsigned long long since_flush = 0;
  zero_hist(m);
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_item = atomicAdd(P.counter, 1);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(s_item);  // uniform: item and slice in SGPRs
    if (it >= P.n_items) break;
    const Item w = P.items[it];
    const unsigned long long work = (unsigned long long)B * (unsigned long long)(w.j1 - w.j0);
    if (w.slice != cur || since_flush + work > (unsigned long long)P.flush_limit) {
      if (cur >= 0) {
        flush_hist(m, P, s);
        __syncthreads();
      }
      if (w.slice != cur) {
        cur = w.slice;
        s = P.slices[cur];
        // sphere and flat point fields: the binary32 key table of certified distance bins (the
        // binary64 key table is read from global memory by the rare pairs that need it)
        load_tables(m, P, s, G == kSphere || !P.bl);
      }
      since_flush = 0;
      __syncthreads();  // tables and cleared histogram visible before the item's pairs
    }
    since_flush += work;
    const int i = w.i0 + tile_row<J>();
    const bool iv = i < s.N;
    double ri[S];
#pragma unroll
    for (int k = 0; k < S; ++k) ri[k] = iv ? rec[(s.pt_off + i) * S + k] : 0.0;
    if (iv) {
      if (G == kSphere || !P.bl) {
        if (w.j0 < w.i0 + B) run_cols64p<G, D, J, OPK, true>(P, s, m, ri, i, w.j0, w.j1, rec);
        else run_cols64p<G, D, J, OPK, false>(P, s, m, ri, i, w.j0, w.j1, rec);
      } else {
        if (w.j0 < w.i0 + B) run_cols64<G, D, J, OPK, true>(P, s, m, ri, i, w.j0, w.j1, rec);
        else run_cols64<G, D, J, OPK, false>(P, s, m, ri, i, w.j0, w.j1, rec);
      }
    }
  }
  __syncthreads();
  if (cur >= 0) flush_hist(m, P, s);
}

// ---------------------------------------------------------------------------------------------
// hist32: binary32 pair arithmetic with certified binning, uncertain pairs queued
// ---------------------------------------------------------------------------------------------
DEV float frsq(float x) { return __builtin_amdgcn_rsqf(x); }

// Nearly coincident points go to the binary64 re-bin: the binary32 records carry positions as
// hi + lo (2^-48 relative), a direction error of 2^-48 / separation that below a separation of
// 2^-22 (sphere: radians; flat: times max|x|) would outgrow the value bound's margin and the
// reference's own binary64 accuracy (2^-53 / separation).  Smallest sin^2(sigma) (sphere) and
// squared distance (flat) evaluated in binary32:
inline constexpr float kCloseS2 = 0x1p-44f;
DEV float close_key(const Slice &s) { return __builtin_fmaxf(1e-30f, 0x1p12f * s.xerr * s.xerr); }

DEV int cell32(float key, const Slice &s) {
  const int bits = __float_as_int(key);
  const int c = (bits >> (23 - s.M32)) - s.base32;
  // clamp to [0, ncell32 - 1] in one instruction (the compiler forms max + min)
  int r;
  asm("v_med3_i32 %0, %1, 0, %2" : "=v"(r) : "v"(c), "s"(s.ncell32 - 1));
  return r;
}


// value-axis bin from a binary32 value with absolute error bound dv = dva * dvb; sets amb when
// undecidable.  For uniform edges (VU) dva carries the factor 1 / bin width (the bound in bin
// widths is then one multiply-add away).  Returns the 0-based value bin g (valid bins 0 .. nE-2;
// anything else is out of range) and g as a float in gf.
template <bool VU>
DEV int val_bin32(float v, float dva, float dvb, const Slice &s, const float *E32, int nE, bool &amb,
                  float &gf) {
  if constexpr (VU) {  // uniform value edges (s.v_uniform; chosen once per column range)
    // t = v / w - v0 / w in one fma (its product exact, one rounding)
    const float t = __builtin_fmaf(v, s.vinvf, s.vtc);
    // error of t: value bound (1/w > 0), host slack (edge non-uniformity, rounding of v0, 1/w and
    // -v0 / w), and the rounding of the fma and the two offsets below (a few ulp of |t|).  Undecidable iff an integer lies in (t - eps, t + eps].  Far outside the edges
    // (|t| >= 2^21) eps exceeds 1, so such values are undecidable too (and re-binned); the
    // float-to-int conversion saturates, so no clamp is needed.  A non-finite v is undecidable
    // without a test of its own: t = +-Inf makes eps infinite and t - eps or t + eps NaN, and a
    // NaN fails the comparison below.
    const float eps = __builtin_fmaf(dva, dvb, __builtin_fmaf(4.8e-7f, __builtin_fabsf(t), s.vslack));
    // floor(t - eps) != floor(t + eps) exactly when t - eps < floor(t + eps) (NaN: undecidable)
    const float gh = __builtin_floorf(t + eps);
    amb = amb || !(t - eps >= gh);
    gf = gh;
    return (int)gh;
  }
  // a non-finite value is undecidable here; the lookup below still runs, its bin is ignored
  amb = amb || !(__builtin_fabsf(v) < __builtin_inff());
  int lo = 0;
  if (s.v_log) {
    // log-spaced edges (positive from edge vk0 on): the first edge >= v guessed from log2(v),
    // then a walk to it, which alone decides (v <= 0 and NaN start from edge 0)
    float tg = __builtin_fmaf(__builtin_amdgcn_logf(v) - s.vl0f, s.vlinvf, (float)s.vk0 + 1.0f);
    tg = __builtin_fminf(__builtin_fmaxf(tg, 0.0f), (float)nE);
    lo = (int)tg;
    while (lo > 0 && !(E32[lo - 1] < v)) --lo;
    while (lo < nE && E32[lo] < v) ++lo;
  } else {
    int hi = nE;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (E32[mid] < v) lo = mid + 1; else hi = mid;
    }
  }
  // nearest edges must be farther than the error bound (edges rounded to binary32: slack)
  const float gap = __builtin_fmaf(dva, dvb, __builtin_fmaf(s.vslack, __builtin_fabsf(v), s.vabs)) + 1e-37f;
  if (lo > 0 && v - E32[lo - 1] <= gap) amb = true;
  if (lo < nE && E32[lo] - v <= gap) amb = true;
  gf = (float)(lo - 1);
  return lo - 1;
}

// binary32 value of the operator from du_L, du_T (signed transverse), ||du||^2 - du_L^2.
DEV float op_value32(int op, int nl, int nt, int D, float L, float T, float tn2, float n2) {
  switch (op) {
    case 0: {
      float ex = 1.0f;
      if (nl == 1) ex = L;
      else if (nl == 2) ex = L * L;
      else if (nl == 3) ex = L * L * L;
      if (nt == 1) ex = ex * T;
      else if (nt == 2) ex = ex * tn2;
      else if (nt == 3) ex = ex * (T * T * T);
      return ex;
    }
    case 1: return n2;
    case 2: return L * n2;
    case 3: return tn2 / (float)(D - 1);
    case 4: return L * (tn2 / (float)(D - 1));
    default: {
      if (nl == 2) return n2;
      const float r = __builtin_sqrtf(n2);
      return nl == 1 ? r : r * r * r;
    }
  }
}

// Operator class fixed at compile time: 0 = any operator (runtime switch), 1 = du_L
// (ProjectedStructureFunctionType{1,0}), 2 = du_L^2 ({2,0}); binary32 1D kernels also 5 = du_L^3
// ({3,0}), 6 = du_T ({0,1}), 7 = du_T^2 ({0,2}), 8 = du_T^3 ({0,3}), with the operation order of
// the runtime switch (op_value32).
template <int OPK>
DEV float op32(const Params &P, int D, float L, float T, float tn2, float n2) {
  if constexpr (OPK == 1) return L;
  else if constexpr (OPK == 2) return L * L;
  else if constexpr (OPK == 5) return L * L * L;
  else if constexpr (OPK == 6) return T;
  else if constexpr (OPK == 7) return tn2;
  else if constexpr (OPK == 8) return T * T * T;
  else return op_value32(P.op, P.nl, P.nt, D, L, T, tn2, n2);
}

// (|u_i| + |u_j|)^order without a loop (order 0..3)
template <int OPK>
DEV float pow_order(const Params &P, float x) {
  if constexpr (OPK >= 9) {
    // fused: the joint operator's order, selected without a branch
    const float x2 = x * x, x3 = x2 * x;
    const float t = P.order >= 2 ? x2 : (P.order == 1 ? x : 1.0f);
    return P.order == 3 ? x3 : t;
  } else if constexpr (OPK == 1 || OPK == 6) return x;
  else if constexpr (OPK == 2 || OPK == 7) return x * x;
  else if constexpr (OPK == 5 || OPK == 8) return x * x * x;
  else {
    const float x2 = x * x;
    return P.order == 0 ? 1.0f : (P.order == 1 ? x : (P.order == 2 ? x2 : x2 * x));
  }
}

template <int F>
DEV void load_rec(const float4 *col, int idx, float *r) {
#pragma unroll
  for (int k = 0; k < F / 4; ++k) {
    const float4 q = col[idx * (F / 4) + k];
    r[4 * k] = q.x; r[4 * k + 1] = q.y; r[4 * k + 2] = q.z; r[4 * k + 3] = q.w;
  }
}

// Deferred histogram update of the previous pair of this lane: slot (-1: none) and value(s).
// It is issued after the next pair's table read: LDS operations complete in order, so the wait
// for that read does not include the atomics' latency.
struct PendingAdd {
  int h;
  float v, w;
  unsigned q;  // fixed-point joint cells: the quantized value
};

// One pair (i, j): certified distance bin from the binary32 key, operator value, certified value
// bin, and the deferred update.  rj is column j's record; the record of column j + 1 (if
// j + 1 < j_end) is requested into rn.  The body has no branch on the pair's own state: every
// pair is evaluated (out-of-range pairs are simply not counted), so the wave never splits.
// Returns whether the pair is undecidable in binary32 and must be re-binned in binary64; the
// caller collects these flags and queues the pairs every few columns (queue_deferred).
template <int G, int D, bool J, int OPK, bool MASK, bool FX, bool VU, bool SL>
DEV bool pair32(const Params &P, const Slice &s, const Smem &m, const float *ri, int i, int j,
                const float *rj, float *rn, const float4 *next, bool has_next, PendingAdd &pa) {
  constexpr int F = Rec<G, D>::f32;
  // OPK 3 (single pass) and 4 (a plane of its 2D form) need du_L and |du|^2; on the sphere
  // |du|^2 = du_L^2 + |du_T|^2
  constexpr bool SP = (OPK == 3);
  constexpr bool SP2 = (OPK == 4);
  constexpr bool needT = (OPK == 0) || (OPK >= 6) || ((SP || SP2) && G == kSphere);
  // OPK 9, 10, 11: fused outputs -- the joint operator (if J) and P.n1d <= NF 1D operators per
  // pair, every value a product of two basis values (fused_basis_index)
  constexpr bool MULTI = (OPK >= 9);
  constexpr int NF = OPK == 9 ? 1 : (OPK == 10 ? 2 : (OPK == 11 ? 4 : (OPK == 12 ? 6 : kMaxFused)));
  constexpr bool needN2 = (OPK == 0) || MULTI || SP || SP2;
  const int nb = P.n_dist;
  const bool upper = !MASK || j > i;
  bool amb, amb_d;
  float key, L = 0.f, T = 0.f, tn2 = 0.f, n2 = 0.f, amp = 1.0f, umax;
  int db0;  // 0-based distance bin
  // ---- distance key and certified bin
  float dx[3], du[3], d[3];
  if constexpr (G == kFlat) {
#pragma unroll
    for (int k = 0; k < D; ++k) dx[k] = (rj[k] - ri[k]) + (rj[D + k] - ri[D + k]);
    key = dx[0] * dx[0];
#pragma unroll
    for (int k = 1; k < D; ++k) key = __builtin_fmaf(dx[k], dx[k], key);
    umax = ri[4 * D] + rj[4 * D];
  } else {
#pragma unroll
    for (int k = 0; k < 3; ++k) d[k] = (rj[k] - ri[k]) + (rj[3 + k] - ri[3 + k]);
    key = __builtin_fmaf(d[2], d[2], __builtin_fmaf(d[1], d[1], d[0] * d[0]));
    umax = ri[12] + rj[12];
  }
  {
    const int c = cell32(key, s);
    const float2 lh = m.lohi[c];
    const int bm1 = m.blo[c].y;  // bin count below the cell's interval, minus one
    // the previous pair's update, unconditional (a lane with nothing pending holds its scratch
    // slot in pa.h) so that the wait below is a partial one on every path
    __builtin_amdgcn_sched_barrier(0);
    const int hh = pa.h;
    if constexpr (FX) {
      atomicAdd(&m.hq[hh], (1ull << kFxShift) | (unsigned long long)pa.q);
    } else if constexpr (SP) {
      hist_add_sp(m, hh, pa.v, pa.w);
    } else {
      hist_add(m, hh, pa.v);
    }
    // consume the table entry here (an empty asm reading it, ordered before later memory
    // operations): the next column's record is requested only after this point, because a
    // scalar load still in flight would make the wait for the table read wait for it as well
    asm volatile("" ::"v"(lh.x), "v"(lh.y), "v"(bm1) : "memory");
    // two compares: a key that is not in [lo, hi] is above the edge exactly when key >= lo (a
    // NaN key fails both: not ambiguous here, bin blo, and degenerate by the tests below)
    const bool ge = key >= lh.x;
    amb_d = ge & (key <= lh.y);
    db0 = ge ? bm1 + 1 : bm1;
  }
  // scalar-load variant: the request is unconditional (past the last column it reads the record
  // array's one-record tail pad) and pinned here, so that the scheduler does not sink it towards
  // its use and this pair's arithmetic covers its latency
  if constexpr (SL) {
    load_rec<F>(next, 0, rn);
    __builtin_amdgcn_sched_barrier(0);
  } else if (has_next) {
    load_rec<F>(next, 0, rn);
  }
  amb = amb_d;
  const bool db_ok = (unsigned)db0 < (unsigned)nb;
  bool vb_ok = true;
  float v = 0.f;
  int vb0 = 0;
  // value bound inputs (joint histograms): largest |component| of the velocity difference, the
  // flat direction-error factor, and the floor of the difference scale (below)
  float mx = 0.f, xamp = 1.0f, dfloor = 0.f;
  constexpr bool TIGHT = J && (OPK != 1);
  if constexpr (G == kFlat) {
#pragma unroll
    for (int k = 0; k < D; ++k) du[k] = (rj[2 * D + k] - ri[2 * D + k]) + (rj[3 * D + k] - ri[3 * D + k]);
    float p = du[0] * dx[0];
#pragma unroll
    for (int k = 1; k < D; ++k) p = __builtin_fmaf(du[k], dx[k], p);
    const float rinv = frsq(key);
    L = p * rinv;
    if constexpr (J) {
      // the reference's binary64 dx carries an absolute error ~2^-53 max|x|, which turns r's
      // direction by that much over |dx|: a relative error of its du_L of 2^-53 max|x| / |dx|
      xamp = __builtin_fmaf(s.xerr, rinv, 1.0f);
      if constexpr (TIGHT) {
        mx = __builtin_fmaxf(__builtin_fabsf(du[0]), __builtin_fabsf(du[1]));
        if constexpr (D == 3) mx = __builtin_fmaxf(mx, __builtin_fabsf(du[2]));
        dfloor = 0x1p-20f * umax;
      }
    }
    if constexpr (needN2) {
      n2 = du[0] * du[0];
#pragma unroll
      for (int k = 1; k < D; ++k) n2 = __builtin_fmaf(du[k], du[k], n2);
    }
    if constexpr (needT) {
      // signed transverse component (n = zhat x rhat) and transverse energy ||du x rhat||^2,
      // formed from the cross product so that it does not cancel when du is nearly longitudinal
      const float cz = __builtin_fmaf(du[1], dx[0], -du[0] * dx[1]);
      if constexpr (D == 2) {
        T = cz * rinv;
        tn2 = T * T;
      } else {
        T = cz * frsq(__builtin_fmaf(dx[1], dx[1], dx[0] * dx[0]));
        const float cx = __builtin_fmaf(du[2], dx[1], -du[1] * dx[2]);
        const float cy = __builtin_fmaf(du[0], dx[2], -du[2] * dx[0]);
        tn2 = __builtin_fmaf(cz, cz, __builtin_fmaf(cy, cy, cx * cx)) * (rinv * rinv);
      }
    }
    amb = amb || !(key > close_key(s));
  } else {
    // geodesic frame (HelperFunctions.jl:186-200): t_A = (d - (p.d) p)/sin, t_B = (d + (p.d) q)/sin
    // with d = q - p, sin^2(sigma) = |p x d|^2.  U_A is tangent at p and U_B at q, so
    //   du_L = t_B.U_B - t_A.U_A = (d.(U_B - U_A) + (p.d)(q.U_B + p.U_A))/sin = d.dU / sin
    //   du_T = (p x d).dU / sin
    // (the binary32 error of this form stays within the bound used below, like the long form)
    float w[3], dU[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) dU[k] = (rj[6 + k] - ri[6 + k]) + (rj[9 + k] - ri[9 + k]);
    w[0] = __builtin_fmaf(ri[1], d[2], -ri[2] * d[1]);
    w[1] = __builtin_fmaf(ri[2], d[0], -ri[0] * d[2]);
    w[2] = __builtin_fmaf(ri[0], d[1], -ri[1] * d[0]);
    const float s2 = __builtin_fmaf(w[2], w[2], __builtin_fmaf(w[1], w[1], w[0] * w[0]));
    amb = amb || !(s2 > kCloseS2);
    const float inv_s = frsq(s2);
    L = __builtin_fmaf(d[2], dU[2], __builtin_fmaf(d[1], dU[1], d[0] * dU[0])) * inv_s;
    float dw = 0.f;
    if constexpr (needT) {
      T = __builtin_fmaf(w[2], dU[2], __builtin_fmaf(w[1], dU[1], w[0] * dU[0])) * inv_s;
      tn2 = T * T;
      if constexpr (D == 3) {
        dw = (rj[13] - ri[13]) + (rj[14] - ri[14]);
        tn2 = __builtin_fmaf(dw, dw, tn2);
      }
      n2 = __builtin_fmaf(L, L, tn2);
    }
    // binary32 error of du_L, du_T grows like (1 + c^2 / sin(sigma)): near antipodal pairs the
    // numerators cancel and sin(sigma) is small
    amp = __builtin_fmaf(key, inv_s, 1.0f);
    if constexpr (TIGHT) {
      mx = __builtin_fmaxf(__builtin_fmaxf(__builtin_fabsf(dU[0]), __builtin_fabsf(dU[1])), __builtin_fabsf(dU[2]));
      if constexpr (needT && D == 3) mx = __builtin_fmaxf(mx, __builtin_fabsf(dw));
      // the reference's own binary64 frame error for close points grows like 2^-53 / sin(sigma)
      dfloor = umax * __builtin_fmaf(0x1p-26f, inv_s, 0x1p-20f);
    }
  }
  // fused outputs: the basis X = {1, du_L, du_L^2, du_L^3, du_T, tn2, du_T^3, |du|^2,
  // tn2 / (D - 1)}; each value is X[a] * X[b] with wave-uniform a, b (nine entries: the compiler
  // reads them with register indexing, one move each, no compare chain)
  float X[MULTI ? 9 : 1];
  if constexpr (MULTI) {
    const float L2 = L * L;
    X[0] = 1.0f; X[1] = L; X[2] = L2; X[3] = L2 * L; X[4] = T; X[5] = tn2; X[6] = (T * T) * T;
    X[7] = n2; X[8] = (D == 3) ? tn2 * 0.5f : tn2;
  }
  if constexpr (SP) {
    // moments are formed at the accumulation; all four must be finite in binary32
    const float L2 = L * L;
    amb = amb || !(n2 < __builtin_inff()) || !(__builtin_fabsf(L * n2) < __builtin_inff()) ||
          !(__builtin_fabsf(L * L2) < __builtin_inff());
  } else if constexpr (SP2) {
    v = sp_invariant(s.plane, L, n2);
  } else if constexpr (MULTI) {
    v = X[P.ja] * X[P.jb];
  } else {
    v = op32<OPK>(P, D, L, T, tn2, n2);
  }
  if constexpr (J) {
    // |v - v_ref| <= coef * 2^-24 * m^order * amp, m bounding |du| and the binary32 error's
    // scale.  m = ||u_i|| + ||u_j|| always holds.  Where the velocities differ little (close
    // points of a smooth field) the error scales with the difference instead: with dU the
    // difference of the hi/lo velocity records, |du_L| <= |dU| amp and |du| <= sqrt(3) max|dU_k|
    // (flat) or 3 max(|dU_k|, |du_r|) (sphere), plus a floor of 2^-20 (||u_i|| + ||u_j||) that
    // covers the 2^-48 relative error of the hi/lo split and the reference's own binary64
    // rounding (checked by a binary32 emulation against binary64 over near, far, near-antipodal
    // pairs and differences down to 1e-13 of the velocities: at most 0.2 of the coefficients
    // below).  Not for the du_L kernel, whose linear value bins gain nothing from it.
    float mm = umax;
    const float ampv = (G == kFlat) ? xamp : amp;
    if constexpr (TIGHT)
      mm = __builtin_fminf(umax, __builtin_fmaf(G == kFlat ? 1.7320509f : 3.0f, mx, dfloor) * ampv);
    // 2D single-pass planes: coefficient 32 * order, order 2 for |du|^2, du_L^2, du_T^2 and 3
    // for the products with du_L (the difference du_T^2 = |du|^2 - du_L^2 and the products add
    // the errors of both factors, hence twice the coefficient of one operator)
    // the bound is dva * ampv (with uniform value edges, in units of the bin width)
    float dva;
    if constexpr (SP2) {
      const float u2 = mm * mm;
      dva = s.plane < 3 ? 0x1p-18f * u2 : 0x1.8p-18f * u2 * mm;
      if constexpr (VU) dva *= s.vinvf;
    } else if constexpr (VU) {
      if constexpr (OPK == 1 && !TIGHT) {
        // K (|u_i| + |u_j|) as one fma on the row's K |u_i| (loop-invariant): the same bound to
        // an ulp, against which the coefficient keeps a factor of 5 (umax itself is then unused)
        constexpr int UI = (G == kFlat) ? 4 * D : 12;
        const float kv = P.vcoef * s.vinvf;
        dva = __builtin_fmaf(kv, rj[UI], kv * ri[UI]);
      } else {
        dva = (P.vcoef * s.vinvf) * pow_order<OPK>(P, mm);
      }
    } else {
      dva = P.vcoef * pow_order<OPK>(P, mm);
    }
    // both velocities tiny: a bound of order 2 or 3 could underflow in binary32 while it still
    // matters in units of a (tiny) bin, so re-bin in binary64.  Order 1 needs no test: a bound
    // that underflows is below 2^-126 bin widths, far inside the slack, and a NaN fails the
    // value-bin test by itself.
    if constexpr (!(OPK == 1 || OPK == 6)) amb = amb || !(umax >= 0x1p-30f);
    float gf;
    vb0 = val_bin32<VU>(v, dva, ampv, s, m.vE32, P.n_val + 1, amb, gf);
    vb_ok = (unsigned)vb0 < (unsigned)P.n_val;
    if constexpr (FX) {
      // offset within the value bin in units of 2^-22 bin widths, biased, rounded to nearest:
      // v vsf + c_g with c_g ~ 2^23 + kFxBias - a_g vsf (one fma, exact before its rounding)
      // lies in [2^23, 2^24), where binary32 has unit spacing: the fma rounds to the nearest
      // integer and the bits are kFxBase + q (no conversion needed)
      pa.q = __float_as_uint(__builtin_fmaf(v, s.vsf, __builtin_fmaf(-gf, s.fxc1, s.fxc0)));
    }
  } else if constexpr (!SP) {
    amb = amb || !(__builtin_fabsf(v) < __builtin_inff());
  }
  // fused 1D outputs: every value must be finite in binary32 (else the pair is re-binned)
  // fused 1D outputs: NF values per pair, unconditionally (outputs past P.n1d have a = b = 0)
  float mv[MULTI ? NF : 1];
  if constexpr (MULTI) {
    // one finiteness test for all: a sum of magnitudes is finite iff every term is (a sum that
    // overflows from finite terms only re-bins the pair in binary64)
    float mag = 0.f;
#pragma unroll
    for (int k = 0; k < NF; ++k) {
      // the indices are extracted here, by an opaque scalar bit-field extract: hoisted out of the
      // loop, 2 NF indices would hold SGPRs of their own (at NF = 8 they spilled to VGPR lanes,
      // read back with a v_readlane per lookup)
      int ia, ib;
      asm volatile("s_bfe_u32 %0, %1, %2" : "=s"(ia) : "s"(P.mab_a), "i"((4 * k) | (4 << 16)));
      asm volatile("s_bfe_u32 %0, %1, %2" : "=s"(ib) : "s"(P.mab_b), "i"((4 * k) | (4 << 16)));
      mv[k] = X[ia] * X[ib];
      mag += __builtin_fabsf(mv[k]);
    }
    amb = amb || !(mag < __builtin_inff());
  }
  // binned here: certain and in range; queued: undecidable and not certainly out of range.  A
  // fused call without a joint output bins through its 1D histograms only.
  const bool add = upper & !amb & db_ok & vb_ok & !(MULTI && !J);
  const bool defer = upper & amb & (amb_d | db_ok);
  if constexpr (MULTI) {
    // the fused 1D outputs of a certain, in-range distance bin (whatever the joint value bin)
    const bool add1 = upper & !amb & db_ok;
    const int h1 = add1 ? db0 * m.mcop + (int)(__lane_id() & (m.mcop - 1)) : m.mscr;
    atomicAdd(&m.mc[h1], 1u);
    double *const mp = m.ms + h1;
#pragma unroll
    for (int k = 0; k < NF; ++k) {
      // output k's histogram is k mstr slots on: the offset is formed here by an opaque scalar
      // multiply, not hoisted into an SGPR per output (those spilled to VGPR lanes)
      int off = 0;
      if (k > 0) asm volatile("s_mul_i32 %0, %1, %2" : "=s"(off) : "s"(m.mstr), "i"(k));
      atomicAdd(mp + off, (double)mv[k]);
    }
  }
  int h;
  // 24-bit multiplies (full rate): every operand of a slot that is used is a small bin index;
  // values of a pair that is not added may be garbage and are selected away below
  if constexpr (J) {
    // two 24-bit multiply-adds (the compiler forms two multiplies and a three-way add)
    int hv;
    asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(hv) : "v"(vb0), "s"(m.jv), "v"(m.coff));
    asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(h) : "v"(db0), "s"(m.jd), "v"(hv));
  } else {
    h = db0 * m.copies + (__lane_id() & (m.copies - 1));
  }
  if constexpr (FX) {
    // a value below a quarter bin width goes to the binary64 cell of its (distance, value) bin
    // at once (few lanes, so a branch): the fixed-point error (up to 2^-22 bin widths) would not
    // be small relative to it.  The other pairs take the deferred fixed-point update.  The clamp
    // keeps the cell in range; the host's choice of gz already puts every such bin inside it
    // (a range test in its place measured 0.8 % slower)
    const int k = min(max(vb0 - s.gz, 0), kZeroBins - 1);
    const bool small = add & (__builtin_fabsf(v) < s.vsmall);
    if (small) {
      const int z = db0 * kZeroBins + k;
      atomicAdd(&m.zs[z], (double)v);
      atomicAdd(&m.zc[z], 1u);
    }
    pa.h = (add & !small) ? h : m.scr;
  } else {
    pa.h = add ? h : m.scr;
  }
  pa.v = SP ? L : v;
  pa.w = n2;
  // scalar-load variant: an empty asm that reads the last three words of the next record makes
  // the compiler fetch the 16-float record in one 64-byte request rather than a split 8+4+1 (the
  // scalar data cache is request-bound here).  It sits at the end of the pair so that the wait
  // it implies comes after this pair's arithmetic, not right after the request.
  if constexpr (SL && F == 16) asm volatile("" ::"s"(rn[13]), "s"(rn[14]), "s"(rn[15]));
  return defer;
}

// Exclusive prefix sum of n (< 64) over the active lanes of the wave, from one ballot per bit;
// *total receives the sum.  Works with any set of active lanes.
DEV unsigned wave_prefix(unsigned n, unsigned *total) {
  const unsigned long long lt = (1ull << __lane_id()) - 1ull;
  unsigned tot = 0, pre = 0;
#pragma unroll
  for (int b = 0; b < 6; ++b) {
    const unsigned long long bal = __ballot((n >> b) & 1u);
    tot += (unsigned)__popcll(bal) << b;
    pre += (unsigned)__popcll(bal & lt) << b;
  }
  *total = tot;
  return pre;
}

// One atomic add of `total` for the whole wave, by its first active lane; returns the old value
// to every lane.
template <typename T>
DEV T wave_reserve(T *counter, T total) {
  const int leader = __ffsll((long long)__ballot(1)) - 1;
  T base = 0;
  if ((int)__lane_id() == leader) base = atomicAdd(counter, total);
  return __shfl(base, leader);
}

// the wave's index in its workgroup, in an SGPR
DEV int wave_index() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

// Orders the wave's LDS accesses to its queue segment: entries written by one lane are read by
// another lane of the same wave.
DEV void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Move the first `fill` entries of the wave's LDS segment to the global queue: one reservation
// for the wave, the entries spread over its active lanes.
DEV void queue_wave_move(const Params &P, const Smem &m, unsigned fill, int slice_id, int fields) {
  const unsigned long long act = __ballot(1);
  const unsigned na = (unsigned)__popcll(act);
  const unsigned rank = (unsigned)__popcll(act & ((1ull << __lane_id()) - 1ull));
  const unsigned long long base = wave_reserve<unsigned long long>(P.pending_count, (unsigned long long)fill);
  wave_lds_sync();
  const int2 *seg = m.q + wave_index() * m.qw;
  for (unsigned k = rank; k < fill; k += na) {
    const unsigned long long idx = base + k;
    if (idx < (unsigned long long)P.pending_cap) {
      const int2 e = seg[k];
      P.pending[idx] = Pending{slice_id, e.x, e.y, fields};
    }
  }
  wave_lds_sync();
}

// Queue the pairs (i, jg + step t) whose bit t is set in dm for the binary64 re-bin, as record
// indices (the re-bin kernel maps them to the caller's indices).  Called by every active lane of
// the wave after a group of columns, outside the pair loop: a branch there (with its global atomic
// and lane exchange) splits the loop body into pieces the compiler cannot overlap.  Each wave owns
// a segment of the workgroup's LDS queue and moves it to the global queue itself when the next
// group would not fit, and when the slice changes (queue_wave_flush): one global reservation per
// segment rather than per group.  The global queue counter is shared by every wave of the device,
// and under that contention a returning atomic costs thousands of cycles; the waves of a
// workgroup never wait on each other for the queue.  A group larger than a segment goes to the
// global queue directly.
DEV void queue_deferred(const Params &P, const Smem &m, unsigned dm, int i, int jg, int slice_id,
                        int fields = 0, int step = 1) {
  if (__ballot(dm != 0u) == 0ull) return;  // the common case: nothing to queue in the wave
  const unsigned n = (unsigned)__popc(dm);
  unsigned tot;
  const unsigned pre = wave_prefix(n, &tot);  // tot > 0: some lane has a bit set
  const unsigned qw = (unsigned)m.qw;
  if (tot <= qw) {
    unsigned *fillp = m.qn + wave_index();
    unsigned fill = (unsigned)__builtin_amdgcn_readfirstlane((int)*fillp);
    if (fill + tot > qw) {
      queue_wave_move(P, m, fill, slice_id, fields);
      fill = 0u;
    }
    int2 *seg = m.q + wave_index() * m.qw + fill + pre;
    unsigned e = 0;
    while (dm) {
      const int t = __ffs((int)dm) - 1;
      dm &= dm - 1u;
      seg[e++] = make_int2(i, jg + step * t);
    }
    *fillp = fill + tot;  // every active lane stores the same value
    return;
  }
  unsigned long long g = wave_reserve<unsigned long long>(P.pending_count, (unsigned long long)tot) + pre;
  while (dm) {
    const int t = __ffs((int)dm) - 1;
    dm &= dm - 1u;
    if (g < (unsigned long long)P.pending_cap) P.pending[g] = Pending{slice_id, i, jg + step * t, fields};
    ++g;
  }
}

// Move the wave's queued entries (pairs of slice slice_id) to the global queue.  Called by whole
// waves, before the workgroup moves to another slice and before it ends.
DEV void queue_wave_flush(const Params &P, const Smem &m, int slice_id, int fields = 0) {
  if (m.qw == 0) return;
  unsigned *fillp = m.qn + wave_index();
  const unsigned fill = (unsigned)__builtin_amdgcn_readfirstlane((int)*fillp);
  if (fill == 0u) return;
  queue_wave_move(P, m, fill, slice_id, fields);
  *fillp = 0u;
}

template <int G, int D, bool J, int OPK, bool MASK, bool FX, bool VU, bool SL>
DEV void run_chunk32_vu(const Params &P, const Slice &s, const Smem &m, const float *ri, int i,
                        int j_begin, int j_end, const float4 *col, int col_j0, int slice_id) {
  constexpr int F = Rec<G, D>::f32;
  constexpr int R = F / 4;  // float4 per record
  if (j_begin >= j_end) return;
  float ra[F], rb[F];
  // cp walks the column records two at a time; the loads use constant offsets from it
  const float4 *cp = col + (j_begin - col_j0) * R;
  load_rec<F>(cp, 0, ra);
  PendingAdd pa{m.scr, 0.f, 0.f, 0u};
  int j = j_begin;
  // groups of 32 columns: the flags are shifted in, so bit t of dm flags pair (i, je - 1 - t)
  // for the re-bin queue (one shift-add per pair)
  while (j < j_end) {
    const int jg = j;
    const int je = min(j_end, jg + 32);
    unsigned dm = 0u;
    for (; j + 1 < je; j += 2, cp += 2 * R) {
      dm = (dm << 1) | (unsigned)pair32<G, D, J, OPK, MASK, FX, VU, SL>(P, s, m, ri, i, j, ra, rb, cp + R, true, pa);
      dm = (dm << 1) |
           (unsigned)pair32<G, D, J, OPK, MASK, FX, VU, SL>(P, s, m, ri, i, j + 1, rb, ra, cp + 2 * R, j + 2 < j_end, pa);
    }
    // an odd count only occurs in the last group (the others hold 32 columns)
    if (j < je) {
      dm = (dm << 1) | (unsigned)pair32<G, D, J, OPK, MASK, FX, VU, SL>(P, s, m, ri, i, j, ra, rb, cp + R, false, pa);
      ++j;
    }
    queue_deferred(P, m, dm, i, je - 1, slice_id, 0, -1);
  }
  if constexpr (FX) {
    atomicAdd(&m.hq[pa.h], (1ull << kFxShift) | (unsigned long long)pa.q);
  } else if constexpr (OPK == 3) {
    hist_add_sp(m, pa.h, pa.v, pa.w);
  } else {
    hist_add(m, pa.h, pa.v);
  }
}

// the value-edge lookup (uniform or searched) is chosen once per column range, not per pair
template <int G, int D, bool J, int OPK, bool MASK, bool FX, bool SL>
DEV void run_chunk32(const Params &P, const Slice &s, const Smem &m, const float *ri, int i,
                     int j_begin, int j_end, const float4 *col, int col_j0, int slice_id) {
  if (FX || (J && s.v_uniform))
    run_chunk32_vu<G, D, J, OPK, MASK, FX, true, SL>(P, s, m, ri, i, j_begin, j_end, col, col_j0, slice_id);
  else
    run_chunk32_vu<G, D, J, OPK, MASK, FX, false, SL>(P, s, m, ri, i, j_begin, j_end, col, col_j0, slice_id);
}

// SLOAD: column records are read straight from global memory with wave-uniform addresses, which
// the compiler turns into scalar loads (SGPR operands, scalar cache) -- no LDS staging, no
// barriers inside an item.  Otherwise columns are staged through LDS in chunks of P.tj.
// Fixed-point SLOAD kernels are held to 6 waves per SIMD (80 VGPRs): two 768-thread workgroups
// per CU, which their small LDS footprint allows; the few values this spills are per-item ones
// (the fused ones spill a few SGPR-held values of the pair loop too, and still gain: the 256
// ragged 2D slices' du_L^2 and joint du_L from one launch, 73.7 -> 57.6 ms).
template <int G, int D, bool J, int OPK, bool SLOAD, bool FX>
__global__ void __launch_bounds__(1024, (FX && SLOAD) ? 6 : 1) hist32(Params P, const float *__restrict__ rec) {
  extern __shared__ __align__(16) char smem_raw[];
  constexpr int F = Rec<G, D>::f32;
  const Smem m = carve<J>(smem_raw, P, P.stage_bytes);
  __shared__ int s_item;
  const int B = blockDim.x;
  int cur = -1;
  Slice s{};
  unsigned long long since_flush = 0;
  zero_hist(m);
  if ((threadIdx.x & 63) == 0 && m.qcap > 0) m.qn[threadIdx.x >> 6] = 0u;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_item = atomicAdd(P.counter, 1);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(s_item);  // uniform: item and slice in SGPRs
    if (it >= P.n_items) break;
    const Item w = P.items[it];
    const unsigned long long work = (unsigned long long)B * (unsigned long long)(w.j1 - w.j0);
    if (w.slice != cur || since_flush + work > (unsigned long long)P.flush_limit) {
      if (cur >= 0) {
        flush_hist(m, P, s);
        if constexpr (OPK >= 9) flush_multi(m, P, cur);
        __syncthreads();
      }
      if (w.slice != cur) {
        if (cur >= 0) queue_wave_flush(P, m, cur);
        cur = w.slice;
        s = P.slices[cur];
        load_tables(m, P, s, true);
        if (SLOAD) __syncthreads();
      }
      since_flush = 0;
    }
    since_flush += work;
    const int i = w.i0 + tile_row<J>();
    const bool iv = i < s.N;
    float ri[F];
#pragma unroll
    for (int k = 0; k < F; ++k) ri[k] = iv ? rec[(s.pt_off + i) * F + k] : 0.0f;
    if constexpr (SLOAD) {
      if (iv) {
        // the slice offset is wave-uniform by construction; say so, so that column addresses
        // stay in SGPRs (the pointer itself keeps its global address space: scalar loads)
        const uint64_t po = (uint64_t)s.pt_off;
        const int64_t off = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(po >> 32)) << 32) |
                                      (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)po));
        const float4 *col = (const float4 *)(rec + off * F);
        const int jm = min(w.j1, w.i0 + B);  // columns that can still be <= some row of the tile
        if (w.j0 < jm) run_chunk32<G, D, J, OPK, true, FX, true>(P, s, m, ri, i, w.j0, jm, col, 0, cur);
        if (jm < w.j1) run_chunk32<G, D, J, OPK, false, FX, true>(P, s, m, ri, i, max(jm, w.j0), w.j1, col, 0, cur);
      }
    } else {
      for (int jb = w.j0; jb < w.j1; jb += P.tj) {
        const int cnt = min(P.tj, w.j1 - jb);
        __syncthreads();
        const float4 *src = (const float4 *)(rec + (s.pt_off + jb) * F);
        float4 *dst = (float4 *)m.stage;
        for (int k = threadIdx.x; k < cnt * (F / 4); k += B) dst[k] = src[k];
        __syncthreads();
        if (iv) {
          const float4 *col = (const float4 *)m.stage;
          if (jb < w.i0 + B) run_chunk32<G, D, J, OPK, true, FX, false>(P, s, m, ri, i, jb, jb + cnt, col, jb, cur);
          else run_chunk32<G, D, J, OPK, false, FX, false>(P, s, m, ri, i, jb, jb + cnt, col, jb, cur);
        }
      }
    }
  }
  __syncthreads();
  if (cur >= 0) {
    queue_wave_flush(P, m, cur);
    flush_hist(m, P, s);
    if constexpr (OPK >= 9) flush_multi(m, P, cur);
  }
}

// ---------------------------------------------------------------------------------------------
// hist32s: B-strip of fields sharing positions (auxiliary axes, batch_leading.jl:105-140): the
// pair's geometry -- distance key, certified bin, frame -- is evaluated once and each of the
// strip's velocity fields adds its value to its own LDS histogram (1D histograms).  Slices
// s0 .. s0 + nf - 1 are the fields; their binary32 records are in one order (same N, same
// positions), so column j of every field is record j of its slice.  Columns are staged through
// LDS, field-major.  An undecidable pair (distance within its margin, degenerate, or a
// non-finite value in any field) is queued once for all fields.
// ---------------------------------------------------------------------------------------------
template <int G, int D>
struct StripRec {
  // the strip's LDS stage per column: positions (flat x_hi, x_lo; sphere P_hi, P_lo) and per field
  // the velocity (flat u_hi, u_lo; sphere U_hi, U_lo, u_r hi/lo), each padded to whole float4s
  static constexpr int pos = (G == kFlat) ? 2 * D : 6;      // floats of the position part
  static constexpr int PS = (pos + 3) / 4 * 4;
  static constexpr int voff = (G == kFlat) ? 2 * D : 6;     // velocity part's offset in a record
  static constexpr int vel = (G == kFlat) ? 2 * D : (D == 3 ? 8 : 6);
  static constexpr int VS = (vel + 3) / 4 * 4;
};

template <int N4>
DEV void lds_read4(const float *src, float *dst) {
#pragma unroll
  for (int k = 0; k < N4 / 4; ++k) {
    const float4 q = ((const float4 *)src)[k];
    dst[4 * k] = q.x; dst[4 * k + 1] = q.y; dst[4 * k + 2] = q.z; dst[4 * k + 3] = q.w;
  }
}

template <int G, int D, int OPK, bool MASK>
DEV void run_chunk_strip(const Params &P, const Slice &s, const Smem &m, const float *ri,
                         const float (*rv)[StripRec<G, D>::VS], int nf, int i, int jb, int cnt,
                         int slice_id) {
  using SR = StripRec<G, D>;
  constexpr bool needT = (OPK == 0);
  constexpr bool needN2 = (OPK == 0);
  const float *spos = (const float *)m.stage;
  const float *svel = spos + (size_t)P.tj * SR::PS;
  const int tj = P.tj;
  const int nb = P.n_dist;
  const int slot0 = (int)(__lane_id() & (m.copies - 1));
  for (int jg = 0; jg < cnt; jg += 32) {
    const int je = min(cnt, jg + 32);
    unsigned dm = 0u;
    for (int jj = jg; jj < je; ++jj) {
      const int j = jb + jj;
      const bool upper = !MASK || j > i;
      float rj[SR::PS];
      lds_read4<SR::PS>(spos + jj * SR::PS, rj);
      float key, rinv = 0.f, inv_s = 0.f;
      float dx[3], d[3], w[3];
      if constexpr (G == kFlat) {
#pragma unroll
        for (int k = 0; k < D; ++k) dx[k] = (rj[k] - ri[k]) + (rj[D + k] - ri[D + k]);
        key = dx[0] * dx[0];
#pragma unroll
        for (int k = 1; k < D; ++k) key = __builtin_fmaf(dx[k], dx[k], key);
      } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) d[k] = (rj[k] - ri[k]) + (rj[3 + k] - ri[3 + k]);
        key = __builtin_fmaf(d[2], d[2], __builtin_fmaf(d[1], d[1], d[0] * d[0]));
      }
      const int c = cell32(key, s);
      const float2 lh = m.lohi[c];
      const int blo = m.blo[c].x;
      const bool amb_d = (key >= lh.x) & (key <= lh.y);
      const int db = (key > lh.y) ? blo + 1 : blo;
      const bool db_ok = (unsigned)(db - 1) < (unsigned)nb;
      bool amb = amb_d;
      if constexpr (G == kFlat) {
        rinv = frsq(key);
        amb = amb || !(key > close_key(s));
      } else {
        w[0] = __builtin_fmaf(ri[1], d[2], -ri[2] * d[1]);
        w[1] = __builtin_fmaf(ri[2], d[0], -ri[0] * d[2]);
        w[2] = __builtin_fmaf(ri[0], d[1], -ri[1] * d[0]);
        const float s2 = __builtin_fmaf(w[2], w[2], __builtin_fmaf(w[1], w[1], w[0] * w[0]));
        amb = amb || !(s2 > kCloseS2);
        inv_s = frsq(s2);
      }
      float vals[kStripMax];
#pragma unroll
      for (int b = 0; b < kStripMax; ++b) {
        vals[b] = 0.f;
        if (b < nf) {
          float rjb[SR::VS];
          lds_read4<SR::VS>(svel + ((size_t)b * tj + jj) * SR::VS, rjb);
          const float *rib = rv[b];
          float L, T = 0.f, tn2 = 0.f, n2 = 0.f;
          if constexpr (G == kFlat) {
            float du[3];
#pragma unroll
            for (int k = 0; k < D; ++k) du[k] = (rjb[k] - rib[k]) + (rjb[D + k] - rib[D + k]);
            float p = du[0] * dx[0];
#pragma unroll
            for (int k = 1; k < D; ++k) p = __builtin_fmaf(du[k], dx[k], p);
            L = p * rinv;
            if constexpr (needN2) {
              n2 = du[0] * du[0];
#pragma unroll
              for (int k = 1; k < D; ++k) n2 = __builtin_fmaf(du[k], du[k], n2);
            }
            if constexpr (needT) {
              const float cz = __builtin_fmaf(du[1], dx[0], -du[0] * dx[1]);
              if constexpr (D == 2) {
                T = cz * rinv;
                tn2 = T * T;
              } else {
                T = cz * frsq(__builtin_fmaf(dx[1], dx[1], dx[0] * dx[0]));
                const float cx = __builtin_fmaf(du[2], dx[1], -du[1] * dx[2]);
                const float cy = __builtin_fmaf(du[0], dx[2], -du[2] * dx[0]);
                tn2 = __builtin_fmaf(cz, cz, __builtin_fmaf(cy, cy, cx * cx)) * (rinv * rinv);
              }
            }
          } else {
            float dU[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) dU[k] = (rjb[k] - rib[k]) + (rjb[3 + k] - rib[3 + k]);
            L = __builtin_fmaf(d[2], dU[2], __builtin_fmaf(d[1], dU[1], d[0] * dU[0])) * inv_s;
            if constexpr (needT) {
              T = __builtin_fmaf(w[2], dU[2], __builtin_fmaf(w[1], dU[1], w[0] * dU[0])) * inv_s;
              tn2 = T * T;
              if constexpr (D == 3) {
                const float dw = (rjb[6] - rib[6]) + (rjb[7] - rib[7]);
                tn2 = __builtin_fmaf(dw, dw, tn2);
              }
              n2 = __builtin_fmaf(L, L, tn2);
            }
          }
          const float v = op32<OPK>(P, D, L, T, tn2, n2);
          amb = amb || !(__builtin_fabsf(v) < __builtin_inff());
          vals[b] = v;
        }
      }
      const bool add = upper & !amb & db_ok;
      const bool defer = upper & amb & (amb_d | db_ok);
      if (add) {
        // one count for the pair (every field of the strip shares the bin); one sum per field
        const int h = (db - 1) * m.copies + slot0;
        atomicAdd(&m.hc[h], 1u);
#pragma unroll
        for (int b = 0; b < kStripMax; ++b)
          if (b < nf) atomicAdd(&m.hs[b * m.nslots + h], (double)vals[b]);
      }
      if (defer) dm |= 1u << (jj - jg);
    }
    queue_deferred(P, m, dm, i, jb + jg, slice_id, nf);
  }
}

// flush a strip's histograms: the shared counts and each field's sums into the field's slice
DEV void flush_strip(const Smem &m, const Params &P, int s0, int nf) {
  for (int c = threadIdx.x; c < m.nd; c += blockDim.x) {
    unsigned long long n = 0;
    for (int k = 0; k < m.copies; ++k) {
      n += m.hc[c * m.copies + k];
      m.hc[c * m.copies + k] = 0u;
    }
    for (int b = 0; b < nf; ++b) {
      double v = 0.0;
      for (int k = 0; k < m.copies; ++k) {
        const int h = b * m.nslots + c * m.copies + k;
        v += m.hs[h];
        m.hs[h] = 0.0;
      }
      if (n) {
        const Slice sb = P.slices[s0 + b];
        atomicAdd(&P.counts[sb.out_off + c], n);
        atomicAdd(&P.sums[sb.out_off + c], v);
      }
    }
  }
}

template <int G, int D, int OPK>
__global__ void __launch_bounds__(1024) hist32s(Params P, const float *__restrict__ rec) {
  extern __shared__ __align__(16) char smem_raw[];
  using SR = StripRec<G, D>;
  constexpr int F = Rec<G, D>::f32;
  const Smem m = carve<false>(smem_raw, P, P.stage_bytes);
  __shared__ int s_item;
  const int B = blockDim.x;
  int cur = -1, nf = 1;
  Slice s{};
  unsigned long long since_flush = 0;
  zero_hist(m);
  if ((threadIdx.x & 63) == 0 && m.qcap > 0) m.qn[threadIdx.x >> 6] = 0u;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_item = atomicAdd(P.counter, 1);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(s_item);
    if (it >= P.n_items) break;
    const Item w = P.items[it];
    const unsigned long long work = (unsigned long long)B * (unsigned long long)(w.j1 - w.j0);
    if (w.slice != cur || since_flush + work > (unsigned long long)P.flush_limit) {
      if (cur >= 0) {
        flush_strip(m, P, cur, nf);
        __syncthreads();
      }
      if (w.slice != cur) {
        if (cur >= 0) queue_wave_flush(P, m, cur, nf);
        cur = w.slice;
        s = P.slices[cur];
        nf = s.strip_n > 1 ? s.strip_n : 1;
        load_tables(m, P, s, true);
      }
      since_flush = 0;
    }
    since_flush += work;
    const int i = w.i0 + (int)threadIdx.x;
    const bool iv = i < s.N;
    float ri[SR::pos];
#pragma unroll
    for (int k = 0; k < SR::pos; ++k) ri[k] = iv ? rec[(s.pt_off + i) * F + k] : 0.0f;
    float rv[kStripMax][SR::VS];
#pragma unroll
    for (int b = 0; b < kStripMax; ++b) {
#pragma unroll
      for (int k = 0; k < SR::VS; ++k) {
        rv[b][k] = 0.f;
        if (iv && b < nf && k < SR::vel) rv[b][k] = rec[(P.slices[cur + b].pt_off + i) * F + SR::voff + k];
      }
    }
    for (int jb = w.j0; jb < w.j1; jb += P.tj) {
      const int cnt = min(P.tj, w.j1 - jb);
      __syncthreads();
      // stage: positions of the columns, then each field's velocities (compact, float4-aligned)
      float *spos = (float *)m.stage;
      float *svel = spos + (size_t)P.tj * SR::PS;
      for (int t = threadIdx.x; t < cnt * (1 + nf); t += B) {
        const int jj = t % cnt, part = t / cnt;
        if (part == 0) {
          const float *src = rec + (s.pt_off + jb + jj) * F;
#pragma unroll
          for (int k = 0; k < SR::PS; ++k) spos[jj * SR::PS + k] = k < SR::pos ? src[k] : 0.f;
        } else {
          const int b = part - 1;
          const float *src = rec + (P.slices[cur + b].pt_off + jb + jj) * F + SR::voff;
          float *dst = svel + ((size_t)b * P.tj + jj) * SR::VS;
#pragma unroll
          for (int k = 0; k < SR::VS; ++k) dst[k] = k < SR::vel ? src[k] : 0.f;
        }
      }
      __syncthreads();
      if (iv) {
        if (jb < w.i0 + B) run_chunk_strip<G, D, OPK, true>(P, s, m, ri, rv, nf, i, jb, cnt, cur);
        else run_chunk_strip<G, D, OPK, false>(P, s, m, ri, rv, nf, i, jb, cnt, cur);
      }
    }
  }
  __syncthreads();
  if (cur >= 0) {
    queue_wave_flush(P, m, cur, nf);
    flush_strip(m, P, cur, nf);
  }
}

// ---------------------------------------------------------------------------------------------
// recheck: binary64 binning of queued pairs (global-memory tables, global atomics)
// ---------------------------------------------------------------------------------------------
// SP: the single pass; with J, one plane of its 2D form (output cells every sixth)
template <int G, int D, bool J, bool SP>
__global__ void __launch_bounds__(256) recheck(Params P) {
  // Each workgroup takes a contiguous range of the queue.  Entries of the range's first slice are
  // accumulated in an LDS histogram flushed once; entries of other slices (the queue is filled in
  // roughly slice order, so these are few) go straight to the global accumulators.
  extern __shared__ __align__(16) char smem_raw[];
  constexpr int S = Rec<G, D>::d64;
  constexpr bool SP1 = SP && !J;          // 1D single pass: four moments per distance bin
  constexpr int NM = SP1 ? 4 : 1;
  constexpr int64_t OSTR = SP ? kSpRows : 1;
  const int nh = P.n_dist * (P.n_val > 0 ? P.n_val : 1);
  double *hs = (double *)smem_raw;           // NM sums per cell, moment-major
  unsigned *hc = (unsigned *)(hs + NM * nh);
  for (int c = threadIdx.x; c < nh; c += blockDim.x) hc[c] = 0u;
  for (int c = threadIdx.x; c < NM * nh; c += blockDim.x) hs[c] = 0.0;
  const unsigned long long n0 = *P.pending_count;
  const unsigned long long n = n0 < (unsigned long long)P.pending_cap ? n0 : (unsigned long long)P.pending_cap;
  const unsigned long long per = (n + gridDim.x - 1) / gridDim.x;
  const unsigned long long beg = (unsigned long long)blockIdx.x * per;
  const unsigned long long end = beg + per < n ? beg + per : n;
  __syncthreads();
  if (beg < end) {
    const int home = P.pending[beg].slice;
    for (unsigned long long k = beg + threadIdx.x; k < end; k += blockDim.x) {
      const Pending q0 = P.pending[k];
      // a B-strip entry stands for the pair in each of its fields' slices
      const int nfields = q0.fields > 1 ? q0.fields : 1;
      for (int f = 0; f < nfields; ++f) {
      Pending q = q0;
      q.slice += f;
      const Slice s = P.slices[q.slice];
      // queued as binary32 record indices (records may be permuted or spatially sorted): map them
      // to the caller's indices, so that the pair is evaluated as the reference does, with i < j
      constexpr int F = Rec<G, D>::f32;
      const int ci = __float_as_int(P.rec32[(s.pt_off + q.i) * F + Rec<G, D>::id]);
      const int cj = __float_as_int(P.rec32[(s.pt_off + q.j) * F + Rec<G, D>::id]);
      const int64_t qi = ci < cj ? ci : cj, qj = ci < cj ? cj : ci;
      double A[S], Bv[S];
#pragma unroll
      for (int t = 0; t < S; ++t) {
        A[t] = P.rec64[(s.pt_off + qi) * S + t];
        Bv[t] = P.rec64[(s.pt_off + qj) * S + t];
      }
      double v[2];
      const int c = exact_cell<G, D, J, SP ? (J ? 4 : 3) : 0>(P, s, A, Bv, P.cells + s.tab_off,
                                                              P.dedges + s.dedge_off, P.vedges + s.vedge_off, v);
      if (c >= 0) {
        double mv[NM];
        if constexpr (SP1) {
          const double duL2 = v[0] * v[0];
          mv[0] = v[1];
          mv[1] = duL2;
          mv[2] = v[0] * v[1];
          mv[3] = v[0] * duL2;
        } else {
          mv[0] = v[0];
        }
        if (q.slice == home) {
#pragma unroll
          for (int t = 0; t < NM; ++t) atomicAdd(&hs[t * nh + c], mv[t]);
          atomicAdd(&hc[c], 1u);
        } else {
          const int64_t o = s.out_off + OSTR * c;
          atomicAdd(&P.counts[o], 1ull);
#pragma unroll
          for (int t = 0; t < NM; ++t) atomicAdd(&P.sums[o + (SP1 ? sp_row(t) : 0)], mv[t]);
        }
      }
      }
    }
    __syncthreads();
    const int64_t off = P.slices[home].out_off;
    for (int c = threadIdx.x; c < nh; c += blockDim.x) {
      if (hc[c]) {
        const int64_t o = off + OSTR * c;
        atomicAdd(&P.counts[o], (unsigned long long)hc[c]);
#pragma unroll
        for (int t = 0; t < NM; ++t) atomicAdd(&P.sums[o + (SP1 ? sp_row(t) : 0)], hs[t * nh + c]);
      }
    }
  }
}

// Fused outputs: binary64 re-binning of the queued pairs, every output evaluated exactly (the
// joint operator into the joint accumulators, each 1D operator into its own).  Like recheck, each
// workgroup takes a contiguous range of the queue and accumulates the entries of the range's
// first slice in LDS histograms flushed once (the 1D outputs share one count per distance bin);
// entries of other slices go straight to the global accumulators.
template <int G, int D, bool J>
__global__ void __launch_bounds__(256) recheck_multi(Params P) {
  extern __shared__ __align__(16) char smem_raw[];
  constexpr int S = Rec<G, D>::d64;
  constexpr int F = Rec<G, D>::f32;
  const int nh = J ? P.n_dist * P.n_val : 0;  // joint cells
  const int nd = P.n_dist, n1 = P.n1d;
  double *hs = (double *)smem_raw;            // joint sums [nh], then 1D sums [n1][nd]
  double *ms = hs + nh;
  unsigned *hc = (unsigned *)(ms + n1 * nd);  // joint counts [nh], then 1D counts [nd]
  unsigned *mc = hc + nh;
  for (int c = threadIdx.x; c < nh + nd; c += blockDim.x) hc[c] = 0u;
  for (int c = threadIdx.x; c < nh + n1 * nd; c += blockDim.x) hs[c] = 0.0;
  const unsigned long long n0 = *P.pending_count;
  const unsigned long long n = n0 < (unsigned long long)P.pending_cap ? n0 : (unsigned long long)P.pending_cap;
  const unsigned long long per = (n + gridDim.x - 1) / gridDim.x;
  const unsigned long long beg = (unsigned long long)blockIdx.x * per;
  const unsigned long long end = beg + per < n ? beg + per : n;
  __syncthreads();
  if (beg >= end) return;
  const int home = P.pending[beg].slice;
  for (unsigned long long k = beg + threadIdx.x; k < end; k += blockDim.x) {
    const Pending q = P.pending[k];
    const Slice s = P.slices[q.slice];
    const int ci = __float_as_int(P.rec32[(s.pt_off + q.i) * F + Rec<G, D>::id]);
    const int cj = __float_as_int(P.rec32[(s.pt_off + q.j) * F + Rec<G, D>::id]);
    const int64_t qi = ci < cj ? ci : cj, qj = ci < cj ? cj : ci;
    double A[S], Bv[S];
#pragma unroll
    for (int t = 0; t < S; ++t) {
      A[t] = P.rec64[(s.pt_off + qi) * S + t];
      Bv[t] = P.rec64[(s.pt_off + qj) * S + t];
    }
    const Cell *tab = P.cells + s.tab_off;
    const double *dE = P.dedges + s.dedge_off, *vE = P.vedges + s.vedge_off;
    const bool local = (q.slice == home);
    double v[2];
    if constexpr (J) {
      const int c = exact_cell<G, D, true, 0>(P, s, A, Bv, tab, dE, vE, v);
      if (c >= 0) {
        if (local) {
          atomicAdd(&hc[c], 1u);
          atomicAdd(&hs[c], v[0]);
        } else {
          atomicAdd(&P.counts[s.out_off + c], 1ull);
          atomicAdd(&P.sums[s.out_off + c], v[0]);
        }
      }
    }
    for (int o = 0; o < n1; ++o) {
      Params Q = P;
      Q.op = P.mop[o];
      Q.nl = P.mnl[o];
      Q.nt = P.mnt[o];
      Q.n_val = 0;
      const int c = exact_cell<G, D, false, 0>(Q, s, A, Bv, tab, dE, vE, v);
      if (c >= 0) {
        if (local) {
          if (o == 0) atomicAdd(&mc[c], 1u);  // the distance bin is the same for every output
          atomicAdd(&ms[o * nd + c], v[0]);
        } else {
          const int64_t off = ((int64_t)o * P.n_slices + q.slice) * nd + c;
          atomicAdd(&P.mcounts[off], 1ull);
          atomicAdd(&P.msums[off], v[0]);
        }
      }
    }
  }
  __syncthreads();
  const int64_t joff = P.slices[home].out_off;
  for (int c = threadIdx.x; c < nh; c += blockDim.x)
    if (hc[c]) {
      atomicAdd(&P.counts[joff + c], (unsigned long long)hc[c]);
      atomicAdd(&P.sums[joff + c], hs[c]);
    }
  for (int c = threadIdx.x; c < n1 * nd; c += blockDim.x) {
    const int o = c / nd, b = c - o * nd;
    if (mc[b]) {
      const int64_t off = ((int64_t)o * P.n_slices + home) * nd + b;
      atomicAdd(&P.mcounts[off], (unsigned long long)mc[b]);
      atomicAdd(&P.msums[off], ms[c]);
    }
  }
}

__global__ void merge_out(const double *src_s, const unsigned long long *src_c, double *dst_s,
                          unsigned long long *dst_c, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) {
    dst_s[k] += src_s[k];
    dst_c[k] += src_c[k];
  }
}

// merge unless the re-bin queue overflowed (then the accumulators are incomplete: flag it and
// leave the outputs untouched so the host can redo the call in binary64)
__global__ void merge_checked(const double *src_s, const unsigned long long *src_c, double *dst_s,
                              unsigned long long *dst_c, int64_t n,
                              const unsigned long long *pcount, int64_t cap, int *flag) {
  if (pcount && *pcount > (unsigned long long)cap) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *flag = 1;
    return;
  }
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) {
    dst_s[k] += src_s[k];
    dst_c[k] += src_c[k];
  }
}

// _sp1d_derive_rows! (serial_single_pass.jl:208-219) over nbins (6)-row columns: T2 = S2 - L2,
// L1T2 = S3 - L3, count rows 2..6 = row 1.  Assignments, so it is idempotent on accumulated output.
__global__ void derive_sp(double *s, unsigned long long *c, int64_t nbins) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbins) return;
  double *col = s + kSpRows * b;
  unsigned long long *cc = c + kSpRows * b;
  col[2] = col[0] - col[1];
  col[5] = col[3] - col[4];
  const unsigned long long n = cc[0];
  for (int t = 1; t < kSpRows; ++t) cc[t] = n;
}

__global__ void zero_u64(unsigned long long *p, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0ull;
}

__global__ void zero_f64(double *p, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0.0;
}

__global__ void eval_math(int fn, int64_t n, const double *a, const double *b, double *out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  switch (fn) {
    case 0: sf_sincosd(a[k], &out[2 * k], &out[2 * k + 1]); break;
    case 1: sf_sincos(a[k], &out[2 * k], &out[2 * k + 1]); break;
    case 2: out[k] = sf_atan2(a[k], b[k]); break;
    case 3: out[k] = __builtin_sqrt(a[k]); break;
    default: out[k] = a[k] / b[k]; break;
  }
}

__global__ void zero_i32(int *p, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0;
}

// Several buffers cleared by one launch (one row of blocks per buffer): a call's accumulators,
// counters and flags, one kernel instead of one per buffer.
struct ZeroSet {
  int *p[kZeroSetMax];
  int64_t n[kZeroSetMax];  // 32-bit words
};
__global__ void zero_multi(ZeroSet z) {
  const int b = blockIdx.y;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < z.n[b]) z.p[b][k] = 0;
}

}  // namespace sfk

// =============================================================================================
// launchers (called from sf_api.hip)
// =============================================================================================
namespace sfk {

struct PrepSliceHost;

template <int G, int D>
static hipError_t prep_impl(bool mixed, int geom_code, const void *ps, int n_slices, int64_t max_n,
                            double *rec64, float *rec32, hipStream_t st) {
  if (max_n <= 0) return hipSuccess;
  const int T = 256;
  dim3 grid((unsigned)((max_n + T - 1) / T), (unsigned)n_slices);
  prep64<G, D><<<grid, T, 0, st>>>((const PrepSlice *)ps, geom_code, rec64);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !mixed) return e;
  prep32<G, D><<<grid, T, 0, st>>>((const PrepSlice *)ps, rec64, rec32);
  return hipGetLastError();
}

hipError_t launch_prep(int G, int D, bool mixed, int geom_code, const void *ps, int n_slices,
                       int64_t max_n, double *rec64, float *rec32, hipStream_t st) {
  if (G == kFlat) {
    return D == 2 ? prep_impl<kFlat, 2>(mixed, geom_code, ps, n_slices, max_n, rec64, rec32, st)
                  : prep_impl<kFlat, 3>(mixed, geom_code, ps, n_slices, max_n, rec64, rec32, st);
  }
  return D == 2 ? prep_impl<kSphere, 2>(mixed, geom_code, ps, n_slices, max_n, rec64, rec32, st)
                : prep_impl<kSphere, 3>(mixed, geom_code, ps, n_slices, max_n, rec64, rec32, st);
}

template <int G, int D, bool J, bool SL, bool FX>
static const void *hist32_fn(int opk) {
  // fused outputs: scalar-loaded columns only
  if (opk == 9) return (const void *)&hist32<G, D, J, 9, true, FX>;
  if (opk == 10) return (const void *)&hist32<G, D, J, 10, true, FX>;
  if (opk == 11) return (const void *)&hist32<G, D, J, 11, true, FX>;
  if (opk == 12) return (const void *)&hist32<G, D, J, 12, true, FX>;
  if (opk == 13) return (const void *)&hist32<G, D, J, 13, true, FX>;
  if constexpr (!J) {
    if (opk == 3) return (const void *)&hist32<G, D, false, 3, SL, false>;
    if constexpr (SL) {  // 1D, scalar-loaded columns (the default): L^3, T, T^2, T^3
      if (opk == 5) return (const void *)&hist32<G, D, false, 5, true, false>;
      if (opk == 6) return (const void *)&hist32<G, D, false, 6, true, false>;
      if (opk == 7) return (const void *)&hist32<G, D, false, 7, true, false>;
      if (opk == 8) return (const void *)&hist32<G, D, false, 8, true, false>;
    }
  } else {
    if (opk == 4) return (const void *)&hist32<G, D, true, 4, SL, FX>;
  }
  if (opk == 1) return (const void *)&hist32<G, D, J, 1, SL, FX>;
  if (opk == 2) return (const void *)&hist32<G, D, J, 2, SL, FX>;
  return (const void *)&hist32<G, D, J, 0, SL, FX>;
}

template <int G, int D, bool J, bool SL>
static const void *hist32_fx(bool fx, int opk) {
  // fixed-point cells exist for joint histograms only
  if constexpr (J) {
    if (fx) return hist32_fn<G, D, true, SL, true>(opk);
  }
  return hist32_fn<G, D, J, SL, false>(opk);
}

template <int G, int D, bool J>
static const void *hist_fn(bool mixed, int opk, bool sload, bool fx) {
  if (mixed) return sload ? hist32_fx<G, D, J, true>(fx, opk) : hist32_fx<G, D, J, false>(fx, opk);
  if constexpr (!J) {
    if (opk == 3) return (const void *)&hist64<G, D, false, 3>;
  } else {
    if (opk == 4) return (const void *)&hist64<G, D, true, 4>;
  }
  if (opk == 1) return (const void *)&hist64<G, D, J, 1>;
  if (opk == 2) return (const void *)&hist64<G, D, J, 2>;
  return (const void *)&hist64<G, D, J, 0>;
}

static const void *pick_hist(int G, int D, bool J, bool mixed, int opk, bool sl, bool fx) {
  if (G == kFlat) {
    if (D == 2) return J ? hist_fn<kFlat, 2, true>(mixed, opk, sl, fx) : hist_fn<kFlat, 2, false>(mixed, opk, sl, fx);
    return J ? hist_fn<kFlat, 3, true>(mixed, opk, sl, fx) : hist_fn<kFlat, 3, false>(mixed, opk, sl, fx);
  }
  if (D == 2) return J ? hist_fn<kSphere, 2, true>(mixed, opk, sl, fx) : hist_fn<kSphere, 2, false>(mixed, opk, sl, fx);
  return J ? hist_fn<kSphere, 3, true>(mixed, opk, sl, fx) : hist_fn<kSphere, 3, false>(mixed, opk, sl, fx);
}

static int opk_of(const Params &P) {
  if (P.n1d > 0) return fused_opk(P.n1d);
  if (sp2d(P.op, P.n_val)) return 4;
  if (P.nmom > 1) return 3;
  if (P.op == 0 && P.nl == 1 && P.nt == 0) return 1;
  if (P.op == 0 && P.nl == 2 && P.nt == 0) return 2;
  if (P.op == 0 && P.nl == 3 && P.nt == 0) return 5;
  if (P.op == 0 && P.nl == 0 && P.nt == 1) return 6;
  if (P.op == 0 && P.nl == 0 && P.nt == 2) return 7;
  if (P.op == 0 && P.nl == 0 && P.nt == 3) return 8;
  return 0;
}

int hist_occupancy(int G, int D, bool J, bool mixed, int opk, bool sload, bool fx, int block,
                   size_t smem) {
  const void *f = pick_hist(G, D, J, mixed, opk, sload, fx);
  (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, block, smem) != hipSuccess) return 1;
  return nb;
}

int hist_opk(const Params &P) { return opk_of(P); }

template <int G, int D>
static const void *strip_fn(int opk) {
  if (opk == 1) return (const void *)&hist32s<G, D, 1>;
  if (opk == 2) return (const void *)&hist32s<G, D, 2>;
  return (const void *)&hist32s<G, D, 0>;
}

static const void *pick_strip(int G, int D, int opk) {
  if (G == kFlat) return D == 2 ? strip_fn<kFlat, 2>(opk) : strip_fn<kFlat, 3>(opk);
  return D == 2 ? strip_fn<kSphere, 2>(opk) : strip_fn<kSphere, 3>(opk);
}

int strip_occupancy(int G, int D, int opk, int block, size_t smem) {
  const void *f = pick_strip(G, D, opk);
  (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, block, smem) != hipSuccess) return 1;
  return nb;
}

hipError_t launch_hist(int G, int D, bool J, bool mixed, bool sload, const Params &P, int grid,
                       int block, size_t smem, hipStream_t st) {
  const void *f = (mixed && P.strip > 1) ? pick_strip(G, D, opk_of(P))
                                         : pick_hist(G, D, J, mixed, opk_of(P), sload, mixed && P.fx);
  (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  Params Pc = P;
  const void *rec = mixed ? (const void *)P.rec32 : (const void *)P.rec64;
  void *args[] = {&Pc, &rec};
  return hipLaunchKernel(f, dim3(grid), dim3(block), args, smem, st);
}

template <int G, int D>
static const void *recheck_fn(bool J, bool sp) {
  if (J) return sp ? (const void *)&recheck<G, D, true, true> : (const void *)&recheck<G, D, true, false>;
  return sp ? (const void *)&recheck<G, D, false, true> : (const void *)&recheck<G, D, false, false>;
}

template <int G, int D>
static const void *recheck_multi_fn(bool J) {
  return J ? (const void *)&recheck_multi<G, D, true> : (const void *)&recheck_multi<G, D, false>;
}

hipError_t launch_recheck(int G, int D, bool J, const Params &P, int grid, hipStream_t st) {
  if (P.n1d > 0) {
    const void *f = G == kFlat ? (D == 2 ? recheck_multi_fn<kFlat, 2>(J) : recheck_multi_fn<kFlat, 3>(J))
                               : (D == 2 ? recheck_multi_fn<kSphere, 2>(J) : recheck_multi_fn<kSphere, 3>(J));
    Params Pc = P;
    void *args[] = {&Pc};
    const size_t nh = J ? (size_t)P.n_dist * P.n_val : 0;
    const size_t smem = nh * 12 + (size_t)P.n1d * P.n_dist * 8 + (size_t)P.n_dist * 4 + 16;
    (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    return hipLaunchKernel(f, dim3(grid), dim3(256), args, smem, st);
  }
  const bool sp = P.nmom > 1 || sp2d(P.op, P.n_val);
  const void *f;
  if (G == kFlat) f = D == 2 ? recheck_fn<kFlat, 2>(J, sp) : recheck_fn<kFlat, 3>(J, sp);
  else f = D == 2 ? recheck_fn<kSphere, 2>(J, sp) : recheck_fn<kSphere, 3>(J, sp);
  Params Pc = P;
  void *args[] = {&Pc};
  const int nh = P.n_dist * (P.n_val > 0 ? P.n_val : 1);
  const size_t smem = (size_t)nh * (P.nmom > 1 ? 36 : 12) + 16;
  (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  return hipLaunchKernel(f, dim3(grid), dim3(256), args, smem, st);
}

hipError_t launch_merge(const double *src_s, const unsigned long long *src_c, double *dst_s,
                        unsigned long long *dst_c, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  merge_out<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(src_s, src_c, dst_s, dst_c, n);
  return hipGetLastError();
}

hipError_t launch_eval_math(int fn, int64_t n, const double *a, const double *b, double *out,
                            hipStream_t st) {
  eval_math<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(fn, n, a, b, out);
  return hipGetLastError();
}

hipError_t launch_merge_checked(const double *src_s, const unsigned long long *src_c, double *dst_s,
                                unsigned long long *dst_c, int64_t n,
                                const unsigned long long *pcount, int64_t cap, int *flag,
                                hipStream_t st) {
  if (n <= 0) return hipSuccess;
  merge_checked<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(src_s, src_c, dst_s, dst_c, n, pcount,
                                                             cap, flag);
  return hipGetLastError();
}

hipError_t launch_derive_sp(double *s, unsigned long long *c, int64_t nbins, hipStream_t st) {
  if (nbins <= 0) return hipSuccess;
  derive_sp<<<(unsigned)((nbins + 255) / 256), 256, 0, st>>>(s, c, nbins);
  return hipGetLastError();
}

hipError_t launch_zero(void *p, int64_t bytes, hipStream_t st) {
  const int64_t n = bytes / 4;
  if (n <= 0) return hipSuccess;
  zero_i32<<<(unsigned)((n + 255) / 256), 256, 0, st>>>((int *)p, n);
  return hipGetLastError();
}

hipError_t launch_zero_multi(int count, void *const *p, const int64_t *bytes, hipStream_t st) {
  if (count <= 0 || count > kZeroSetMax) return hipErrorInvalidValue;
  ZeroSet z{};
  int64_t nmax = 0;
  for (int b = 0; b < count; ++b) {
    z.p[b] = (int *)p[b];
    z.n[b] = bytes[b] / 4;
    nmax = nmax > z.n[b] ? nmax : z.n[b];
  }
  if (nmax <= 0) return hipSuccess;
  zero_multi<<<dim3((unsigned)((nmax + 255) / 256), (unsigned)count), 256, 0, st>>>(z);
  return hipGetLastError();
}

}  // namespace sfk
