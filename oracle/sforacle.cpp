// sforacle.cpp -- CPU oracle (test infrastructure only; see sforacle.h).
//
// A scalar restatement of the reference's SerialBackend semantics:
//   flat (Euclidean, D in {2,3}):   _pf_simd_pairs!        src/Calculations/serial.jl:68-101
//                                   _pf_2d_simd_pairs!     src/Calculations/serial_2d.jl:49-85
//   sphere (Haversine/SphericalAngle): calculate_structure_function_i!    serial.jl:228-267
//                                   calculate_structure_function_2d_i! serial_2d.jl:260-303
// with the per-pair pieces they call:
//   squared digitize plan           src/BinEdges.jl:439-469,515-550 (bin = #{k: e_k^2 < r^2})
//   digitize / searchsortedfirst    src/HelperFunctions.jl:77-84, src/BinEdges.jl:56
//   _sf_raw                         src/StructureFunctionTypes.jl:227-244
//   operator functors               src/StructureFunctionTypes.jl:58-151
//   pair_frame / geodesic_frame     src/HelperFunctions.jl:186-200,248-276
//   pair_delta (sphere)             src/HelperFunctions.jl:324-348
//   n-hat, transverse_norm2         src/HelperFunctions.jl:450-471,624-683
// Every binary64 operation is written in the reference's order (StaticArrays dot/cross are left
// folds starting from the first product); the build compiles this file with -ffp-contract=off so
// no multiply-add is fused, as Julia does not fuse without muladd/@fastmath.
#include "sforacle.h"

#include <math.h>
#include <string.h>

#include <thread>
#include <vector>

#include "../structurefunctions.jl_amd/csrc/sfmath.h"

namespace {

// Julia isless for Float64: NaN sorts last; -0.0 < +0.0.
inline bool jl_isless(double a, double b) {
  if (a != a) return false;
  if (b != b) return true;
  if (a < b) return true;
  if (a == b) return signbit(a) && !signbit(b);
  return false;
}

// searchsortedfirst(v, x) - 1 for a vector sorted by isless: number of v[k] with isless(v[k], x).
inline int64_t count_isless(const double *v, int64_t n, double x) {
  int64_t lo = 0, hi = n;  // answer in [lo, hi]
  while (lo < hi) {
    int64_t m = lo + (hi - lo) / 2;
    if (jl_isless(v[m], x)) lo = m + 1; else hi = m;
  }
  return lo;
}

// searchsortedfirst(::LinearBinEdges, x) (BinEdges.jl:114-131).  Julia's muladd is taken as a
// fused multiply-add; -f * inv_step is a separate product.
int64_t linear_ssf(const sfo_bin_rule &r, double x) {
  const double f = r.first_edge;
  if (x <= f) return 1;
  const double l = r.last_edge;
  const int64_t n = r.n;
  if (x > l) return n + 1;
  const double t = fma(x, r.inv_step, -f * r.inv_step);
  int64_t idx = (int64_t)floor(t) + 1;
  if (idx < 1) idx = 1;
  if (idx > n) idx = n;
  const double u = fma((double)(idx - 1), r.step_val, f);
  return u < x ? idx + 1 : idx;
}

// digitize(x, edges) = searchsortedfirst(edges, x) - 1 (HelperFunctions.jl:77-84) for a
// LinearBinEdges axis, or one inside InfPaddedBinEdges (BinEdges.jl:306-316: x <= -Inf -> 1,
// x > last -> length + 2, else the inner result + 1).  The reference rejects NaN
// (floor(Int, NaN) throws); here it lands past the last bin, i.e. it is not counted.
int64_t rule_digitize(const sfo_bin_rule &r, double x) {
  const int64_t nb = r.padded ? r.n + 1 : r.n - 1;
  if (x != x) return nb + 1;
  if (r.padded) {
    if (x <= -__builtin_inf()) return 0;
    if (x > r.last_edge) return r.n + 1;
    return linear_ssf(r, x);
  }
  return linear_ssf(r, x) - 1;
}

// distance bin of a flat pair on the point-field kernels: the squared-distance plan
// (BinEdges.jl:410-558): searchsortedfirst(sqedges, r^2) - 1, or with a LinearBinEdges axis
// (SquaredLinearPlan) the rule on sqrt(r^2)
inline int64_t flat_point_bin(const sfo_problem *p, const double *sqedges, double r2) {
  if (p->dist_rule) return rule_digitize(*p->dist_rule, sqrt(r2));
  return count_isless(sqedges, p->n_dist + 1, r2);
}

// digitize(d, distance_bins) of a distance d (flat batch-leading r, or the spherical distance)
inline int64_t dist_bin(const sfo_problem *p, double d) {
  if (p->dist_rule) return rule_digitize(*p->dist_rule, d);
  return count_isless(p->dist_edges, p->n_dist + 1, d);
}

// digitize(v, value_bins)
inline int64_t val_bin(const sfo_problem *p, double v) {
  if (p->val_rule) return rule_digitize(*p->val_rule, v);
  return count_isless(p->val_edges, p->n_val + 1, v);
}

struct Op {
  int op, nl, nt, D;
};

// norm2 / _sum_abs2: 0 + x1*x1 + x2*x2 (+ x3*x3)  (StructureFunctionTypes.jl:23-31,
// HelperFunctions.jl:533-539)
inline double sum_abs2(const double *v, int D) {
  double out = 0.0;
  for (int k = 0; k < D; ++k) out += v[k] * v[k];
  return out;
}

// StaticArrays dot: a1*b1 + a2*b2 (+ a3*b3), left fold from the first product
inline double sdot(const double *a, const double *b, int D) {
  double s = a[0] * b[0];
  for (int k = 1; k < D; ++k) s = s + a[k] * b[k];
  return s;
}

// mdu_t(du, rhat) = dot(du, nhat(rhat))  (HelperFunctions.jl:450-471,640-643)
inline double mdu_t(const double *du, const double *rh, int D) {
  if (D == 2) {
    const double n[2] = {-rh[1], rh[0]};
    return sdot(du, n, 2);
  }
  // cross(zhat, rhat) with zhat = (0,0,1): (0*r3 - 1*r2, 1*r1 - 0*r3, 0*r2 - 0*r1)
  const double c[3] = {0.0 * rh[2] - 1.0 * rh[1], 1.0 * rh[0] - 0.0 * rh[2],
                       0.0 * rh[1] - 0.0 * rh[0]};
  const double nrm = sqrt(sdot(c, c, 3));
  const double n[3] = {c[0] / nrm, c[1] / nrm, c[2] / nrm};
  return sdot(du, n, 3);
}

// literal_pow x^n for n = 1, 2, 3 ((x*x)*x)
inline double ipow(double x, int n) {
  if (n == 1) return x;
  if (n == 2) return x * x;
  return x * x * x;
}

// sf(du, rhat) for every supported operator (StructureFunctionTypes.jl:58-151)
inline double op_apply(const Op &o, const double *du, const double *rh) {
  const int D = o.D;
  switch (o.op) {
    case SFO_OP_PROJECTED: {
      double ex = 1.0;
      if (o.nl != 0) ex = ex * ipow(sdot(du, rh, D), o.nl);
      if (o.nt != 0) {
        if (o.nt == 2) {
          const double L = sdot(du, rh, D);
          ex = ex * (sum_abs2(du, D) - L * L);
        } else {
          ex = ex * ipow(mdu_t(du, rh, D), o.nt);
        }
      }
      return ex;
    }
    case SFO_OP_S2:
      return sum_abs2(du, D);
    case SFO_OP_S3:
      return sdot(du, rh, D) * sum_abs2(du, D);
    case SFO_OP_T2COMP: {
      const double L = sdot(du, rh, D);
      return (sum_abs2(du, D) - L * L) / (double)(D - 1);
    }
    case SFO_OP_L1T2COMP: {
      const double L = sdot(du, rh, D);
      const double L2 = sdot(du, rh, D);
      return L * ((sum_abs2(du, D) - L2 * L2) / (double)(D - 1));
    }
    case SFO_OP_FULLVEC: {
      if (o.nl == 2) return sum_abs2(du, D);
      const double nrm = sqrt(sum_abs2(du, D));
      return ipow(nrm, o.nl);
    }
  }
  return 0.0;
}

// _sf_raw(sf, du, dx, r2) on the flat path (StructureFunctionTypes.jl:227-244)
inline double sf_raw(const Op &o, const double *du, const double *dx, double r2) {
  const int D = o.D;
  if (o.op == SFO_OP_S2) return sum_abs2(du, D);
  if (o.op == SFO_OP_PROJECTED && o.nl == 2 && o.nt == 0) {
    const double p = sdot(du, dx, D);
    return p * p / r2;
  }
  if (o.op == SFO_OP_PROJECTED && o.nl == 0 && o.nt == 2) {
    const double p = sdot(du, dx, D);
    return sum_abs2(du, D) - p * p / r2;
  }
  if (o.op == SFO_OP_T2COMP) {
    const double p = sdot(du, dx, D);
    return (sum_abs2(du, D) - p * p / r2) / (double)(D - 1);
  }
  const double s = sqrt(r2);
  double rh[3];
  for (int k = 0; k < D; ++k) rh[k] = dx[k] / s;
  return op_apply(o, du, rh);
}

struct Acc {
  std::vector<double> s, c, a, g;
  std::vector<uint64_t> n;
  bool comp;
  explicit Acc(size_t nb, bool compensated)
      : s(nb, 0.0), c(nb, 0.0), a(nb, 0.0), g(nb, 0.0), n(nb, 0), comp(compensated) {}
  inline void add(size_t b, double v, double mag) {
    g[b] += mag;
    if (comp) {
      // Neumaier
      const double t = s[b] + v;
      if (fabs(s[b]) >= fabs(v)) c[b] += (s[b] - t) + v; else c[b] += (v - t) + s[b];
      s[b] = t;
    } else {
      s[b] += v;
    }
    a[b] += fabs(v);
    n[b] += 1;
  }
};

struct Prob {
  const sfo_problem *p;
  Op o;
  int order;
  std::vector<double> sqedges;  // e_k * e_k (correctly rounded square, BinEdges.jl:440)
  // per-point sphere frame, hoisted: unit position, local east, local north
  std::vector<double> P, E, Nn;
};

// ---------------------------------------------------------------------------------------------
// spherical pair frame (HelperFunctions.jl:186-200, 265-276) and increments (:324-348)
// ---------------------------------------------------------------------------------------------
inline void sphere_point(int geom, double lon, double lat, double *P, double *E, double *N) {
  double sl, cl, sp, cp;
  if (geom == SFO_GEOM_SPHERE_DEG) {
    sf_sincosd(lon, &sl, &cl);
    sf_sincosd(lat, &sp, &cp);
  } else {
    sf_sincos(lon, &sl, &cl);
    sf_sincos(lat, &sp, &cp);
  }
  // _unit_position(sl, cl, sp, cp) = (cp*cl, cp*sl, sp)
  P[0] = cp * cl; P[1] = cp * sl; P[2] = sp;
  // _local_east_north: E = (-sl, cl, 0), N = (-sp*cl, -sp*sl, cp)
  E[0] = -sl; E[1] = cl; E[2] = 0.0;
  N[0] = -sp * cl; N[1] = -sp * sl; N[2] = cp;
}

struct SpherePair {
  bool ok;
  double distance;
  double tA[2], tB[2];
};

inline SpherePair sphere_frame(const double *p, const double *q, const double *EA, const double *NA,
                               const double *EB, const double *NB, double radius) {
  SpherePair r;
  const double d[3] = {q[0] - p[0], q[1] - p[1], q[2] - p[2]};
  const double s[3] = {p[0] + q[0], p[1] + q[1], p[2] + q[2]};
  const double dp = sdot(d, p, 3);
  const double t_A[3] = {d[0] - dp * p[0], d[1] - dp * p[1], d[2] - dp * p[2]};
  const double t_B[3] = {dp * q[0] + d[0], dp * q[1] + d[1], dp * q[2] + d[2]};
  // cross(p, d)
  const double w[3] = {p[1] * d[2] - p[2] * d[1], p[2] * d[0] - p[0] * d[2],
                       p[0] * d[1] - p[1] * d[0]};
  const double s2 = sdot(w, w, 3);
  const double sigma = 2.0 * sf_atan2(sqrt(sdot(d, d, 3)), sqrt(sdot(s, s, 3)));
  r.ok = s2 > 2.2250738585072014e-308;  // floatmin(Float64)
  const double inv_s = r.ok ? 1.0 / sqrt(s2) : 0.0;
  const double ta[3] = {t_A[0] * inv_s, t_A[1] * inv_s, t_A[2] * inv_s};
  const double tb[3] = {t_B[0] * inv_s, t_B[1] * inv_s, t_B[2] * inv_s};
  r.tA[0] = sdot(ta, EA, 3); r.tA[1] = sdot(ta, NA, 3);
  r.tB[0] = sdot(tb, EB, 3); r.tB[1] = sdot(tb, NB, 3);
  r.distance = radius * sigma;
  return r;
}

// pair_delta(::SphericalGeometry{D}) -> du in the geodesic frame; rhat = e1
inline void sphere_delta(const SpherePair &f, const double *uA, const double *uB, int D, double *du) {
  const double nA[2] = {-f.tA[1], f.tA[0]};
  const double nB[2] = {-f.tB[1], f.tB[0]};
  du[0] = sdot(uB, f.tB, 2) - sdot(uA, f.tA, 2);
  du[1] = sdot(uB, nB, 2) - sdot(uA, nA, 2);
  if (D == 3) du[2] = uB[2] - uA[2];
}

// ||du||^order: the magnitude scale of one pair's value (for error bounds in tests)
inline double mag_of(const double *du, int D, int order) {
  const double n = sqrt(sum_abs2(du, D));
  double m = 1.0;
  for (int k = 0; k < order; ++k) m *= n;
  return m;
}

int validate(const sfo_problem *p) {
  if (!p) return -1;
  if (p->D != 2 && p->D != 3) return -2;
  if (p->geometry == SFO_GEOM_FLAT) {
    if (p->W != p->D) return -3;
  } else if (p->geometry == SFO_GEOM_SPHERE_DEG || p->geometry == SFO_GEOM_SPHERE_RAD) {
    if (p->W != 2) return -3;
    if (!(p->radius > 0.0) || !(p->radius < __builtin_inf())) return -4;
  } else {
    return -5;
  }
  if (p->N < 0) return -6;
  if (p->N > 0 && (!p->x || !p->u)) return -6;
  if (p->op < 0 || p->op > SFO_OP_FULLVEC) return -7;
  if (p->op == SFO_OP_PROJECTED && (p->nl < 0 || p->nt < 0 || p->nl + p->nt > 3)) return -7;
  if (p->op == SFO_OP_FULLVEC && (p->nl < 1 || p->nl > 3)) return -7;
  if (p->n_dist < 1 || !p->dist_edges) return -8;
  for (int k = 0; k <= p->n_dist; ++k) {
    if (p->dist_edges[k] != p->dist_edges[k]) return -8;
    if (k > 0 && jl_isless(p->dist_edges[k], p->dist_edges[k - 1])) return -8;
  }
  // flat: squared plan needs non-negative edges; a leading -Inf is InfPaddedBinEdges' catch-all
  // underflow bin (SquaredInfPaddedPlan, BinEdges.jl:434-437,520,553-558)
  if (p->geometry == SFO_GEOM_FLAT)
    for (int k = 0; k <= p->n_dist; ++k)
      if (p->dist_edges[k] < 0.0 && !(k == 0 && p->dist_edges[0] == -__builtin_inf())) return -8;
  // a LinearBinEdges rule must give the axis' bin count
  if (p->dist_rule && (p->dist_rule->n < 2 ||
                       p->n_dist != (p->dist_rule->padded ? p->dist_rule->n + 1 : p->dist_rule->n - 1)))
    return -8;
  if (p->val_rule && (p->val_rule->n < 2 ||
                      p->n_val != (p->val_rule->padded ? p->val_rule->n + 1 : p->val_rule->n - 1)))
    return -9;
  if (p->n_val < 0) return -9;
  if (p->n_val > 0) {
    if (!p->val_edges) return -9;
    for (int k = 0; k <= p->n_val; ++k) {
      if (p->val_edges[k] != p->val_edges[k]) return -9;
      if (k > 0 && jl_isless(p->val_edges[k], p->val_edges[k - 1])) return -9;
    }
  }
  return 0;
}

// One outer row i: pairs (i, j > i) accumulated into acc.
void row(const Prob &P, int64_t i, Acc &acc) {
  const sfo_problem *p = P.p;
  const int D = p->D, W = p->W;
  const int64_t N = p->N;
  const int nb = p->n_dist;
  const int nv = p->n_val;
  if (p->geometry == SFO_GEOM_FLAT) {
    double Xi[3], Ui[3];
    for (int k = 0; k < D; ++k) { Xi[k] = p->x[k + W * i]; Ui[k] = p->u[k + D * i]; }
    for (int64_t j = i + 1; j < N; ++j) {
      double dx[3] = {0.0, 0.0, 0.0}, du[3] = {0.0, 0.0, 0.0};
      for (int k = 0; k < D; ++k) {
        dx[k] = p->x[k + W * j] - Xi[k];
        du[k] = p->u[k + D * j] - Ui[k];
      }
      const double r2 = sdot(dx, dx, D);
      double val;
      int64_t b;
      if (p->batch_leading) {
        // pair_frame / pair_direction / pair_delta (HelperFunctions.jl:248-330), as
        // _bl_shared_1d! and _bl_joint2d_shared! use them (batch_leading.jl:105-215)
        const double r = sqrt(r2);
        double rh[3] = {0.0, 0.0, 0.0};
        for (int k = 0; k < D; ++k) rh[k] = dx[k] / r;
        val = op_apply(P.o, du, rh);
        b = dist_bin(p, r);
      } else {
        val = sf_raw(P.o, du, dx, r2);
        b = flat_point_bin(p, P.sqedges.data(), r2);
      }
      if (1 <= b && b <= nb) {
        const double mag = mag_of(du, D, P.order);
        if (nv == 0) {
          acc.add((size_t)(b - 1), val, mag);
        } else {
          const int64_t vb = val_bin(p, val);
          if (1 <= vb && vb <= nv) acc.add((size_t)((b - 1) + (int64_t)nb * (vb - 1)), val, mag);
        }
      }
    }
    return;
  }
  // sphere: scalar per-i kernel
  const double *pA = &P.P[3 * i], *EA = &P.E[3 * i], *NA = &P.Nn[3 * i];
  const double *uA = &p->u[D * i];
  const double rh[3] = {1.0, 0.0, 0.0};  // pair_direction(::SphericalGeometry{D}) = e1
  for (int64_t j = i + 1; j < N; ++j) {
    const SpherePair f = sphere_frame(pA, &P.P[3 * j], EA, NA, &P.E[3 * j], &P.Nn[3 * j], p->radius);
    const int64_t b = dist_bin(p, f.distance);
    if (f.ok && 1 <= b && b < nb + 1) {
      double du[3];
      sphere_delta(f, uA, &p->u[D * j], D, du);
      const double val = op_apply(P.o, du, rh);
      const double mag = mag_of(du, D, P.order);
      if (nv == 0) {
        acc.add((size_t)(b - 1), val, mag);
      } else {
        const int64_t vb = val_bin(p, val);
        if (1 <= vb && vb < nv + 1) acc.add((size_t)((b - 1) + (int64_t)nb * (vb - 1)), val, mag);
      }
    }
  }
}

void prepare(Prob &P) {
  const sfo_problem *p = P.p;
  P.o = Op{p->op, p->nl, p->nt, p->D};
  switch (p->op) {
    case SFO_OP_PROJECTED: P.order = p->nl + p->nt; break;
    case SFO_OP_S2: case SFO_OP_T2COMP: P.order = 2; break;
    case SFO_OP_S3: case SFO_OP_L1T2COMP: P.order = 3; break;
    default: P.order = p->nl; break;
  }
  P.sqedges.resize(p->n_dist + 1);
  for (int k = 0; k <= p->n_dist; ++k) {
    const double e = p->dist_edges[k];
    P.sqedges[k] = (e == -__builtin_inf()) ? e : e * e;  // -Inf stays below every r^2
  }
  if (p->geometry != SFO_GEOM_FLAT) {
    P.P.resize(3 * p->N); P.E.resize(3 * p->N); P.Nn.resize(3 * p->N);
    for (int64_t i = 0; i < p->N; ++i)
      sphere_point(p->geometry, p->x[2 * i], p->x[2 * i + 1], &P.P[3 * i], &P.E[3 * i], &P.Nn[3 * i]);
  }
}

}  // namespace

extern "C" int sfo_hist(const sfo_problem *p, int nthreads, int compensated, int64_t i_begin,
                        int64_t i_end, int64_t i_step, double *sums, uint64_t *counts,
                        double *abs_sums, double *mag_sums) {
  const int rc = validate(p);
  if (rc) return rc;
  if (!sums || !counts) return -10;
  if (i_step < 1) return -11;
  if (i_begin < 0) i_begin = 0;
  if (i_end > p->N) i_end = p->N;
  Prob P;
  P.p = p;
  prepare(P);
  const size_t ncell = (size_t)p->n_dist * (size_t)(p->n_val > 0 ? p->n_val : 1);
  if (nthreads < 1) nthreads = 1;
  std::vector<Acc> accs;
  accs.reserve(nthreads);
  for (int t = 0; t < nthreads; ++t) accs.emplace_back(ncell, compensated != 0);
  auto work = [&](int t) {
    int64_t k = 0;
    for (int64_t i = i_begin; i < i_end; i += i_step, ++k) {
      if ((int)(k % nthreads) == t) row(P, i, accs[t]);
    }
  };
  if (nthreads == 1) {
    work(0);
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t) th.emplace_back(work, t);
    for (auto &x : th) x.join();
  }
  for (size_t b = 0; b < ncell; ++b) {
    // combine per-thread partials (value + compensation), then add into caller buffers
    double s = 0.0, c = 0.0, a = 0.0, g = 0.0;
    uint64_t n = 0;
    for (int t = 0; t < nthreads; ++t) {
      const double parts[2] = {accs[t].s[b], accs[t].c[b]};
      for (double v : parts) {
        const double tt = s + v;
        if (fabs(s) >= fabs(v)) c += (s - tt) + v; else c += (v - tt) + s;
        s = tt;
      }
      a += accs[t].a[b];
      g += accs[t].g[b];
      n += accs[t].n[b];
    }
    sums[b] += s + c;
    counts[b] += n;
    if (abs_sums) abs_sums[b] += a;
    if (mag_sums) mag_sums[b] += g;
  }
  return 0;
}

extern "C" int64_t sfo_rule_digitize(const sfo_bin_rule *r, double x) { return rule_digitize(*r, x); }

extern "C" void sfo_sincosd(double x, double *s, double *c) { sf_sincosd(x, s, c); }
extern "C" void sfo_sincos(double x, double *s, double *c) { sf_sincos(x, s, c); }
extern "C" double sfo_atan2(double y, double x) { return sf_atan2(y, x); }

extern "C" int sfo_pair(const sfo_problem *p, int64_t i, int64_t j, double *distance, double *value,
                        int64_t *dbin, int64_t *vbin) {
  if (validate(p)) return -1;
  if (i < 0 || j < 0 || i >= p->N || j >= p->N) return -1;
  const int D = p->D, W = p->W;
  Op o{p->op, p->nl, p->nt, D};
  *vbin = -1;
  if (p->geometry == SFO_GEOM_FLAT) {
    double dx[3] = {0.0, 0.0, 0.0}, du[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < D; ++k) {
      dx[k] = p->x[k + W * j] - p->x[k + W * i];
      du[k] = p->u[k + D * j] - p->u[k + D * i];
    }
    const double r2 = sdot(dx, dx, D);
    *distance = sqrt(r2);
    if (p->batch_leading) {
      double rh[3] = {0.0, 0.0, 0.0};
      for (int k = 0; k < D; ++k) rh[k] = dx[k] / *distance;
      *value = op_apply(o, du, rh);
      *dbin = dist_bin(p, *distance);
    } else {
      *value = sf_raw(o, du, dx, r2);
      std::vector<double> sq(p->n_dist + 1);
      for (int k = 0; k <= p->n_dist; ++k) {
        const double e = p->dist_edges[k];
        sq[k] = (e == -__builtin_inf()) ? e : e * e;
      }
      *dbin = flat_point_bin(p, sq.data(), r2);
    }
    if (p->n_val > 0) *vbin = val_bin(p, *value);
    return 1;
  }
  double Pa[3], Ea[3], Na[3], Pb[3], Eb[3], Nb[3];
  sphere_point(p->geometry, p->x[2 * i], p->x[2 * i + 1], Pa, Ea, Na);
  sphere_point(p->geometry, p->x[2 * j], p->x[2 * j + 1], Pb, Eb, Nb);
  const SpherePair f = sphere_frame(Pa, Pb, Ea, Na, Eb, Nb, p->radius);
  double du[3];
  sphere_delta(f, &p->u[D * i], &p->u[D * j], D, du);
  const double rh[3] = {1.0, 0.0, 0.0};
  *distance = f.distance;
  *value = op_apply(o, du, rh);
  *dbin = dist_bin(p, f.distance);
  if (p->n_val > 0) *vbin = val_bin(p, *value);
  return f.ok ? 1 : 0;
}

// ---------------------------------------------------------------------------------------------
// Single pass: the six invariants S2, L2, T2, S3, L3, L1T2 in one pair pass
// (calculate_structure_functions_single_pass, src/Calculations/serial_single_pass.jl:602-700).
// Flat D in {2,3}: _pf_sp_simd_pairs! (:231-271): inv_r = inv(sqrt(r2)), du_L = dot(du,dx)*inv_r,
// |du|^2 = dot(du,du); sphere: _sp1d_pairs! (:315-345): du_L = dot(du, rhat), |du|^2 = dot(du,du).
// Rows 1,2,4,5 (S2, L2, S3, L3) and count row 1 are accumulated; then, over the whole buffer,
// T2 = S2 - L2, L1T2 = S3 - L3 and count rows 2..6 = count row 1 (_sp1d_derive_rows!, :208-219).
// Output layout (6, n_dist) column-major.
// ---------------------------------------------------------------------------------------------
extern "C" int sfo_hist_single_pass(const sfo_problem *p, int nthreads, int64_t i_begin,
                                    int64_t i_end, int64_t i_step, double *sums,
                                    uint64_t *counts) {
  sfo_problem q = *p;
  q.op = SFO_OP_S2;
  q.nl = q.nt = 0;
  q.n_val = 0;
  q.val_edges = nullptr;
  const int rc = validate(&q);
  if (rc) return rc;
  if (!sums || !counts || i_step < 1) return -10;
  if (i_begin < 0) i_begin = 0;
  if (i_end > q.N) i_end = q.N;
  Prob P;
  P.p = &q;
  prepare(P);
  const int nb = q.n_dist, D = q.D, W = q.W;
  if (nthreads < 1) nthreads = 1;
  // per thread: 4 moments (compensated) + count per bin
  struct SP { std::vector<double> s, c; std::vector<uint64_t> n; };
  std::vector<SP> acc(nthreads);
  for (auto &a : acc) { a.s.assign(4 * nb, 0.0); a.c.assign(4 * nb, 0.0); a.n.assign(nb, 0); }
  auto add = [](SP &a, int b, int m, double v) {
    double &s = a.s[4 * b + m], &c = a.c[4 * b + m];
    const double t = s + v;
    if (fabs(s) >= fabs(v)) c += (s - t) + v; else c += (v - t) + s;
    s = t;
  };
  auto work = [&](int t) {
    SP &a = acc[t];
    int64_t k = 0;
    for (int64_t i = i_begin; i < i_end; i += i_step, ++k) {
      if ((int)(k % nthreads) != t) continue;
      for (int64_t j = i + 1; j < q.N; ++j) {
        double duL, dn2;
        int64_t b;
        if (q.geometry == SFO_GEOM_FLAT) {
          double dx[3] = {0, 0, 0}, du[3] = {0, 0, 0};
          for (int d = 0; d < D; ++d) {
            dx[d] = q.x[d + W * j] - q.x[d + W * i];
            du[d] = q.u[d + D * j] - q.u[d + D * i];
          }
          const double r2 = sdot(dx, dx, D);
          if (q.batch_leading) {
            // _bl_sp1d_shared! / _bl_sp2d_shared! (batch_leading.jl:270-330): du_L = dot(du, rhat)
            const double r = sqrt(r2);
            double rh[3] = {0.0, 0.0, 0.0};
            for (int d = 0; d < D; ++d) rh[d] = dx[d] / r;
            duL = sdot(du, rh, D);
            b = dist_bin(&q, r);
          } else {
            const double inv_r = 1.0 / sqrt(r2);
            duL = sdot(du, dx, D) * inv_r;
            b = flat_point_bin(&q, P.sqedges.data(), r2);
          }
          dn2 = sdot(du, du, D);
          if (!(1 <= b && b <= nb)) continue;
        } else {
          const SpherePair f = sphere_frame(&P.P[3 * i], &P.P[3 * j], &P.E[3 * i], &P.Nn[3 * i],
                                            &P.E[3 * j], &P.Nn[3 * j], q.radius);
          b = dist_bin(&q, f.distance);
          if (!(f.ok && 1 <= b && b <= nb)) continue;
          double du[3];
          sphere_delta(f, &q.u[D * i], &q.u[D * j], D, du);
          const double rh[3] = {1.0, 0.0, 0.0};
          duL = sdot(du, rh, D);
          dn2 = sdot(du, du, D);
        }
        const double duL2 = duL * duL;
        add(a, (int)(b - 1), 0, dn2);
        add(a, (int)(b - 1), 1, duL2);
        add(a, (int)(b - 1), 2, duL * dn2);
        add(a, (int)(b - 1), 3, duL * duL2);
        a.n[b - 1] += 1;
      }
    }
  };
  if (nthreads == 1) {
    work(0);
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t) th.emplace_back(work, t);
    for (auto &x : th) x.join();
  }
  for (int b = 0; b < nb; ++b) {
    double m[4];
    uint64_t n = 0;
    for (int mm = 0; mm < 4; ++mm) {
      double s = 0.0, c = 0.0;
      for (int t = 0; t < nthreads; ++t) {
        const double parts[2] = {acc[t].s[4 * b + mm], acc[t].c[4 * b + mm]};
        for (double v : parts) {
          const double tt = s + v;
          if (fabs(s) >= fabs(v)) c += (s - tt) + v; else c += (v - tt) + s;
          s = tt;
        }
      }
      m[mm] = s + c;
    }
    for (int t = 0; t < nthreads; ++t) n += acc[t].n[b];
    sums[0 + 6 * b] += m[0];
    sums[1 + 6 * b] += m[1];
    sums[3 + 6 * b] += m[2];
    sums[4 + 6 * b] += m[3];
    counts[6 * b] += n;
  }
  // _sp1d_derive_rows! (serial_single_pass.jl:208-219): assignments over the whole buffer
  for (int b = 0; b < nb; ++b) {
    sums[2 + 6 * b] = sums[0 + 6 * b] - sums[1 + 6 * b];
    sums[5 + 6 * b] = sums[3 + 6 * b] - sums[4 + 6 * b];
    for (int r = 1; r < 6; ++r) counts[r + 6 * b] = counts[6 * b];
  }
  return 0;
}

// ---------------------------------------------------------------------------------------------
// Single pass, 2D (joint) form: the six invariants each binned by (distance, own value)
// (calculate_structure_functions_single_pass_2d, src/Calculations/serial_single_pass.jl:1161-1181;
// pair loops _sp2d_simd_pairs! :873-916 for flat, _sp2d_pairs! :820-850 for the sphere).  Per
// counted pair: du_L and |du|^2 as in the 1D form, du_L2 = du_L^2, du_T2 = |du|^2 - du_L2, values
// (|du|^2, du_L2, du_T2, du_L |du|^2, du_L du_L2, du_L du_T2); value t goes to bin
// digitize(value, vb_t) when 1 <= bin <= n_val (_sp2d_scatter!, :852-864).  val_edges6 holds the
// six edge vectors back to back, n_val + 1 each.  Output (6, n_dist, n_val) column-major.
// ---------------------------------------------------------------------------------------------
extern "C" int sfo_hist_single_pass_2d(const sfo_problem *p, int nthreads, int64_t i_begin,
                                       int64_t i_end, int64_t i_step, const double *val_edges6,
                                       double *sums, uint64_t *counts, double *abs_sums,
                                       double *mag_sums) {
  if (!p || p->n_val < 1 || !val_edges6) return -10;
  sfo_problem q = *p;
  const int nv = q.n_val;
  q.op = SFO_OP_S2;
  q.nl = q.nt = 0;
  q.n_val = 0;
  q.val_edges = nullptr;
  const int rc = validate(&q);
  if (rc) return rc;
  for (int t = 0; t < 6; ++t) {
    sfo_problem qt = *p;
    qt.op = SFO_OP_S2;
    qt.nl = qt.nt = 0;
    qt.val_edges = val_edges6 + (size_t)t * (nv + 1);
    const int rt = validate(&qt);
    if (rt) return rt;
  }
  if (!sums || !counts || i_step < 1) return -10;
  if (i_begin < 0) i_begin = 0;
  if (i_end > q.N) i_end = q.N;
  Prob P;
  P.p = &q;
  prepare(P);
  const int nb = q.n_dist, D = q.D, W = q.W;
  const int64_t ncell = (int64_t)6 * nb * nv;
  if (nthreads < 1) nthreads = 1;
  struct Acc { std::vector<double> s, c, a, g; std::vector<uint64_t> n; };
  std::vector<Acc> acc(nthreads);
  for (auto &a : acc) {
    a.s.assign(ncell, 0.0); a.c.assign(ncell, 0.0); a.n.assign(ncell, 0);
    a.a.assign(ncell, 0.0); a.g.assign(ncell, 0.0);
  }
  auto work = [&](int th) {
    Acc &a = acc[th];
    int64_t k = 0;
    for (int64_t i = i_begin; i < i_end; i += i_step, ++k) {
      if ((int)(k % nthreads) != th) continue;
      for (int64_t j = i + 1; j < q.N; ++j) {
        double duL, dn2, scale;
        int64_t b;
        if (q.geometry == SFO_GEOM_FLAT) {
          double dx[3] = {0, 0, 0}, du[3] = {0, 0, 0};
          for (int d = 0; d < D; ++d) {
            dx[d] = q.x[d + W * j] - q.x[d + W * i];
            du[d] = q.u[d + D * j] - q.u[d + D * i];
          }
          const double r2 = sdot(dx, dx, D);
          if (q.batch_leading) {
            const double r = sqrt(r2);
            double rh[3] = {0.0, 0.0, 0.0};
            for (int d = 0; d < D; ++d) rh[d] = dx[d] / r;
            duL = sdot(du, rh, D);
            b = dist_bin(&q, r);
          } else {
            const double inv_r = 1.0 / sqrt(r2);
            duL = sdot(du, dx, D) * inv_r;
            b = flat_point_bin(&q, P.sqedges.data(), r2);
          }
          dn2 = sdot(du, du, D);
          scale = sqrt(dn2);
          if (!(1 <= b && b <= nb)) continue;
        } else {
          const SpherePair f = sphere_frame(&P.P[3 * i], &P.P[3 * j], &P.E[3 * i], &P.Nn[3 * i],
                                            &P.E[3 * j], &P.Nn[3 * j], q.radius);
          b = dist_bin(&q, f.distance);
          if (!(f.ok && 1 <= b && b <= nb)) continue;
          double du[3];
          sphere_delta(f, &q.u[D * i], &q.u[D * j], D, du);
          const double rh[3] = {1.0, 0.0, 0.0};
          duL = sdot(du, rh, D);
          dn2 = sdot(du, du, D);
          // the frame transport makes binary32 errors absolute in the velocities' size, growing
          // like 1 + c^2 / sin(sigma) towards antipodal pairs
          const double sig = f.distance / q.radius, c2 = 4.0 * sin(0.5 * sig) * sin(0.5 * sig);
          scale = (sqrt(sdot(&q.u[D * i], &q.u[D * i], D)) + sqrt(sdot(&q.u[D * j], &q.u[D * j], D))) *
                  (1.0 + c2 / fmax(sin(sig), 1e-300));
        }
        const double duL2 = duL * duL;
        const double duT2 = dn2 - duL2;
        const double vals[6] = {dn2, duL2, duT2, duL * dn2, duL * duL2, duL * duT2};
        for (int t = 0; t < 6; ++t) {
          const int64_t vb = count_isless(val_edges6 + (size_t)t * (nv + 1), nv + 1, vals[t]);
          if (!(1 <= vb && vb <= nv)) continue;
          const int64_t cell = t + 6 * ((b - 1) + (int64_t)nb * (vb - 1));
          double &sm = a.s[cell], &cm = a.c[cell];
          const double v = vals[t], tt = sm + v;
          if (fabs(sm) >= fabs(v)) cm += (sm - tt) + v; else cm += (v - tt) + sm;
          sm = tt;
          a.n[cell] += 1;
          a.a[cell] += fabs(v);
          a.g[cell] += t < 3 ? scale * scale : scale * scale * scale;  // the error scale^order
        }
      }
    }
  };
  if (nthreads == 1) {
    work(0);
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t) th.emplace_back(work, t);
    for (auto &x : th) x.join();
  }
  for (int64_t c = 0; c < ncell; ++c) {
    double sm = 0.0, cm = 0.0;
    uint64_t n = 0;
    for (int t = 0; t < nthreads; ++t) {
      const double parts[2] = {acc[t].s[c], acc[t].c[c]};
      for (double v : parts) {
        const double tt = sm + v;
        if (fabs(sm) >= fabs(v)) cm += (sm - tt) + v; else cm += (v - tt) + sm;
        sm = tt;
      }
      n += acc[t].n[c];
    }
    sums[c] += sm + cm;
    counts[c] += n;
    for (int t = 0; t < nthreads; ++t) {
      if (abs_sums) abs_sums[c] += acc[t].a[c];
      if (mag_sums) mag_sums[c] += acc[t].g[c];
    }
  }
  return 0;
}
