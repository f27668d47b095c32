// sf_api.hip -- host side of libsfmi355x.so: validation, bin tables, work items, launches.
//
// Validation mirrors what the reference checks before its O(N^2) loop:
//   _validate_array_shape          src/Calculations/shapes.jl:67-105  (D in {2,3}, W, N match)
//   _assert_counts_representable   src/Calculations/serial.jl:111-120 (N(N-1)/2 fits the count type)
//   pair_geometry_for              src/HelperFunctions.jl:390-417     (known metric only)
//   operator validity              src/StructureFunctionTypes.jl:43-151
// plus the preconditions the reference assumes silently: edges sorted (searchsortedfirst on an
// unsorted vector is meaningless) and, for the flat squared plan (BinEdges.jl:440), non-negative.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/sfmi355x.h"
#include "sf_internal.h"

namespace sfk {
// kernel launchers (sf_kernels.hip)
hipError_t launch_prep(int G, int D, bool mixed, int geom_code, const void *prep_slices,
                       int n_slices, int64_t max_n, double *rec64, float *rec32, hipStream_t st);
hipError_t launch_hist(int G, int D, bool J, bool mixed, bool sload, const Params &P, int grid,
                       int block, size_t smem, hipStream_t st);
hipError_t launch_recheck(int G, int D, bool J, const Params &P, int grid, hipStream_t st);
hipError_t launch_merge(const double *src_s, const unsigned long long *src_c, double *dst_s,
                        unsigned long long *dst_c, int64_t n, hipStream_t st);
hipError_t launch_merge_checked(const double *src_s, const unsigned long long *src_c, double *dst_s,
                                unsigned long long *dst_c, int64_t n,
                                const unsigned long long *pcount, int64_t cap, int *flag,
                                hipStream_t st);
hipError_t launch_zero(void *p, int64_t bytes, hipStream_t st);
hipError_t launch_zero_multi(int count, void *const *p, const int64_t *bytes, hipStream_t st);
hipError_t launch_derive_sp(double *s, unsigned long long *c, int64_t nbins, hipStream_t st);
hipError_t launch_eval_math(int fn, int64_t n, const double *a, const double *b, double *out,
                            hipStream_t st);
hipError_t cull_workspace_bytes(int64_t N, size_t *bytes);
hipError_t launch_cull(const double *x, int64_t N, int W, int D, int geom, int rows, void *ws,
                       int32_t *perm, double *info, hipStream_t st);
int hist_occupancy(int G, int D, bool J, bool mixed, int opk, bool sload, bool fx, int block,
                   size_t smem);
int hist_opk(const Params &P);
int strip_occupancy(int G, int D, int opk, int block, size_t smem);
struct PrepSliceHost {
  const double *x, *u;
  int64_t N, pt_off;
  const int32_t *perm;
  int64_t mul;
};
}  // namespace sfk

using namespace sfk;

// Device scratch of the allocation-free entry (sf_hist_batch_device_noalloc): records, bin tables,
// work items, accumulators and the re-bin queue of one call, carved in that order.  One copy per
// device, held until the process ends; every call on a device shares it.
static __device__ __align__(256) char scratch[SF_NOALLOC_SCRATCH_BYTES];

namespace {

// A bump allocator over a caller-provided device region (the allocation-free entry's scratch).
struct Arena {
  char *base = nullptr;
  size_t cap = 0, used = 0;
};

template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t cap = 0;  // elements
  bool borrowed = false;  // p points into an Arena: never freed
  // the next n elements of the arena (256-byte aligned); hipErrorOutOfMemory when they do not fit
  hipError_t take(size_t n, Arena &a) {
    if (p && !borrowed) return hipErrorInvalidValue;  // would leak an owned allocation
    const size_t want = std::max<size_t>(n, 1);
    const size_t off = (a.used + 255) & ~(size_t)255;
    if (off > a.cap || want > (a.cap - off) / sizeof(T)) return hipErrorOutOfMemory;
    p = (T *)(a.base + off);
    cap = want;
    borrowed = true;
    a.used = off + want * sizeof(T);
    return hipSuccess;
  }
  hipError_t ensure(size_t n) {
    if (borrowed) return hipErrorInvalidValue;
    if (n <= cap && p) return hipSuccess;
    if (p) {
      hipError_t e = hipFree(p);
      p = nullptr;
      cap = 0;
      if (e != hipSuccess) return e;
    }
    size_t want = std::max<size_t>(n, 1);
    hipError_t e = hipMalloc((void **)&p, want * sizeof(T));
    if (e == hipSuccess) cap = want;
    return e;
  }
  void release() {
    if (p && !borrowed) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    borrowed = false;
  }
};

bool jl_isless(double a, double b) {
  if (a != a) return false;
  if (b != b) return true;
  if (a < b) return true;
  if (a == b) return signbit(a) && !signbit(b);
  return false;
}

int op_order(const sf_problem &p) {
  switch (p.op) {
    case SF_OP_PROJECTED: return p.nl + p.nt;
    case SF_OP_S2: return 2;
    case SF_OP_S3: return 3;
    case SF_OP_T2COMP: return 2;
    case SF_OP_L1T2COMP: return 3;
    case SF_OP_SINGLE_PASS: return 3;
    default: return p.nl;
  }
}

float round_down_f(long double x) {
  float f = (float)x;
  if ((long double)f > x) f = nextafterf(f, -INFINITY);
  return f;
}
float round_up_f(long double x) {
  float f = (float)x;
  if ((long double)f < x) f = nextafterf(f, INFINITY);
  return f;
}

// Cell partition of the positive key axis: binades [2^elo, 2^ehi) cut into 2^M cells each.
struct Partition {
  int M;
  int elo, ehi;
  int ncell() const { return (ehi - elo) << M; }
  // real lower bound of cell c (c = 0 also stands for everything below)
  long double low(int c) const {
    const int e = elo + (c >> M);
    const int m = c & ((1 << M) - 1);
    return ldexpl(1.0L + (long double)m / (long double)(1 << M), e);
  }
};

Partition make_partition(const std::vector<long double> &keys, int cap_cells, int Mmax) {
  long double mn = INFINITY, mx = 0;
  for (long double k : keys)
    if (k > 0 && k < INFINITY) {
      mn = std::min(mn, k);
      mx = std::max(mx, k);
    }
  Partition P;
  if (!(mn < INFINITY)) {  // no positive finite key
    P.elo = 0;
    P.ehi = 1;
    P.M = 0;
    return P;
  }
  int e0, e1;
  frexpl(mn, &e0);  // mn in [2^(e0-1), 2^e0)
  frexpl(mx, &e1);
  P.elo = e0 - 2;   // one spare binade below the smallest key
  P.ehi = e1 + 1;   // and above the largest
  P.elo = std::max(P.elo, -1020);
  P.ehi = std::min(P.ehi, 1020);
  P.M = 0;
  for (int M = Mmax; M >= 0; --M) {
    if (((P.ehi - P.elo) << M) <= cap_cells) {
      P.M = M;
      break;
    }
  }
  return P;
}

}  // namespace

namespace {
// A plan holds everything one batch shape needs on the device: per-point records, bin tables,
// work items, accumulators and the re-bin queue.  Building it does the host work and the
// uploads; launching it enqueues kernels only, so a launch can be captured into a hipGraph and
// replayed while the caller rewrites x/u in place (time slices).
struct Plan {
  sf_ctx *ctx = nullptr;
  int n = 0, G = 0, D = 0, geom_code = 0, ncell = 0, S64 = 0, F32 = 0;
  int nout = 0;  // output slices (n is the internal slice count: six per slice for 2D single pass)
  bool J = false, mixed = false, sload = false, sp = false, fx = false;
  int block = 256, tj = 256, copies = 1, grid = 1, ncell_max = 1, n_items = 0;
  int strip = 0;            // B-strip launch (fields per strip), 0: off
  int n_items64 = 0;        // B-strip: the per-slice items the binary64 kernel runs instead
  int64_t total_pts = 0, max_n = 0, total_pairs = 0, cap = 0;
  Params P{};
  DevBuf<double> rec64;
  DevBuf<float> rec32;
  DevBuf<Slice> slices;
  DevBuf<Item> items;
  DevBuf<Cell> cells;
  DevBuf<double> dedges, vedges;
  DevBuf<float> vedges32;
  DevBuf<double> acc_s;
  DevBuf<unsigned long long> acc_c;
  DevBuf<int> counter;
  DevBuf<unsigned long long> pcount;
  DevBuf<Item> items64;
  DevBuf<Pending> pending;
  DevBuf<PrepSliceHost> prep;
  DevBuf<int> flag;
  DevBuf<int32_t> perm;   // spatial order of the binary32 records (tile culling)
  bool culled = false;    // some work items were dropped by tile culling
  Arena *arena = nullptr;  // set: every buffer is taken from this region (no device allocation)
  // fused outputs: 1D operators evaluated with the main (joint) one, their accumulators and the
  // caller's outputs they merge into
  int n1d = 0, mcopies = 0;
  DevBuf<double> msums;
  DevBuf<unsigned long long> mcounts;
  double *mdst_s[kMaxFused] = {};
  unsigned long long *mdst_c[kMaxFused] = {};
  hipGraphExec_t graph = nullptr;
  double *g_sums = nullptr;
  unsigned long long *g_counts = nullptr;
  hipStream_t g_stream = nullptr;
  void release() {
    if (graph) (void)hipGraphExecDestroy(graph);
    graph = nullptr;
    rec64.release(); rec32.release(); slices.release(); items.release(); cells.release();
    dedges.release(); vedges.release(); vedges32.release(); acc_s.release(); acc_c.release();
    counter.release(); pcount.release(); pending.release(); prep.release(); flag.release();
    items64.release();
    perm.release();
    msums.release();
    mcounts.release();
  }
};

}  // namespace

namespace {
// Launch shape of the histogram kernel; 0 keeps the built-in choice.  block: threads per
// workgroup (rows of an i-tile); j_chunk: columns per work item; bpc: workgroups per CU; variant:
// column records through LDS (1), scalar loads (2) or B-strips (3); copies: privatised copies of
// the LDS histogram (1D: lane groups, a power of two up to 64; joint: 1 or 2, by lane halves);
// layout: joint histogram slots value-bin fastest (1) or distance-bin fastest (2), odd row stride.
struct Tuning {
  int block = 0, j_chunk = 0, bpc = 0, variant = 0, copies = 0, layout = 0;
  bool any() const { return block || j_chunk || bpc || variant || copies || layout; }
};

// Autotuner key: (geometry, D, joint, precision code, operator class, 2D single-pass planes,
// log2 of the pair count of the caller's slices, n_dist, n_val).  Shared by the tuner and the
// plan builder so that an entry the tuner stores is the one a later call looks up.
using TuneKey = std::tuple<int, int, int, int, int, int, int, int, int>;
int hist_opk_of(const sf_problem &p0) {
  Params Pk{};
  Pk.op = p0.op;
  Pk.nl = p0.nl;
  Pk.nt = p0.nt;
  Pk.n_val = p0.n_val;
  Pk.nmom = (p0.op == SF_OP_SINGLE_PASS && p0.n_val == 0) ? 4 : 1;
  return hist_opk(Pk);
}
TuneKey tune_key(const sf_problem &p0, int64_t input_pairs) {
  int lg = 0;
  for (int64_t v = input_pairs; v > 1; v >>= 1) ++lg;
  const int G = p0.geometry == SF_GEOM_FLAT ? kFlat : kSphere;
  return TuneKey(G, p0.D, p0.n_val > 0 ? 1 : 0, p0.precision, hist_opk_of(p0),
                 sp2d(p0.op, p0.n_val) ? 1 : 0, lg, p0.n_dist, p0.n_val);
}
}  // namespace

struct sf_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int n_cu = 0;
  size_t max_smem = 0;
  int clock_khz = 0;   // peak shader clock
  std::string err;
  int64_t stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  bool ev_pending = false;
  Tuning tune;  // sf_set_tuning override (all zero: none)
  // largest re-bin queue a call of this context has needed (entries): later plans start there
  int64_t queue_hint = 0;
  // autotuner cache: the call's shape (tune_key) -> the fastest launch shape measured for it
  std::map<TuneKey, Tuning> tuned;
  // device buffers: host-input staging and the plan reused by one-shot calls
  DevBuf<double> d_x, d_u;
  Plan oneshot;
  // the allocation-free entry: its plan (buffers borrowed from the scratch array) and this
  // device's address of the array
  Plan noalloc;
  char *scratch_ptr = nullptr;
  Plan fused;  // the fused-output entry's plan
  // spatial sort and tile culling: device workspace and the slice/tile boxes read back
  DevBuf<char> cull_ws;
  DevBuf<double> cull_info;

  int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    err = buf;
    return code;
  }
};

#define HIPCHK(expr)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return ctx->fail(SF_E_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));     \
  } while (0)

namespace {

// A LinearBinEdges rule must describe the materialised edge vector it comes with: n edges (plus
// -Inf / +Inf when padded), first and last edges equal to its fields, a finite non-negative step.
int validate_rule(sf_ctx *ctx, const sf_bin_rule *r, int nbins, const double *edges, const char *what) {
  if (r->kind != SF_RULE_LINEAR) return ctx->fail(SF_E_ARG, "%s: unknown bin rule kind %d", what, r->kind);
  if (r->padded != 0 && r->padded != 1) return ctx->fail(SF_E_ARG, "%s: padded must be 0 or 1", what);
  const int64_t want = r->padded ? r->n + 1 : r->n - 1;
  if (r->n < 2 || want != nbins)
    return ctx->fail(SF_E_DIM, "%s: LinearBinEdges of length %lld does not give %d bins", what, (long long)r->n, nbins);
  if (!(r->first_edge > -INFINITY && r->last_edge < INFINITY && r->first_edge <= r->last_edge))
    return ctx->fail(SF_E_ARG, "%s: LinearBinEdges needs finite first <= last", what);
  if (!(r->step_val >= 0.0 && r->step_val < INFINITY) || r->inv_step != r->inv_step)
    return ctx->fail(SF_E_ARG, "%s: LinearBinEdges needs a finite non-negative step", what);
  const int o = r->padded ? 1 : 0;
  if (edges[o] != r->first_edge || edges[o + r->n - 1] != r->last_edge)
    return ctx->fail(SF_E_ARG, "%s: the edge vector does not hold the LinearBinEdges range", what);
  if (r->padded && !(edges[0] == -INFINITY && edges[nbins] == INFINITY))
    return ctx->fail(SF_E_ARG, "%s: a padded rule needs -Inf and +Inf around the range", what);
  return SF_OK;
}

int validate_one(sf_ctx *ctx, const sf_problem *p, bool host_ptrs, int idx) {
  if (!p) return ctx->fail(SF_E_ARG, "slice %d: null problem", idx);
  if (p->D != 2 && p->D != 3)
    return ctx->fail(SF_E_DIM, "expected spatial/velocity dimension D=2 or D=3 on axis 1; got D=%d", p->D);
  if (p->geometry == SF_GEOM_FLAT) {
    if (p->W != p->D)
      return ctx->fail(SF_E_DIM,
                       "this geometry locates a point with %d coordinate(s) on axis 1 of x, but got "
                       "size(x,1)=%d (velocity dimension D=%d from size(u,1))", p->D, p->W, p->D);
  } else if (p->geometry == SF_GEOM_SPHERE_DEG || p->geometry == SF_GEOM_SPHERE_RAD) {
    if (p->W != 2)
      return ctx->fail(SF_E_DIM,
                       "this geometry locates a point with 2 coordinate(s) on axis 1 of x, but got "
                       "size(x,1)=%d (velocity dimension D=%d from size(u,1))", p->W, p->D);
    if (p->geometry == SF_GEOM_SPHERE_DEG && !(p->radius > 0.0 && p->radius < INFINITY))
      return ctx->fail(SF_E_ARG, "Haversine radius must be positive and finite; got %g", p->radius);
  } else {
    return ctx->fail(SF_E_ARG, "no pair geometry is defined for geometry code %d", p->geometry);
  }
  if (p->N < 0) return ctx->fail(SF_E_DIM, "N must be non-negative; got %lld", (long long)p->N);
  if (p->N >= (1LL << 31) - 4096)
    return ctx->fail(SF_E_UNSUPPORTED, "N=%lld exceeds this backend's 2^31 point limit", (long long)p->N);
  if (p->N > 0 && (!p->x || !p->u)) return ctx->fail(SF_E_ARG, "slice %d: null x or u", idx);
  switch (p->op) {
    case SF_OP_PROJECTED:
      if (p->nl < 0 || p->nt < 0 || p->nl + p->nt > 3)
        return ctx->fail(SF_E_UNSUPPORTED,
                         "ProjectedStructureFunctionType{%d,%d}: this backend supports NL,NT >= 0, NL+NT <= 3",
                         p->nl, p->nt);
      break;
    case SF_OP_S2: case SF_OP_S3: case SF_OP_T2COMP: case SF_OP_L1T2COMP:
      break;
    case SF_OP_FULLVEC:
      if (p->nl < 1 || p->nl > 3)
        return ctx->fail(SF_E_UNSUPPORTED, "FullVectorStructureFunctionType{%d}: NF must be 1..3", p->nl);
      break;
    case SF_OP_SINGLE_PASS:  // n_val > 0: the 2D form, six edge vectors (checked below)
      break;
    default:
      return ctx->fail(SF_E_ARG, "unknown operator code %d (derived operators are not pairwise)", p->op);
  }
  if (p->n_dist < 1 || !p->dist_edges)
    return ctx->fail(SF_E_ARG, "distance_bins must have at least 2 edges");
  for (int k = 0; k <= p->n_dist; ++k) {
    const double e = p->dist_edges[k];
    if (e != e) return ctx->fail(SF_E_ARG, "distance_bins contains NaN");
    if (k > 0 && jl_isless(e, p->dist_edges[k - 1]))
      return ctx->fail(SF_E_ARG, "distance_bins must be sorted ascending");
  }
  if (p->geometry == SF_GEOM_FLAT)
    for (int k = 0; k <= p->n_dist; ++k)
      if (p->dist_edges[k] < 0.0 && !(k == 0 && p->dist_edges[0] == -INFINITY))
        return ctx->fail(SF_E_ARG,
                         "Euclidean distance_bins must be non-negative (squared-distance plan); only a "
                         "leading -Inf (InfPaddedBinEdges catch-all) is allowed");
  if (p->n_val < 0) return ctx->fail(SF_E_ARG, "n_val must be >= 0");
  if (p->n_val > 0) {
    if (!p->val_edges) return ctx->fail(SF_E_ARG, "value_bins missing");
    // the single pass's 2D form carries one edge vector per invariant, back to back
    const int nvec = sp2d(p->op, p->n_val) ? kSpRows : 1;
    for (int t = 0; t < nvec; ++t) {
      const double *ve = p->val_edges + (size_t)t * (p->n_val + 1);
      for (int k = 0; k <= p->n_val; ++k) {
        const double e = ve[k];
        if (e != e) return ctx->fail(SF_E_ARG, "value_bins contains NaN");
        if (k > 0 && jl_isless(e, ve[k - 1]))
          return ctx->fail(SF_E_ARG, "value_bins must be sorted ascending");
      }
    }
  }
  if (p->precision != SF_PREC_FP64 && p->precision != SF_PREC_MIXED && p->precision != SF_PREC_MIXED_FIXED)
    return ctx->fail(SF_E_ARG, "precision must be SF_PREC_FP64, SF_PREC_MIXED or SF_PREC_MIXED_FIXED");
  if (p->count_bits != 32 && p->count_bits != 64)
    return ctx->fail(SF_E_ARG, "count_bits must be 32 or 64");
  if (p->semantics != SF_SEM_POINT && p->semantics != SF_SEM_BATCH)
    return ctx->fail(SF_E_ARG, "semantics must be SF_SEM_POINT or SF_SEM_BATCH");
  if (p->dist_rule) {
    const int rc = validate_rule(ctx, p->dist_rule, p->n_dist, p->dist_edges, "distance_bins");
    if (rc) return rc;
  }
  if (p->val_rule) {
    if (p->n_val == 0) return ctx->fail(SF_E_ARG, "val_rule given without value_bins");
    if (p->op == SF_OP_SINGLE_PASS)
      return ctx->fail(SF_E_UNSUPPORTED, "the single pass's 2D form takes explicit value edge vectors only");
    const int rc = validate_rule(ctx, p->val_rule, p->n_val, p->val_edges, "value_bins");
    if (rc) return rc;
  }
  const unsigned __int128 pairs = (unsigned __int128)p->N * (unsigned __int128)(p->N > 0 ? p->N - 1 : 0) / 2;
  const unsigned __int128 cmax = p->count_bits == 32 ? 0xffffffffull : 0xffffffffffffffffull;
  if (pairs > cmax)
    return ctx->fail(SF_E_COUNT,
                     "count_eltype=UInt%d cannot represent the worst-case pair count for N=%lld; "
                     "pass count_eltype=UInt64 or Int64.", p->count_bits, (long long)p->N);
  if (host_ptrs && p->geometry == SF_GEOM_SPHERE_RAD) {
    for (int64_t k = 0; k < 2 * p->N; ++k)
      if (fabs(p->x[k]) > 1.6e6)
        return ctx->fail(SF_E_UNSUPPORTED, "SphericalAngle coordinates beyond |1.6e6| rad are not supported");
  }
  return SF_OK;
}

int validate_batch(sf_ctx *ctx, int n, const sf_problem *s, bool host_ptrs) {
  if (n < 0) return ctx->fail(SF_E_ARG, "n_slices must be >= 0");
  if (n > 0 && !s) return ctx->fail(SF_E_ARG, "null slices");
  for (int k = 0; k < n; ++k) {
    int rc = validate_one(ctx, &s[k], host_ptrs, k);
    if (rc) return rc;
    if (k > 0) {
      const sf_problem &a = s[0], &b = s[k];
      if (a.D != b.D || a.W != b.W || a.geometry != b.geometry || a.op != b.op || a.nl != b.nl ||
          a.nt != b.nt || a.n_dist != b.n_dist || a.n_val != b.n_val || a.precision != b.precision ||
          a.semantics != b.semantics ||
          (a.geometry == SF_GEOM_SPHERE_DEG && a.radius != b.radius))
        return ctx->fail(SF_E_ARG, "slice %d: all slices must share D, geometry, operator, bin counts and precision", k);
    }
  }
  return SF_OK;
}

// Host-side tables for one slice.
struct SliceTables {
  std::vector<Cell> tab64, tab32;
  std::vector<double> dedges;  // exact keys
  Slice desc{};
};

// binary64 decision points of a LinearBinEdges rule lie within this distance of the edge values
long double rule_slack(const sf_bin_rule *r) {
  if (!r) return 0.0L;
  return ldexpl(1.0L, -48) * (fabsl((long double)r->first_edge) + fabsl((long double)r->last_edge));
}

LinRule lin_rule_of(const sf_bin_rule *r) {
  LinRule L{};
  if (!r) return L;
  L.on = 1;
  L.padded = r->padded;
  L.n = r->n;
  L.f = r->first_edge;
  L.l = r->last_edge;
  L.inv = r->inv_step;
  L.step = r->step_val;
  return L;
}

void build_tables(const sf_problem &p, double xmax, SliceTables &T) {
  const int nE = p.n_dist + 1;
  const bool flat = p.geometry == SF_GEOM_FLAT;
  const long double R = flat ? 1.0L : (p.geometry == SF_GEOM_SPHERE_DEG ? (long double)p.radius : 1.0L);
  // ---- binary64 exact keys and table
  T.dedges.resize(nE);
  std::vector<long double> k64(nE);
  for (int k = 0; k < nE; ++k) {
    const double e = p.dist_edges[k];
    // flat, point-field kernels: squared edges (a leading -Inf, InfPaddedBinEdges' underflow
    // edge, stays -Inf); flat, batch-leading kernels: the edges themselves (r = sqrt(r^2) is
    // binned); sphere: the edges themselves, -0.0 -> +0.0 (keys are > 0)
    if (flat) T.dedges[k] = p.semantics == SF_SEM_BATCH ? e : (e == -INFINITY ? e : e * e);
    else T.dedges[k] = e == 0.0 ? 0.0 : e;
    k64[k] = T.dedges[k];
  }
  Partition P64 = make_partition(k64, 1024, 6);
  const int n64 = P64.ncell();
  T.tab64.resize(n64);
  for (int c = 0; c < n64; ++c) {
    const long double lowl = (c == 0) ? 0.0L : P64.low(c);
    const double low = (double)lowl;  // cell bounds are exact binary64 values
    int b = 0;
    if (c == 0) {
      // keys below the window: only non-positive edges are below them
      for (int k = 0; k < nE; ++k) if (T.dedges[k] <= 0.0) b = k + 1;
    } else {
      for (int k = 0; k < nE; ++k) if (T.dedges[k] < low) b = k + 1;
    }
    T.tab64[c].bx = b;
    T.tab64[c].lo = INFINITY;
    T.tab64[c].hi = INFINITY;
    T.tab64[c].blo = 0;
  }
  // ---- binary32 keys: flat r^2 = e^2; sphere chord^2 = (2 sin(e / 2R))^2
  std::vector<long double> C(nE);
  for (int k = 0; k < nE; ++k) {
    const long double e = p.dist_edges[k];
    if (flat) {
      C[k] = (e == -INFINITY) ? e : e * e;
    } else if (e <= 0) {
      C[k] = e;
    } else {
      const long double th = std::min(e / R, (long double)M_PIl);
      const long double c = 2.0L * sinl(0.5L * th);
      C[k] = c * c;
    }
  }
  // margins: relative 2^-18 (64 ulp) plus an absolute term for the hi/lo split of the inputs
  const long double drel = ldexpl(1.0L, -20);
  const long double e1 = flat ? ldexpl(1.0L, -44) * std::max((long double)xmax, 1e-300L) : ldexpl(1.0L, -44);
  // a LinearBinEdges rule decides within a few binary64 ulps of max(|first|, |last|) of each
  // materialised edge (its rounding of the step, 1 / step and the two multiply-adds); the margin
  // covers 2^-48 (|first| + |last|), in key-space units
  const long double dr = rule_slack(p.dist_rule) / R;
  std::vector<float> lo(nE), hi(nE);
  for (int k = 0; k < nE; ++k) {
    if (C[k] == INFINITY) { lo[k] = INFINITY; hi[k] = INFINITY; continue; }
    const long double a = 2.0L * sqrtl(fabsl(C[k])) * (e1 + dr) + (e1 + dr) * (e1 + dr) + 1e-30L;
    lo[k] = round_down_f(C[k] - fabsl(C[k]) * drel - a);
    hi[k] = round_up_f(C[k] + fabsl(C[k]) * drel + a);
  }
  std::vector<long double> ckeys(C.begin(), C.end());
  Partition P32 = make_partition(ckeys, kCell32Max, 6);
  // keep every cell's lower bound representable in binary32
  P32.elo = std::max(P32.elo, -125);
  P32.ehi = std::min(P32.ehi, 127);
  if (P32.ehi <= P32.elo) P32.ehi = P32.elo + 1;
  const int n32 = P32.ncell();  // <= kCell32Max: at most 252 binades, M chosen to fit
  T.tab32.resize(n32);
  for (int c = 0; c < n32; ++c) {
    const float Fc = (c == 0) ? -INFINITY : (float)P32.low(c);
    const float Fn = (c + 1 >= n32) ? INFINITY : (float)P32.low(c + 1);
    int nint = 0, which = -1, below = 0;
    for (int k = 0; k < nE; ++k) {
      if (C[k] == -INFINITY || hi[k] < Fc) { ++below; continue; }  // -Inf: below every key
      if (lo[k] >= Fn) continue;
      ++nint;
      which = k;
    }
    Cell &cl = T.tab32[c];
    cl.bx = 0;
    cl.blo = below;
    if (nint == 0) {
      cl.lo = INFINITY;
      cl.hi = INFINITY;
    } else if (nint == 1) {
      cl.lo = lo[which];
      cl.hi = hi[which];
    } else {
      cl.lo = -INFINITY;
      cl.hi = INFINITY;
    }
  }
  Slice &s = T.desc;
  s.drule = lin_rule_of(p.dist_rule);
  s.vrule = p.n_val > 0 ? lin_rule_of(p.val_rule) : LinRule{};
  s.M64 = P64.M;
  s.ncell64 = n64;
  s.base64 = ((int64_t)(P64.elo + 1023) << P64.M);  // cell 0 starts at 2^elo
  s.M32 = P32.M;
  s.ncell32 = n32;
  s.base32 = ((P32.elo + 127) << P32.M);
  // value axis
  s.v_uniform = 0;
  s.v_negzero = 0;
  s.v0 = 0; s.vinv = 0; s.v0f = 0; s.vinvf = 0; s.vtc = 0; s.vslack = 0; s.vabs = 0;
  s.v_fx = 0; s.vwf = 0; s.vsf = 0; s.fxc1 = 0; s.fxc0 = 0; s.plane = 0; s.vsmall = 0; s.gz = 0;
  s.v_log = 0; s.vk0 = 0; s.vl0f = 0; s.vlinvf = 0;
  if (p.n_val > 0) {
    const int nv = p.n_val;
    const double *E = p.val_edges;
    for (int k = 0; k <= nv; ++k) if (E[k] == 0.0 && signbit(E[k])) s.v_negzero++;
    const double v0 = E[0], v1 = E[nv];
    if (v0 > -INFINITY && v1 < INFINITY && v1 > v0) {
      const long double w = ((long double)v1 - v0) / nv;
      long double dev = 0;
      for (int k = 0; k <= nv; ++k) dev = std::max(dev, fabsl(((long double)E[k] - v0) / w - k));
      if (dev < 0.25L) {
        s.v_uniform = 1;
        s.v0 = v0;
        s.vinv = (double)(1.0L / w);
        s.v0f = (float)v0;
        s.vinvf = (float)(1.0L / w);
        // exact product of two binary32 numbers, rounded once: within 2^-24 |v0 / w| of
        // -v0f / w, which the slack's 4u (1 + |v0 / w|) term covers
        s.vtc = (float)(-(double)s.v0f * (double)s.vinvf);
        // slack in t units: edge non-uniformity + rounding of v0 and 1/w in binary32
        const long double u = ldexpl(1.0L, -24);
        s.vslack = round_up_f(dev + 4.0L * u * (1.0L + fabsl((long double)v0 / w)) + 4.0L * u * nv + 1e-6L +
                              rule_slack(p.val_rule) / w);
        // fixed-point joint cells need every value of bin g within [-1/64, 1 + 1/64] widths of
        // the binary32 grid point fmaf(g, vwf, v0f)
        if (dev < 1.0L / 64) {
          s.v_fx = 1;
          s.vwf = (float)w;
          s.vsf = kFxScale / s.vwf;
          s.fxc1 = (float)((double)s.vwf * (double)s.vsf);
          s.fxc0 = (float)(8388608.0 + (double)kFxBias - (double)s.v0f * (double)s.vsf);
          // a certified bin g of a value with |v| < w/4 is within half a bin of [-w/4, w/4] (the
          // binary32 value's error is below half a bin, or the pair is re-binned), so g - gz is
          // 1 .. 3 with gz = floor((-w/4 - v0)/w) - 2; the kZeroBins = 5 cells gz .. gz + 4 leave
          // a bin of margin on either side
          s.vsmall = (float)(0.25L * w);
          const long double gzl = floorl((-0.25L * w - (long double)v0) / w) - 2.0L;
          s.gz = (int32_t)std::max(-1.0e9L, std::min(1.0e9L, gzl));
        }
      }
    }
    if (!s.v_uniform) s.vslack = (float)ldexpl(1.0L, -21);  // relative slack for binary32 edges
    s.vabs = round_up_f(2.0L * rule_slack(p.val_rule));
    if (!s.v_uniform) {
      // log-spaced from the first positive edge on (a leading 0 or negative edges allowed): the
      // binary32 kernel guesses the bin from log2(v) and walks to it; the walk alone decides, so
      // the guess only needs to be close
      int k0 = 0;
      while (k0 <= nv && !(E[k0] > 0.0)) ++k0;
      if (k0 + 3 <= nv && E[nv] < INFINITY) {
        const long double l0 = log2l((long double)E[k0]), l1 = log2l((long double)E[nv]);
        const long double st = (l1 - l0) / (nv - k0);
        long double dev = 0;
        for (int k = k0; k <= nv; ++k) dev = std::max(dev, fabsl((log2l((long double)E[k]) - l0) / st - (k - k0)));
        if (st > 0 && dev < 0.25L) {
          s.v_log = 1;
          s.vk0 = k0;
          s.vl0f = (float)l0;
          s.vlinvf = (float)(1.0L / st);
        }
      }
    }
  }
}

int64_t pairs_of(int64_t n) { return n > 1 ? n * (n - 1) / 2 : 0; }

// output cells per slice: (n_dist), (n_dist, n_val), or (6, n_dist) for the single pass
int out_cells(const sf_problem &p) {
  if (p.op == SF_OP_SINGLE_PASS) return kSpRows * p.n_dist * (p.n_val > 0 ? p.n_val : 1);
  return p.n_dist * (p.n_val > 0 ? p.n_val : 1);
}

// _sp1d_derive_rows! (serial_single_pass.jl:208-219) on host memory
void derive_sp_host(double *s, uint64_t *c, int64_t nbins) {
  for (int64_t b = 0; b < nbins; ++b) {
    double *col = s + kSpRows * b;
    uint64_t *cc = c + kSpRows * b;
    col[2] = col[0] - col[1];
    col[5] = col[3] - col[4];
    for (int t = 1; t < kSpRows; ++t) cc[t] = cc[0];
  }
}

}  // namespace

// ----------------------------------------------------------------------------------------------
// C ABI
// ----------------------------------------------------------------------------------------------
extern "C" int sf_create(int device, sf_ctx **out) {
  if (!out) return SF_E_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return SF_E_DEVICE;
  if (device < 0 || device >= ndev) return SF_E_ARG;
  sf_ctx *ctx = new sf_ctx();
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess) { delete ctx; return SF_E_DEVICE; }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) { delete ctx; return SF_E_DEVICE; }
  ctx->n_cu = prop.multiProcessorCount;
  ctx->max_smem = prop.sharedMemPerBlock;
  ctx->clock_khz = prop.clockRate;
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && v > 0)
    ctx->max_smem = (size_t)v;
  if (hipStreamCreate(&ctx->stream) != hipSuccess) { delete ctx; return SF_E_DEVICE; }
  // this device's copy of the allocation-free entry's scratch array (the array's name is a host
  // stand-in; only this address is used on the host)
  if (hipGetSymbolAddress((void **)&ctx->scratch_ptr, HIP_SYMBOL(scratch)) != hipSuccess) {
    ctx->scratch_ptr = nullptr;
    (void)hipGetLastError();
  }
  *out = ctx;
  return SF_OK;
}

extern "C" void sf_destroy(sf_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->d_x.release();
  ctx->d_u.release();
  ctx->oneshot.release();
  ctx->noalloc.release();
  ctx->fused.release();
  ctx->cull_ws.release();
  ctx->cull_info.release();
  for (int k = 0; k < 4; ++k)
    if (ctx->ev[k]) (void)hipEventDestroy(ctx->ev[k]);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

extern "C" const char *sf_last_error(const sf_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

extern "C" int sf_device_info(sf_ctx *ctx, int64_t out[4]) {
  if (!ctx || !out) return SF_E_ARG;
  out[0] = ctx->n_cu;
  out[1] = ctx->clock_khz;
  out[2] = (int64_t)ctx->max_smem;
  out[3] = 64;
  return SF_OK;
}

extern "C" int sf_last_stats(sf_ctx *ctx, int64_t out[8]) {
  if (!ctx || !out) return SF_E_ARG;
  if (ctx->ev_pending) {
    ctx->ev_pending = false;
    float t01 = 0, t12 = 0, t03 = 0;
    if (hipEventSynchronize(ctx->ev[3]) == hipSuccess) {
      (void)hipEventElapsedTime(&t01, ctx->ev[0], ctx->ev[1]);
      (void)hipEventElapsedTime(&t12, ctx->ev[1], ctx->ev[2]);
      (void)hipEventElapsedTime(&t03, ctx->ev[0], ctx->ev[3]);
    }
    ctx->stats[4] = (int64_t)(t01 * 1e6);
    ctx->stats[5] = (int64_t)(t12 * 1e6);
    ctx->stats[6] = (int64_t)(t03 * 1e6);
  }
  for (int k = 0; k < 8; ++k) out[k] = ctx->stats[k];
  return SF_OK;
}

extern "C" int sf_check(int32_t n_slices, const sf_problem *slices, char *msg, int32_t msg_len) {
  sf_ctx tmp;  // carries the message only; no device state is touched
  const int rc = validate_batch(&tmp, n_slices, slices, true);
  if (msg && msg_len > 0) {
    strncpy(msg, tmp.err.c_str(), (size_t)msg_len - 1);
    msg[msg_len - 1] = 0;
  }
  return rc;
}

extern "C" int sf_eval_math(sf_ctx *ctx, int32_t fn, int64_t n, const double *a, const double *b,
                            double *out) {
  if (!ctx) return SF_E_ARG;
  if (fn < 0 || fn > 4 || n < 0 || (n > 0 && (!a || !out))) return ctx->fail(SF_E_ARG, "bad arguments");
  if ((fn == 2 || fn == 4) && n > 0 && !b) return ctx->fail(SF_E_ARG, "fn needs b");
  if (n == 0) return SF_OK;
  HIPCHK(hipSetDevice(ctx->device));
  const int nout = (fn <= 1) ? 2 : 1;
  DevBuf<double> da, db, dout;
  HIPCHK(da.ensure((size_t)n));
  HIPCHK(db.ensure((size_t)n));
  HIPCHK(dout.ensure((size_t)n * nout));
  hipStream_t st = ctx->stream;
  HIPCHK(hipMemcpyAsync(da.p, a, sizeof(double) * n, hipMemcpyHostToDevice, st));
  if (b) HIPCHK(hipMemcpyAsync(db.p, b, sizeof(double) * n, hipMemcpyHostToDevice, st));
  HIPCHK(launch_eval_math(fn, n, da.p, db.p, dout.p, st));
  HIPCHK(hipMemcpyAsync(out, dout.p, sizeof(double) * n * nout, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  da.release(); db.release(); dout.release();
  return SF_OK;
}

extern "C" int sf_set_tuning(sf_ctx *ctx, int32_t block, int32_t j_chunk, int32_t blocks_per_cu,
                             int32_t variant, int32_t copies, int32_t layout) {
  if (!ctx) return SF_E_ARG;
  if (block != 0 && (block < 64 || block > 1024 || block % 64 != 0))
    return ctx->fail(SF_E_ARG, "block must be 0 or a multiple of 64 in [64, 1024]");
  if (j_chunk < 0 || blocks_per_cu < 0) return ctx->fail(SF_E_ARG, "tuning values must be >= 0");
  if (variant < 0 || variant > 3) return ctx->fail(SF_E_ARG, "variant must be 0, 1, 2 or 3");
  if (copies < 0 || copies > 64 || (copies & (copies - 1)) != 0)
    return ctx->fail(SF_E_ARG, "copies must be 0 or a power of two up to 64");
  if (layout < 0 || layout > 2) return ctx->fail(SF_E_ARG, "layout must be 0, 1 or 2");
  ctx->tune = Tuning{block, j_chunk, blocks_per_cu, variant, copies, layout};
  return SF_OK;
}

namespace {

// ---- spatial sort and tile culling ----------------------------------------------------------
// When every distance edge of a slice is finite, a pair farther apart than the last edge is not
// counted.  Points are put in Morton (Z-curve) order of their key-space coordinates (flat: x;
// sphere: the unit vector), so that a tile of consecutive points is spatially compact, and a
// work item (row tile, column chunk) whose bounding boxes are farther apart than the last edge is
// dropped.  Only the binary32 records are reordered (each keeps its caller index); the binary64
// records, and with them every re-binned pair, stay in the caller's order.
struct Box {
  double lo[3], hi[3];
};

struct CullInfo {
  bool active = false;
  double limit = 0;               // a pair of boxes farther apart than this holds no counted pair
  std::vector<Box> tiles;         // bounding box of each `rows`-record tile, in record order
};

// Spatial order and tile boxes of one slice, computed on the device (launch_cull): the record
// order goes to perm (device, N entries); ci.active says whether culling is worth it.
int cull_prepare(sf_ctx *ctx, const sf_problem &p, const double *dx, double xmax, int rows,
                 hipStream_t st, CullInfo &ci, int32_t *perm) {
  ci.active = false;
  const int64_t N = p.N;
  const double emax = p.dist_edges[p.n_dist];
  if (N < 4 * (int64_t)rows || !(emax < INFINITY)) return SF_OK;
  const bool flat = p.geometry == SF_GEOM_FLAT;
  // largest distance in key-space units: flat |dx|; sphere chord 2 sin(theta / 2)
  double reach;
  if (flat) {
    reach = std::max(emax, 0.0);
  } else {
    const double R = p.geometry == SF_GEOM_SPHERE_DEG ? p.radius : 1.0;
    const double th = std::min(std::max(emax, 0.0) / R, M_PI);
    reach = 2.0 * sin(0.5 * th);
  }
  // cheap refusals before touching the data: the sphere's extent is at most 2, a flat slice's
  // at most 2 max|x| (when the caller measured it)
  if (!flat && reach >= 1.0) return SF_OK;
  if (flat && xmax > 0 && reach >= xmax) return SF_OK;
  const int64_t ntile = (N + rows - 1) / rows;
  size_t wsb = 0;
  HIPCHK(cull_workspace_bytes(N, &wsb));
  HIPCHK(ctx->cull_ws.ensure(wsb));
  HIPCHK(ctx->cull_info.ensure((size_t)(8 + 6 * ntile)));
  const int geom = flat ? 0 : (p.geometry == SF_GEOM_SPHERE_DEG ? 1 : 2);
  HIPCHK(launch_cull(dx, N, p.W, p.D, geom, rows, ctx->cull_ws.p, perm, ctx->cull_info.p, st));
  std::vector<double> info((size_t)(8 + 6 * ntile));
  HIPCHK(hipMemcpyAsync(info.data(), ctx->cull_info.p, sizeof(double) * info.size(), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (info[6] != 0.0) return SF_OK;  // non-finite coordinates: no culling
  // margin: relative 1e-6 plus the rounding of coordinates of magnitude cmax
  ci.limit = reach * (1.0 + 1e-6) + 1e-13 * std::max(1.0, info[7]);
  // worth it only if the reach is small against the extent
  double ext = 0;
  for (int k = 0; k < 3; ++k) ext = std::max(ext, info[3 + k] - info[k]);
  if (!(ci.limit < 0.5 * ext)) return SF_OK;
  ci.tiles.resize((size_t)ntile);
  for (int64_t t = 0; t < ntile; ++t)
    for (int k = 0; k < 3; ++k) {
      ci.tiles[t].lo[k] = info[8 + 6 * t + k];
      ci.tiles[t].hi[k] = info[8 + 6 * t + 3 + k];
    }
  ci.active = true;
  return SF_OK;
}

// true if no point of box a is within `limit` of a point of box b
bool boxes_apart(const Box &a, const Box &b, double limit) {
  double g2 = 0;
  for (int k = 0; k < 3; ++k) {
    const double g = std::max(0.0, std::max(b.lo[k] - a.hi[k], a.lo[k] - b.hi[k]));
    g2 += g * g;
  }
  return g2 > limit * limit;
}

// Stride of the binary32 record order: record k holds point (k * mul) mod N.  Points that the
// caller stores next to each other (a drifter track, a grid row) tend to fall in the same
// (distance, value) cell against a common column point, so a wave whose 64 rows are neighbours
// sends most of its LDS atomics to one or two cells and they serialize.  A stride near N times
// the golden ratio conjugate (a Kronecker sequence) hands each wave rows spread over the whole
// slice.  The pair set and the cells are unchanged; only the order of the additions differs.
int64_t interleave_mul(int64_t N) {
  if (N < 128) return 0;
  int64_t m = (int64_t)((double)N * 0.6180339887498949);
  auto gcd = [](int64_t a, int64_t b) {
    while (b) {
      const int64_t t = a % b;
      a = b;
      b = t;
    }
    return a;
  };
  while (gcd(m, N) != 1) ++m;
  return m;
}

// Fused outputs: the 1D operators a fused plan evaluates besides its main operator.
struct MultiSpec {
  int n1d = 0;
  int op[kMaxFused] = {}, nl[kMaxFused] = {}, nt[kMaxFused] = {};
};

// a plan buffer: allocated, or (allocation-free entry) taken from the plan's arena
#define PLBUF(buf, n)                                                                             \
  do {                                                                                            \
    const hipError_t e_ = pl.arena ? (buf).take((size_t)(n), *pl.arena) : (buf).ensure((size_t)(n)); \
    if (e_ == hipErrorOutOfMemory && pl.arena)                                                    \
      return ctx->fail(SF_E_UNSUPPORTED,                                                          \
                       "this call needs more device scratch than the allocation-free entry's %llu MiB " \
                       "(at %s); use sf_hist_batch_device", (unsigned long long)(SF_NOALLOC_SCRATCH_BYTES >> 20), #buf); \
    if (e_ != hipSuccess) return ctx->fail(SF_E_DEVICE, "%s: %s", #buf, hipGetErrorString(e_));   \
  } while (0)

// Host work and uploads for one batch.  x/u of each slice are device pointers.
int plan_build(sf_ctx *ctx, Plan &pl, int n_in, const sf_problem *sl_in, const double *const *dx_in,
               const double *const *du_in, const double *xmax_in, hipStream_t st, bool cull,
               const MultiSpec *multi = nullptr) {
  // the single pass's 2D form runs every input slice as six plane slices (one joint histogram per
  // invariant, its own value edges); plane t of input slice s writes output cells t, t + 6, ...
  const bool planes = sp2d(sl_in[0].op, sl_in[0].n_val);
  std::vector<sf_problem> ex_sl;
  std::vector<const double *> ex_x, ex_u;
  std::vector<double> ex_m;
  int n = n_in;
  const sf_problem *sl = sl_in;
  const double *const *dx = dx_in;
  const double *const *du = du_in;
  const double *xmax = xmax_in;
  if (planes) {
    for (int s = 0; s < n_in; ++s) {
      for (int t = 0; t < kSpRows; ++t) {
        sf_problem q = sl_in[s];
        q.val_edges = sl_in[s].val_edges + (size_t)t * (sl_in[s].n_val + 1);
        ex_sl.push_back(q);
        ex_x.push_back(dx_in[s]);
        ex_u.push_back(du_in[s]);
        ex_m.push_back(xmax_in[s]);
      }
    }
    n = kSpRows * n_in;
    sl = ex_sl.data();
    dx = ex_x.data();
    du = ex_u.data();
    xmax = ex_m.data();
  }
  const sf_problem &p0 = sl[0];
  pl.ctx = ctx;
  pl.n = n;
  pl.nout = n_in;
  const bool flat = p0.geometry == SF_GEOM_FLAT;
  pl.G = flat ? kFlat : kSphere;
  pl.D = p0.D;
  pl.geom_code = p0.geometry;
  pl.J = p0.n_val > 0;
  pl.mixed = p0.precision != SF_PREC_FP64;
  pl.ncell = out_cells(p0);  // per input slice
  pl.sp = p0.op == SF_OP_SINGLE_PASS && !planes;
  pl.n1d = multi ? multi->n1d : 0;
  // B-strip: slices that share positions (the auxiliary axes of one field: same x, N, edges) and
  // bin 1D histograms in binary32 evaluate each pair's geometry once for up to kStripMax fields.
  // Opt-in (tuning variant 3): the per-slice 1D kernel with 16 histogram copies outruns it (eight
  // sphere fields, N = 65536: 828 against 609 Gpairs/s), its column velocities being LDS reads.
  pl.strip = 0;
  if (ctx->tune.variant == 3 && p0.precision != SF_PREC_FP64 && p0.n_val == 0 &&
      p0.op != SF_OP_SINGLE_PASS && n_in >= 2) {
    bool same = true;
    for (int s = 1; s < n && same; ++s)
      same = sl[s].x == p0.x && sl[s].N == p0.N && xmax[s] == xmax[0] &&
             memcmp(sl[s].dist_edges, p0.dist_edges, sizeof(double) * (p0.n_dist + 1)) == 0 &&
             (sl[s].dist_rule == nullptr) == (p0.dist_rule == nullptr) &&
             (!p0.dist_rule || memcmp(sl[s].dist_rule, p0.dist_rule, sizeof(sf_bin_rule)) == 0);
    if (same) pl.strip = kStripMax;
  }
  const int strip = pl.strip;
  pl.S64 = rec_d64(pl.G, pl.D);
  pl.F32 = rec_f32(pl.G, pl.D);
  const int G = pl.G, D = pl.D, ncell = pl.ncell, S64 = pl.S64, F32 = pl.F32;
  const bool J = pl.J, mixed = pl.mixed;

  // ---- records
  pl.total_pts = 0;
  pl.max_n = 0;
  pl.total_pairs = 0;
  std::vector<PrepSliceHost> prep(n);
  for (int s = 0; s < n; ++s) {
    prep[s].x = dx[s];
    prep[s].u = du[s];
    prep[s].N = sl[s].N;
    prep[s].pt_off = pl.total_pts;
    prep[s].perm = nullptr;
    prep[s].mul = mixed ? interleave_mul(sl[s].N) : 0;
    pl.total_pts += sl[s].N;
    pl.max_n = std::max<int64_t>(pl.max_n, sl[s].N);
    pl.total_pairs += pairs_of(sl[s].N);
  }
  PLBUF(pl.rec64, ((size_t)pl.total_pts * S64 + 16));
  // + one record (F32 <= 16): the binary32 pair loop requests the record after its last column
  if (mixed) PLBUF(pl.rec32, ((size_t)pl.total_pts * F32 + 16));

  // ---- tables
  std::vector<Slice> sdesc(n);
  std::vector<Cell> cells;
  std::vector<double> dedges, vedges;
  std::vector<float> vedges32;
  int ncell_max = 1;
  for (int s = 0; s < n; ++s) {
    SliceTables T;
    build_tables(sl[s], xmax[s], T);
    Slice d = T.desc;
    d.N = sl[s].N;
    d.pt_off = prep[s].pt_off;
    d.out_off = planes ? (int64_t)(s / kSpRows) * ncell + s % kSpRows : (int64_t)s * ncell;
    d.plane = planes ? s % kSpRows : 0;
    d.xerr = flat ? round_up_f(ldexpl((long double)xmax[s], -28)) : 0.0f;
    d.strip_n = (strip && s % strip == 0) ? std::min(strip, n - s) : 0;
    d.tab_off = (int32_t)cells.size();
    cells.insert(cells.end(), T.tab64.begin(), T.tab64.end());
    d.tab32_off = (int32_t)cells.size();
    cells.insert(cells.end(), T.tab32.begin(), T.tab32.end());
    d.dedge_off = (int32_t)dedges.size();
    dedges.insert(dedges.end(), T.dedges.begin(), T.dedges.end());
    d.vedge_off = (int32_t)vedges.size();
    if (J) {
      for (int k = 0; k <= p0.n_val; ++k) {
        vedges.push_back(sl[s].val_edges[k]);
        vedges32.push_back((float)sl[s].val_edges[k]);
      }
    }
    ncell_max = std::max(ncell_max, std::max(d.ncell64, d.ncell32));
    sdesc[s] = d;
  }
  if (vedges.empty()) { vedges.push_back(0.0); vedges32.push_back(0.0f); }
  pl.ncell_max = ncell_max;
  // fixed-point joint cells (binary32 kernel, one operator, value edges uniform in every slice)
  pl.fx = p0.precision == SF_PREC_MIXED_FIXED && J && !pl.sp;
  for (int s = 0; s < n && pl.fx; ++s) pl.fx = sdesc[s].v_uniform && sdesc[s].v_fx;

  // ---- launch geometry: one histogram per workgroup; large joint histograms get 1024-thread
  // workgroups (one per CU, 16 waves share the table), small 1D histograms 256-thread workgroups
  // with per-lane interleaved copies (several workgroups per CU)
  // launch shape: explicit tuning (sf_set_tuning) > autotuned entry for this key > defaults
  Params Pk{};
  Pk.op = p0.op;
  Pk.nl = p0.nl;
  Pk.nt = p0.nt;
  Pk.n_val = p0.n_val;
  Pk.nmom = pl.sp ? 4 : 1;
  Pk.n1d = pl.n1d;
  const int nmom = Pk.nmom;
  Tuning tn = ctx->tune;
  // the autotuner times one-operator launches; a fused launch (more LDS, more work per pair) keeps
  // the built-in shapes rather than a shape tuned for its main operator alone
  if (!tn.any() && pl.n1d == 0) {
    int64_t in_pairs = 0;
    for (int s = 0; s < n_in; ++s) in_pairs += pairs_of(sl_in[s].N);
    const auto it = ctx->tuned.find(tune_key(sl_in[0], in_pairs));
    if (it != ctx->tuned.end()) tn = it->second;
  }
  const int t_block = tn.block, t_jchunk = tn.j_chunk, t_bpc = tn.bpc, t_variant = tn.variant;
  // fixed-point joint: 768 threads, two workgroups per CU; other joint: 1024 threads, one per CU
  // 1D single pass: 512 threads (its 36-byte slots fill the LDS at two or three workgroups per CU;
  // more waves per workgroup then pay: flat 2D 595 against 492 Gpairs/s, sphere 644 against 529).
  // One-operator 1D kernels keep 256: 512 gained 5 % on flat 2D at N = 262144, nothing at 65536,
  // and lost on small slices (N = 4096: 51 against 88; the 256 ragged slices: 1019 against 1031)
  // Fused 1D outputs without a joint one: 512 threads (their histograms, 16 copies each, fill the
  // LDS at two or three workgroups per CU, like the single pass's)
  int block = t_block ? t_block : (J ? (pl.fx ? 768 : 1024) : ((pl.sp || pl.n1d) ? 512 : 256));
  // column records staged per LDS chunk; the fixed-point joint launch keeps two workgroups per CU
  const int tj = std::min(block, (pl.fx || strip) ? 128 : 256);
  // column records: variant 1 stages them through LDS, variant 2 reads them with scalar loads
  // (binary32 kernel only); 0 picks the default, scalar loads: joint histograms measured 212 vs
  // 204 Gpairs/s at N = 65536 (they also keep the record out of VGPRs, which the two-workgroup
  // fixed-point launch needs); 1D histograms with 16 copies, N = 262144, sphere 974 vs 908 and
  // flat 2D 1158 vs 958 (no LDS stage: more workgroups per CU)
  const bool sload = mixed && !strip && t_variant != 1;
  // a B-strip stages every field's column records; the binary64 kernel stages none (scalar loads)
  const int stage_bytes = (sload || !mixed) ? 0 : tj * F32 * 4 * (strip ? strip : 1);
  int copies = 1;
  if (!J) {
    // per-lane-group copies of a 1D histogram (lanes l and l + copies share one): 16 measured
    // best on both binary32 1D kernels (flat 2D du_L^2 at N = 65536: 636 / 643 / 840 / 708 / 597
    // Gpairs/s for 64 / 32 / 16 / 8 / 4 copies; sphere at N = 262144: 601 / 605 / 890 / 732 / 610).
    // A fused launch without a joint output does not use this histogram (its outputs have their
    // own): one copy.
    copies = pl.n1d ? 1 : (tn.copies ? tn.copies : 16);
    while (copies > 1 &&
           (size_t)smem_layout(p0.n_dist, 0, ncell_max, stage_bytes, copies, nmom, 0, 0, strip).total * 2 > ctx->max_smem)
      copies >>= 1;
  } else if (mixed && tn.copies == 2) {
    copies = 2;  // joint histogram privatised by lane halves (binary32 kernel only)
  }
  const int jlayout = (J && tn.layout) ? tn.layout : 1;
  // fused 1D outputs: 16 lane-group copies of their histograms (lanes l, l + 16, l + 32, l + 48
  // share one), fewer only if the LDS cannot hold them.  Their binary64 LDS atomics bound the
  // fused kernel (one per output and pair): beside the sphere's 64 x 96 joint histogram, 4, 8, 16
  // and 32 copies measured 199, 176, 169 and 169 ms for four 1D outputs at N = 262144
  int mcop = 0;
  if (pl.n1d) {
    mcop = 16;
    while (mcop > 1 && (size_t)smem_layout(p0.n_dist, p0.n_val, ncell_max, stage_bytes, copies, nmom, pl.fx, 0,
                                           strip, jlayout, pl.n1d, mcop).total > ctx->max_smem)
      mcop >>= 1;
  }
  pl.mcopies = mcop;
  const SmemLayout L0 = smem_layout(p0.n_dist, p0.n_val, ncell_max, stage_bytes, copies, nmom, pl.fx, 0, strip,
                                    jlayout, pl.n1d, mcop);
  // the binary64 kernel keeps one copy of a joint histogram
  // (the binary64 kernel reads its column records with scalar loads: no LDS stage)
  const SmemLayout L64 = smem_layout(p0.n_dist, p0.n_val, ncell_max, 0, J ? 1 : copies, nmom, 0, 0, 0,
                                     jlayout);
  if ((size_t)std::max(L0.total, L64.total) > ctx->max_smem)
    return ctx->fail(SF_E_UNSUPPORTED,
                     "histogram of %d cells needs %d bytes of LDS; the device allows %zu per workgroup",
                     ncell, std::max(L0.total, L64.total), ctx->max_smem);
  auto occupancy = [&](int blk, size_t bytes) {
    return strip ? strip_occupancy(G, D, hist_opk(Pk), blk, bytes)
                 : hist_occupancy(G, D, J, mixed, hist_opk(Pk), sload, pl.fx, blk, bytes);
  };
  int bpc = occupancy(block, (size_t)L0.total);
  if (bpc < 1) bpc = 1;
  if (!t_block && pl.fx && bpc * block < 1024) {
    // registers allow only one 768-thread workgroup per CU: use 1024 threads instead
    block = 1024;
    bpc = std::max(1, occupancy(block, (size_t)L0.total));
  }
  // binary32 kernel: the LDS re-bin queue, as large as fits without costing a workgroup per CU
  // (0: every queued pair takes the global queue's atomic)
  int qcap = 0;
  if (mixed) {
    for (int q = kQueueCap; q >= 64 && qcap == 0; q >>= 1) {
      const int t = smem_layout(p0.n_dist, p0.n_val, ncell_max, stage_bytes, copies, nmom, pl.fx, q, strip,
                                jlayout, pl.n1d, mcop).total;
      if ((size_t)t <= ctx->max_smem && occupancy(block, (size_t)t) >= bpc) qcap = q;
    }
  }
  if (t_bpc) bpc = std::min(bpc, t_bpc);
  const int max_blocks = bpc * ctx->n_cu;

  // ---- work items: (slice, i-tile, j-chunk); slices in descending N
  int jchunk = t_jchunk;
  if (jchunk <= 0) {
    const double target_items = 16.0 * max_blocks;
    double per_item = (double)pl.total_pairs / target_items;  // pairs per item
    int jc = (int)(per_item / block);
    jc = std::max(jc, block);
    jc = std::min(jc, 64 * block);
    jchunk = (jc + block - 1) / block * block;
  }
  // fixed-point cells: one item must stay below the flush limit (the LDS histogram is flushed
  // between items only)
  if (pl.fx) jchunk = std::max(1, std::min<int>(jchunk, (int)(kFxFlushLimit / block)));
  // spatial sort and tile culling (binary32 kernel, one-shot calls): column chunks of one tile
  std::vector<CullInfo> cinfo(n);
  pl.culled = false;
  if (cull && mixed && !strip) {
    // the record order of slice s lands at pl.perm + pt_off (used only where culling is active)
    PLBUF(pl.perm, ((size_t)pl.total_pts));
    for (int s = 0; s < n; ++s) {
      const int rc = cull_prepare(ctx, sl[s], dx[s], xmax[s], block, st, cinfo[s], pl.perm.p + prep[s].pt_off);
      if (rc) return rc;
      if (cinfo[s].active) prep[s].perm = pl.perm.p + prep[s].pt_off;
    }
  }
  PLBUF(pl.prep, (n));
  HIPCHK(hipMemcpyAsync(pl.prep.p, prep.data(), sizeof(PrepSliceHost) * n, hipMemcpyHostToDevice, st));

  std::vector<int> order(n);
  for (int s = 0; s < n; ++s) order[s] = s;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return sl[a].N > sl[b].N; });
  std::vector<Item> items, items64;
  // fixed-point cells: a workgroup flushes its LDS histogram only between items, so one item may
  // add at most kFxFlushLimit updates (block rows x its columns).  A culled item spans one column
  // tile of `block` columns; at block 1024 that is 2^20 updates, so such tiles are split in halves
  // (the tile of a column is still j0 / block).
  int64_t jc_cull = block;
  while (pl.fx && jc_cull > 1 && (int64_t)block * jc_cull > kFxFlushLimit) jc_cull >>= 1;
  for (int s : order) {
    const int64_t N = sl[s].N;
    const CullInfo &ci = cinfo[s];
    const int64_t jc = ci.active ? jc_cull : jchunk;
    const size_t first = items.size();
    for (int64_t i0 = 0; i0 < N - 1; i0 += block) {
      for (int64_t j0 = i0; j0 < N; j0 += jc) {
        if (ci.active && j0 > i0 && boxes_apart(ci.tiles[i0 / block], ci.tiles[j0 / block], ci.limit)) {
          pl.culled = true;
          continue;
        }
        Item it;
        it.slice = s;
        it.i0 = (int32_t)i0;
        it.j0 = (int32_t)j0;
        it.j1 = (int32_t)std::min<int64_t>(N, j0 + jc);
        // a B-strip's items belong to its leading slice
        if (!strip || s % strip == 0) items.push_back(it);
        if (strip) items64.push_back(it);
      }
    }
    // full items of the slice first (a diagonal item costs its full rectangle): the short last
    // column range of every row tile and the short last row tile end the queue, so workgroups
    // finish together (a linear-time partition; a one-shot call builds its items every time)
    const int64_t rows = block;
    std::stable_partition(items.begin() + first, items.end(), [N, rows, jc](const Item &a) {
      return N - a.i0 >= rows && a.j1 - a.j0 == jc;
    });
  }
  if (pl.fx) {
    // every item must fit the fixed-point cells' exact range (checked, not assumed)
    for (const Item &it : items)
      if ((int64_t)block * (int64_t)(it.j1 - it.j0) > kFxFlushLimit)
        return ctx->fail(SF_E_DEVICE, "internal: work item of %d x %d pair updates exceeds the fixed-point flush limit",
                         block, it.j1 - it.j0);
  }
  if (items.empty()) items.push_back(Item{0, 0, 0, 0});
  if (strip && items64.empty()) items64.push_back(Item{0, 0, 0, 0});
  pl.n_items64 = (int)items64.size();
  pl.n_items = (int)items.size();
  pl.grid = std::max(1, std::min(pl.n_items, max_blocks));
  pl.block = block;
  pl.tj = tj;
  pl.copies = copies;
  pl.sload = sload;

  // ---- upload
  PLBUF(pl.slices, (n));
  PLBUF(pl.items, (items.size()));
  PLBUF(pl.cells, (cells.size()));
  PLBUF(pl.dedges, (dedges.size()));
  PLBUF(pl.vedges, (vedges.size()));
  PLBUF(pl.vedges32, (vedges32.size()));
  PLBUF(pl.acc_s, ((size_t)pl.nout * ncell));
  PLBUF(pl.acc_c, ((size_t)pl.nout * ncell));
  if (pl.n1d) {
    PLBUF(pl.msums, ((size_t)pl.n1d * pl.nout * p0.n_dist));
    PLBUF(pl.mcounts, ((size_t)pl.n1d * pl.nout * p0.n_dist));
  }
  PLBUF(pl.counter, (4));
  PLBUF(pl.pcount, (4));
  PLBUF(pl.flag, (4));
  HIPCHK(hipMemcpyAsync(pl.slices.p, sdesc.data(), sizeof(Slice) * n, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(pl.items.p, items.data(), sizeof(Item) * items.size(), hipMemcpyHostToDevice, st));
  if (strip) {
    PLBUF(pl.items64, (items64.size()));
    HIPCHK(hipMemcpyAsync(pl.items64.p, items64.data(), sizeof(Item) * items64.size(), hipMemcpyHostToDevice, st));
  }
  HIPCHK(hipMemcpyAsync(pl.cells.p, cells.data(), sizeof(Cell) * cells.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(pl.dedges.p, dedges.data(), sizeof(double) * dedges.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(pl.vedges.p, vedges.data(), sizeof(double) * vedges.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(pl.vedges32.p, vedges32.data(), sizeof(float) * vedges32.size(), hipMemcpyHostToDevice, st));
  // host vectors go out of scope: make the asynchronous copies complete first
  HIPCHK(hipStreamSynchronize(st));

  pl.cap = 0;
  if (mixed) {
    pl.cap = std::max<int64_t>(1LL << 22, pl.total_pairs / 2000);
    pl.cap = std::min<int64_t>(pl.cap, 1LL << 27);
    pl.cap = std::max<int64_t>(pl.cap, std::min<int64_t>(ctx->queue_hint, kQueueMax));
    if (pl.arena) {
      // the re-bin queue gets the rest of the scratch (an overflow is redone in binary64)
      const size_t off = (pl.arena->used + 255) & ~(size_t)255;
      const int64_t room = off < pl.arena->cap ? (int64_t)((pl.arena->cap - off) / sizeof(Pending)) : 0;
      pl.cap = std::min<int64_t>(pl.cap, room);
      if (pl.cap < 1024)
        return ctx->fail(SF_E_UNSUPPORTED,
                         "this call leaves no room for the re-bin queue in the allocation-free entry's "
                         "%llu MiB scratch; use sf_hist_batch_device", (unsigned long long)(SF_NOALLOC_SCRATCH_BYTES >> 20));
    }
    PLBUF(pl.pending, ((size_t)pl.cap));
  }

  Params &P = pl.P;
  P = Params{};
  P.slices = pl.slices.p;
  P.items = pl.items.p;
  P.n_items = pl.n_items;
  P.block_rows = block;
  P.counter = pl.counter.p;
  P.rec64 = pl.rec64.p;
  P.rec32 = pl.rec32.p;
  P.cells = pl.cells.p;
  P.dedges = pl.dedges.p;
  P.vedges = pl.vedges.p;
  P.vedges32 = pl.vedges32.p;
  P.sums = pl.acc_s.p;
  P.counts = pl.acc_c.p;
  P.pending = pl.pending.p;
  P.pending_count = pl.pcount.p;
  P.pending_cap = pl.cap;
  P.stats = nullptr;
  P.ncell_max = ncell_max;
  P.tj = tj;
  P.copies = copies;
  P.jlayout = jlayout;
  P.stage_bytes = stage_bytes;
  P.n_dist = p0.n_dist;
  P.n_val = p0.n_val;
  P.op = p0.op;
  P.nl = p0.nl;
  P.nt = p0.nt;
  P.order = op_order(p0);
  P.radius = (p0.geometry == SF_GEOM_SPHERE_DEG) ? p0.radius : 1.0;
  // |dv| <= vcoef * (|u_i| + |u_j|)^order * amp  (flat amp = 1; sphere amp = 1 + c^2 / sin(sigma))
  P.vcoef = (float)ldexp(16.0 * std::max(1, P.order), -24);
  P.nmom = nmom;
  P.fx = pl.fx ? 1 : 0;
  P.qcap = qcap;
  P.qw = qcap / (block / 64);
  P.bl = (flat && p0.semantics == SF_SEM_BATCH) ? 1 : 0;
  P.strip = strip;
  P.n1d = pl.n1d;
  P.mcopies = mcop;
  P.n_slices = pl.nout;
  P.mab_a = P.mab_b = 0;
  for (int k = 0; k < pl.n1d; ++k) {
    P.mop[k] = multi->op[k];
    P.mnl[k] = multi->nl[k];
    P.mnt[k] = multi->nt[k];
    int a = 0, b = 0;
    if (!fused_basis_index(multi->op[k], multi->nl[k], multi->nt[k], &a, &b))
      return ctx->fail(SF_E_DEVICE, "internal: fused output %d is not a basis product", k);
    P.mab_a |= (uint32_t)a << (4 * k);
    P.mab_b |= (uint32_t)b << (4 * k);
  }  // unused outputs: fields 0, 1 * 1
  P.ja = P.jb = 0;
  if (pl.n1d && J && !fused_basis_index(p0.op, p0.nl, p0.nt, &P.ja, &P.jb))
    return ctx->fail(SF_E_DEVICE, "internal: the fused joint operator is not a basis product");
  P.msums = pl.n1d ? pl.msums.p : nullptr;
  P.mcounts = pl.n1d ? pl.mcounts.p : nullptr;
  P.flush_limit = pl.fx ? kFxFlushLimit : 0x7fffffffLL;
  return SF_OK;
}

// Kernel-only launch sequence: records from the current x/u, zeroed accumulators, histogram,
// re-bin (mixed), merge into out (skipped, with *flag set, if the re-bin queue overflowed).
// ev (optional) brackets the histogram and re-bin kernels for timing.
int plan_launch(Plan &pl, bool use_mixed, double *out_s, unsigned long long *out_c, hipStream_t st,
                hipEvent_t *ev) {
  sf_ctx *ctx = pl.ctx;
  if (pl.total_pairs == 0) return SF_OK;
  const int64_t ncells = (int64_t)pl.nout * pl.ncell;
  HIPCHK(launch_prep(pl.G, pl.D, use_mixed, pl.geom_code, pl.prep.p, pl.n, pl.max_n, pl.rec64.p,
                     pl.rec32.p, st));
  const int64_t mcells = (int64_t)pl.n1d * pl.nout * pl.P.n_dist;
  if (pl.n1d && !use_mixed)  // fused outputs exist in the binary32 kernel only
    return ctx->fail(SF_E_UNSUPPORTED, "internal: fused outputs launched on the binary64 kernel");
  {
    // accumulators, the work counter, the re-bin queue count and the overflow flag: one launch
    void *zp[kZeroSetMax] = {pl.acc_s.p, pl.acc_c.p, pl.counter.p, pl.pcount.p, pl.flag.p};
    int64_t zb[kZeroSetMax] = {(int64_t)sizeof(double) * ncells, (int64_t)sizeof(unsigned long long) * ncells,
                               (int64_t)sizeof(int) * 4, (int64_t)sizeof(unsigned long long) * 4,
                               (int64_t)sizeof(int) * 4};
    int nz = 5;
    if (pl.n1d) {
      zp[nz] = pl.msums.p;
      zb[nz++] = (int64_t)sizeof(double) * mcells;
      zp[nz] = pl.mcounts.p;
      zb[nz++] = (int64_t)sizeof(unsigned long long) * mcells;
    }
    HIPCHK(launch_zero_multi(nz, zp, zb, st));
  }
  const bool use_sload = use_mixed && pl.sload;
  const int stage = (use_sload || !use_mixed) ? 0 : pl.tj * pl.F32 * 4;
  Params P = pl.P;
  P.stage_bytes = stage;
  if (!use_mixed) {  // the binary64 kernel keeps binary64 sums and 32-bit counts, no queue
    P.fx = 0;
    if (P.n_val > 0) P.copies = 1;
    P.qcap = 0;
    P.qw = 0;
    P.flush_limit = 0x7fffffffLL;
    if (pl.strip) {  // one slice at a time
      P.strip = 0;
      P.items = pl.items64.p;
      P.n_items = pl.n_items64;
    }
  }
  const int stage_mixed = stage * ((use_mixed && pl.strip) ? pl.strip : 1);
  P.stage_bytes = stage_mixed;
  const SmemLayout L = smem_layout(P.n_dist, P.n_val, P.ncell_max, stage_mixed, P.copies, P.nmom, P.fx, P.qcap,
                                   P.strip, P.jlayout, P.n1d, P.mcopies);
  if (ev) HIPCHK(hipEventRecord(ev[0], st));
  HIPCHK(launch_hist(pl.G, pl.D, pl.J, use_mixed, use_sload, P, pl.grid, pl.block, (size_t)L.total, st));
  if (ev) HIPCHK(hipEventRecord(ev[1], st));
  // four 256-thread workgroups per CU: the re-bin is a latency-bound gather (2D single pass,
  // 2.7e8 queued pairs: 22.6 / 16.7 / 19.6 ms with one / four / eight per CU)
  if (use_mixed) HIPCHK(launch_recheck(pl.G, pl.D, pl.J, P, 4 * ctx->n_cu, st));
  if (ev) HIPCHK(hipEventRecord(ev[2], st));
  // (a fused call without a joint output passes no out_s / out_c)
  if (out_s)
    HIPCHK(launch_merge_checked(pl.acc_s.p, pl.acc_c.p, out_s, out_c, ncells,
                                use_mixed ? pl.pcount.p : nullptr, pl.cap, pl.flag.p, st));
  const int64_t per = (int64_t)pl.nout * pl.P.n_dist;
  for (int k = 0; k < pl.n1d; ++k)
    HIPCHK(launch_merge_checked(pl.msums.p + k * per, pl.mcounts.p + k * per, pl.mdst_s[k], pl.mdst_c[k], per,
                                pl.pcount.p, pl.cap, pl.flag.p, st));
  if (pl.sp) HIPCHK(launch_derive_sp(out_s, out_c, (int64_t)pl.nout * pl.P.n_dist, st));
  if (ev) HIPCHK(hipEventRecord(ev[3], st));
  return SF_OK;
}

// One-shot: build (into the context's cached plan, reusing its buffers), launch, and on a
// re-bin queue overflow redo the call with the binary64 kernel.
int run_batch(sf_ctx *ctx, int n, const sf_problem *sl, const double *const *dx,
              const double *const *du, const double *xmax, double *out_s,
              unsigned long long *out_c, hipStream_t st) {
  Plan &pl = ctx->oneshot;
  int rc = plan_build(ctx, pl, n, sl, dx, du, xmax, st, true);
  if (rc) return rc;
  for (int k = 0; k < 8; ++k) ctx->stats[k] = 0;
  ctx->stats[0] = pl.total_pairs;
  ctx->stats[2] = pl.n_items;
  ctx->stats[3] = pl.grid;
  ctx->stats[7] = pl.block;
  ctx->ev_pending = false;
  if (pl.total_pairs == 0) return SF_OK;
  for (int k = 0; k < 4; ++k)
    if (!ctx->ev[k]) HIPCHK(hipEventCreate(&ctx->ev[k]));
  rc = plan_launch(pl, pl.mixed, out_s, out_c, st, ctx->ev);
  if (rc) return rc;
  if (pl.mixed) {
    unsigned long long pc = 0;
    HIPCHK(hipMemcpyAsync(&pc, pl.pcount.p, sizeof(pc), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    ctx->stats[1] = (int64_t)pc;
    if ((int64_t)pc > pl.cap && (int64_t)pc <= kQueueMax) {
      // the merge was skipped: the queue now knows its size, so grow it and run the binary32
      // kernel once more (and remember the size for later calls of this context)
      ctx->queue_hint = std::max<int64_t>(ctx->queue_hint, std::min<int64_t>((int64_t)pc + pc / 4, kQueueMax));
      pl.cap = ctx->queue_hint;
      HIPCHK(pl.pending.ensure((size_t)pl.cap));
      pl.P.pending = pl.pending.p;
      pl.P.pending_cap = pl.cap;
      rc = plan_launch(pl, true, out_s, out_c, st, ctx->ev);
      if (rc) return rc;
      HIPCHK(hipMemcpyAsync(&pc, pl.pcount.p, sizeof(pc), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      ctx->stats[1] = (int64_t)pc;
    }
    if ((int64_t)pc > pl.cap) {
      // the merge was skipped; redo everything with the binary64 kernel (which evaluates every
      // pair in the caller's order: rebuild the work items without culling first)
      if (pl.culled) {
        rc = plan_build(ctx, pl, n, sl, dx, du, xmax, st, false);
        if (rc) return rc;
      }
      rc = plan_launch(pl, false, out_s, out_c, st, ctx->ev);
      if (rc) return rc;
    }
  }
  ctx->ev_pending = true;
  return SF_OK;
}

int host_entry(sf_ctx *ctx, int n, const sf_problem *sl, double *sums, uint64_t *counts) {
  ctx->err.clear();
  int rc = validate_batch(ctx, n, sl, true);
  if (rc) return rc;
  if (n == 0) return SF_OK;
  if (!sums || !counts) return ctx->fail(SF_E_ARG, "null sums or counts");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const sf_problem &p0 = sl[0];
  const int ncell = out_cells(p0);
  int64_t tx = 0, tu = 0;
  for (int s = 0; s < n; ++s) {
    tx += (int64_t)sl[s].W * sl[s].N;
    tu += (int64_t)sl[s].D * sl[s].N;
  }
  HIPCHK(ctx->d_x.ensure((size_t)tx + 2));
  HIPCHK(ctx->d_u.ensure((size_t)tu + 2));
  std::vector<const double *> px(n), pu(n);
  std::vector<double> xmax(n, 0.0);
  int64_t ox = 0, ou = 0;
  for (int s = 0; s < n; ++s) {
    const int64_t nx = (int64_t)sl[s].W * sl[s].N, nu = (int64_t)sl[s].D * sl[s].N;
    if (nx) HIPCHK(hipMemcpyAsync(ctx->d_x.p + ox, sl[s].x, sizeof(double) * nx, hipMemcpyHostToDevice, st));
    if (nu) HIPCHK(hipMemcpyAsync(ctx->d_u.p + ou, sl[s].u, sizeof(double) * nu, hipMemcpyHostToDevice, st));
    px[s] = ctx->d_x.p + ox;
    pu[s] = ctx->d_u.p + ou;
    double m = 0;
    for (int64_t k = 0; k < nx; ++k) m = std::max(m, fabs(sl[s].x[k]));
    xmax[s] = m;
    ox += nx;
    ou += nu;
  }
  // outputs: accumulate into zeroed device staging, then add on the host
  DevBuf<double> os;
  DevBuf<unsigned long long> oc;
  HIPCHK(os.ensure((size_t)n * ncell));
  HIPCHK(oc.ensure((size_t)n * ncell));
  HIPCHK(launch_zero(os.p, sizeof(double) * (int64_t)n * ncell, st));
  HIPCHK(launch_zero(oc.p, sizeof(unsigned long long) * (int64_t)n * ncell, st));
  rc = run_batch(ctx, n, sl, px.data(), pu.data(), xmax.data(), os.p, oc.p, st);
  if (rc) { os.release(); oc.release(); return rc; }
  std::vector<double> hs((size_t)n * ncell);
  std::vector<unsigned long long> hc((size_t)n * ncell);
  hipError_t e1 = hipMemcpyAsync(hs.data(), os.p, sizeof(double) * hs.size(), hipMemcpyDeviceToHost, st);
  hipError_t e2 = hipMemcpyAsync(hc.data(), oc.p, sizeof(unsigned long long) * hc.size(), hipMemcpyDeviceToHost, st);
  hipError_t e3 = hipStreamSynchronize(st);
  os.release();
  oc.release();
  if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess)
    return ctx->fail(SF_E_DEVICE, "device to host copy failed: %s",
                     hipGetErrorString(e3 != hipSuccess ? e3 : (e1 != hipSuccess ? e1 : e2)));
  for (size_t k = 0; k < hs.size(); ++k) {
    sums[k] += hs[k];
    counts[k] += hc[k];
  }
  if (p0.op == SF_OP_SINGLE_PASS && p0.n_val == 0) derive_sp_host(sums, counts, (int64_t)n * p0.n_dist);
  return SF_OK;
}

}  // namespace

extern "C" int sf_hist(sf_ctx *ctx, const sf_problem *p, double *sums, uint64_t *counts) {
  if (!ctx) return SF_E_ARG;
  return host_entry(ctx, 1, p, sums, counts);
}

extern "C" int sf_hist_batch(sf_ctx *ctx, int32_t n_slices, const sf_problem *slices, double *sums,
                             uint64_t *counts) {
  if (!ctx) return SF_E_ARG;
  return host_entry(ctx, n_slices, slices, sums, counts);
}

extern "C" int sf_hist_batch_device(sf_ctx *ctx, int32_t n, const sf_problem *sl, double *d_sums,
                                    uint64_t *d_counts, void *stream) {
  if (!ctx) return SF_E_ARG;
  ctx->err.clear();
  int rc = validate_batch(ctx, n, sl, false);
  if (rc) return rc;
  if (n == 0) return SF_OK;
  if (!d_sums || !d_counts) return ctx->fail(SF_E_ARG, "null sums or counts");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  std::vector<const double *> px(n), pu(n);
  std::vector<double> xmax(n, 0.0);
  for (int s = 0; s < n; ++s) {
    px[s] = sl[s].x;
    pu[s] = sl[s].u;
    // coordinate magnitude for the binary32 margin: read x back (O(N)) only for the mixed flat path
    if (sl[s].precision != SF_PREC_FP64 && sl[s].geometry == SF_GEOM_FLAT && sl[s].N > 0) {
      std::vector<double> hx((size_t)sl[s].W * sl[s].N);
      HIPCHK(hipMemcpyAsync(hx.data(), sl[s].x, sizeof(double) * hx.size(), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      double m = 0;
      for (double v : hx) m = std::max(m, fabs(v));
      xmax[s] = m;
    }
  }
  return run_batch(ctx, n, sl, px.data(), pu.data(), xmax.data(), d_sums,
                   (unsigned long long *)d_counts, st);
}

extern "C" int sf_hist_device(sf_ctx *ctx, const sf_problem *p, double *d_sums, uint64_t *d_counts,
                              void *stream) {
  return sf_hist_batch_device(ctx, 1, p, d_sums, d_counts, stream);
}

// ----------------------------------------------------------------------------------------------
// fused outputs: one pair evaluation, several operators
// ----------------------------------------------------------------------------------------------
// What a fused call does with its outputs: each output's own problems (validated as its own call
// would be), the joint output fused into the launch (-1: none), the fused 1D outputs in launch
// order, and the outputs that run as their own calls (binary64 precision: all of them; else those
// whose binary32 value is not a product of two basis values, FullVector{1} and {3}).
struct FusedSetup {
  std::vector<std::vector<sf_problem>> per;
  int joint = -1;
  std::vector<int> ids;
  std::vector<int> separate;
  MultiSpec ms;
  const std::vector<sf_problem> &main() const { return per[joint >= 0 ? joint : ids[0]]; }
  bool launch() const { return joint >= 0 || !ids.empty(); }
};

int fused_setup(sf_ctx *ctx, int32_t n, const sf_problem *sl, int32_t n_out, const sf_output *outs,
                FusedSetup &fs) {
  if (n_out < 1 || !outs) return ctx->fail(SF_E_ARG, "a fused call needs at least one output");
  int joint = -1, n1 = 0;
  for (int k = 0; k < n_out; ++k) {
    if (outs[k].op == SF_OP_SINGLE_PASS)
      return ctx->fail(SF_E_ARG, "output %d: the single pass is not a fused output (it is its own fused form)", k);
    if (outs[k].n_val > 0) {
      if (joint >= 0) return ctx->fail(SF_E_UNSUPPORTED, "a fused call takes at most one joint output");
      joint = k;
    } else {
      ++n1;
    }
  }
  if (n1 > SF_MAX_FUSED) return ctx->fail(SF_E_UNSUPPORTED, "a fused call takes at most %d 1D outputs", SF_MAX_FUSED);
  fs.per.assign((size_t)n_out, {});
  for (int k = 0; k < n_out; ++k) {
    fs.per[k].assign(sl, sl + std::max(n, 0));
    for (sf_problem &q : fs.per[k]) {
      q.op = outs[k].op;
      q.nl = outs[k].nl;
      q.nt = outs[k].nt;
      q.n_val = outs[k].n_val;
      if (outs[k].n_val == 0 || outs[k].val_edges) {  // else each slice's own value edges
        q.val_edges = outs[k].val_edges;
        q.val_rule = outs[k].val_rule;
      }
    }
    const int rc = validate_batch(ctx, n, fs.per[k].data(), false);
    if (rc) return rc;
  }
  fs.joint = -1;
  fs.ids.clear();
  fs.separate.clear();
  fs.ms = MultiSpec();
  if (n > 0 && sl[0].precision == SF_PREC_FP64) {
    for (int k = 0; k < n_out; ++k) fs.separate.push_back(k);
    return SF_OK;
  }
  int a = 0, b = 0;
  if (joint >= 0) {
    if (fused_basis_index(outs[joint].op, outs[joint].nl, outs[joint].nt, &a, &b)) fs.joint = joint;
    else fs.separate.push_back(joint);
  }
  for (int k = 0; k < n_out; ++k) {
    if (k == joint) continue;
    if (!fused_basis_index(outs[k].op, outs[k].nl, outs[k].nt, &a, &b)) {
      fs.separate.push_back(k);
      continue;
    }
    fs.ms.op[fs.ms.n1d] = outs[k].op;
    fs.ms.nl[fs.ms.n1d] = outs[k].nl;
    fs.ms.nt[fs.ms.n1d] = outs[k].nt;
    ++fs.ms.n1d;
    fs.ids.push_back(k);
  }
  return SF_OK;
}

// coordinate magnitudes of flat slices (the binary32 kernels' margins), read from the device
int slice_xmax(sf_ctx *ctx, int n, const sf_problem *sl, hipStream_t st, double scale, std::vector<double> &xmax) {
  xmax.assign((size_t)n, 0.0);
  for (int s = 0; s < n; ++s) {
    if (sl[s].geometry == SF_GEOM_FLAT && sl[s].N > 0) {
      std::vector<double> hx((size_t)sl[s].W * sl[s].N);
      HIPCHK(hipMemcpyAsync(hx.data(), sl[s].x, sizeof(double) * hx.size(), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      double m = 0;
      for (double v : hx) m = std::max(m, fabs(v));
      xmax[s] = scale * m;
    }
  }
  return SF_OK;
}

extern "C" int sf_hist_multi_device(sf_ctx *ctx, int32_t n, const sf_problem *sl, int32_t n_out,
                                    const sf_output *outs, double *const *d_sums, uint64_t *const *d_counts,
                                    void *stream) {
  if (!ctx) return SF_E_ARG;
  ctx->err.clear();
  if (n_out < 1 || !outs || !d_sums || !d_counts) return ctx->fail(SF_E_ARG, "a fused call needs at least one output");
  for (int k = 0; k < n_out; ++k)
    if (!d_sums[k] || !d_counts[k]) return ctx->fail(SF_E_ARG, "output %d: null sums or counts", k);
  FusedSetup fs;
  int rc = fused_setup(ctx, n, sl, n_out, outs, fs);
  if (rc) return rc;
  if (n == 0) return SF_OK;
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  for (int k : fs.separate) {
    rc = sf_hist_batch_device(ctx, n, fs.per[k].data(), d_sums[k], d_counts[k], st);
    if (rc) return rc;
  }
  if (!fs.launch()) return SF_OK;
  const std::vector<int> &separate = fs.separate;
  const std::vector<std::vector<sf_problem>> &per = fs.per;
  const int joint = fs.joint;
  Plan &pl = ctx->fused;
  for (size_t j = 0; j < fs.ids.size(); ++j) {
    pl.mdst_s[j] = d_sums[fs.ids[j]];
    pl.mdst_c[j] = (unsigned long long *)d_counts[fs.ids[j]];
  }
  // the plan's main problem: the joint output's (its value bins), or the first 1D output's,
  // whose own histogram then stays unused
  const std::vector<sf_problem> &main = fs.main();
  std::vector<const double *> px(n), pu(n);
  std::vector<double> xmax;
  for (int s = 0; s < n; ++s) {
    px[s] = sl[s].x;
    pu[s] = sl[s].u;
  }
  rc = slice_xmax(ctx, n, sl, st, 1.0, xmax);
  if (rc) return rc;
  rc = plan_build(ctx, pl, n, main.data(), px.data(), pu.data(), xmax.data(), st, true, &fs.ms);
  if (rc) return rc;
  for (int k = 0; k < 8; ++k) ctx->stats[k] = 0;
  ctx->stats[0] = pl.total_pairs;
  ctx->stats[2] = pl.n_items;
  ctx->stats[3] = pl.grid;
  ctx->stats[7] = pl.block;
  ctx->ev_pending = false;
  if (pl.total_pairs == 0) return SF_OK;
  for (int k = 0; k < 4; ++k)
    if (!ctx->ev[k]) HIPCHK(hipEventCreate(&ctx->ev[k]));
  double *js = joint >= 0 ? d_sums[joint] : nullptr;
  unsigned long long *jc = joint >= 0 ? (unsigned long long *)d_counts[joint] : nullptr;
  unsigned long long pc = 0;
  for (int attempt = 0; attempt < 2; ++attempt) {
    rc = plan_launch(pl, true, js, jc, st, ctx->ev);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(&pc, pl.pcount.p, sizeof(pc), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    ctx->stats[1] = (int64_t)pc;
    if ((int64_t)pc <= pl.cap || (int64_t)pc > kQueueMax) break;
    // the merges were skipped: grow the re-bin queue to the size reached and run again
    ctx->queue_hint = std::max<int64_t>(ctx->queue_hint, std::min<int64_t>((int64_t)pc + pc / 4, kQueueMax));
    pl.cap = ctx->queue_hint;
    HIPCHK(pl.pending.ensure((size_t)pl.cap));
    pl.P.pending = pl.pending.p;
    pl.P.pending_cap = pl.cap;
  }
  ctx->ev_pending = true;
  if ((int64_t)pc > pl.cap) {
    // still too many pairs to re-bin: every fused output with the binary64 kernel, one after
    // the other (the separate outputs are done)
    for (int k = 0; k < n_out; ++k) {
      if (std::find(separate.begin(), separate.end(), k) != separate.end()) continue;
      std::vector<sf_problem> q = per[k];
      for (sf_problem &e : q) e.precision = SF_PREC_FP64;
      rc = sf_hist_batch_device(ctx, n, q.data(), d_sums[k], d_counts[k], st);
      if (rc) return rc;
    }
  }
  return SF_OK;
}

// Host-pointer form: x, u and the outputs are host memory (staged like sf_hist_batch).
extern "C" int sf_hist_multi(sf_ctx *ctx, int32_t n, const sf_problem *sl, int32_t n_out, const sf_output *outs,
                             double *const *sums, uint64_t *const *counts) {
  if (!ctx) return SF_E_ARG;
  ctx->err.clear();
  if (n < 0 || (n > 0 && !sl)) return ctx->fail(SF_E_ARG, "bad slices");
  if (n_out < 1 || n_out > SF_MAX_FUSED + 1 || !outs || !sums || !counts)
    return ctx->fail(SF_E_ARG, "a fused call needs 1 to %d outputs", SF_MAX_FUSED + 1);
  {
    // the host arrays are checked here; each output's operator and bins by sf_hist_multi_device
    std::vector<sf_problem> chk(sl, sl + n);
    for (sf_problem &q : chk) {
      q.op = outs[0].op;
      q.nl = outs[0].nl;
      q.nt = outs[0].nt;
      q.n_val = outs[0].n_val;
      if (outs[0].n_val == 0 || outs[0].val_edges) {
        q.val_edges = outs[0].val_edges;
        q.val_rule = outs[0].val_rule;
      }
    }
    const int rc = validate_batch(ctx, n, chk.data(), true);
    if (rc) return rc;
  }
  if (n == 0) return SF_OK;
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  int64_t tx = 0, tu = 0;
  for (int s = 0; s < n; ++s) {
    tx += (int64_t)sl[s].W * sl[s].N;
    tu += (int64_t)sl[s].D * sl[s].N;
  }
  HIPCHK(ctx->d_x.ensure((size_t)tx + 2));
  HIPCHK(ctx->d_u.ensure((size_t)tu + 2));
  std::vector<sf_problem> dev(sl, sl + n);
  int64_t ox = 0, ou = 0;
  for (int s = 0; s < n; ++s) {
    const int64_t nx = (int64_t)sl[s].W * sl[s].N, nu = (int64_t)sl[s].D * sl[s].N;
    if (nx) HIPCHK(hipMemcpyAsync(ctx->d_x.p + ox, sl[s].x, sizeof(double) * nx, hipMemcpyHostToDevice, st));
    if (nu) HIPCHK(hipMemcpyAsync(ctx->d_u.p + ou, sl[s].u, sizeof(double) * nu, hipMemcpyHostToDevice, st));
    dev[s].x = ctx->d_x.p + ox;
    dev[s].u = ctx->d_u.p + ou;
    ox += nx;
    ou += nu;
  }
  std::vector<DevBuf<double>> os((size_t)n_out);
  std::vector<DevBuf<unsigned long long>> oc((size_t)n_out);
  std::vector<double *> ds((size_t)n_out);
  std::vector<uint64_t *> dc((size_t)n_out);
  std::vector<int64_t> cells((size_t)n_out);
  int rc = SF_OK;
  for (int k = 0; k < n_out && !rc; ++k) {
    cells[k] = (int64_t)n * sl[0].n_dist * (outs[k].n_val > 0 ? outs[k].n_val : 1);
    if (os[k].ensure((size_t)cells[k]) != hipSuccess || oc[k].ensure((size_t)cells[k]) != hipSuccess)
      rc = ctx->fail(SF_E_DEVICE, "output buffers");
    else if (launch_zero(os[k].p, sizeof(double) * cells[k], st) != hipSuccess ||
             launch_zero(oc[k].p, sizeof(unsigned long long) * cells[k], st) != hipSuccess)
      rc = ctx->fail(SF_E_DEVICE, "output zeroing");
    ds[k] = os[k].p;
    dc[k] = (uint64_t *)oc[k].p;
  }
  if (!rc) rc = sf_hist_multi_device(ctx, n, dev.data(), n_out, outs, ds.data(), dc.data(), st);
  for (int k = 0; k < n_out && !rc; ++k) {
    std::vector<double> hs((size_t)cells[k]);
    std::vector<unsigned long long> hc((size_t)cells[k]);
    hipError_t e1 = hipMemcpyAsync(hs.data(), os[k].p, sizeof(double) * hs.size(), hipMemcpyDeviceToHost, st);
    hipError_t e2 = hipMemcpyAsync(hc.data(), oc[k].p, sizeof(unsigned long long) * hc.size(), hipMemcpyDeviceToHost, st);
    hipError_t e3 = hipStreamSynchronize(st);
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
      rc = ctx->fail(SF_E_DEVICE, "device to host copy failed");
      break;
    }
    for (int64_t c = 0; c < cells[k]; ++c) {
      sums[k][c] += hs[c];
      counts[k][c] += hc[c];
    }
  }
  for (int k = 0; k < n_out; ++k) {
    os[k].release();
    oc[k].release();
  }
  return rc;
}

// ----------------------------------------------------------------------------------------------
// allocation-free entry: every device buffer of the call comes from one static array per device
// ----------------------------------------------------------------------------------------------
extern "C" int sf_hist_batch_device_noalloc(sf_ctx *ctx, int32_t n, const sf_problem *sl, double *d_sums,
                                            uint64_t *d_counts, void *stream) {
  if (!ctx) return SF_E_ARG;
  ctx->err.clear();
  int rc = validate_batch(ctx, n, sl, false);
  if (rc) return rc;
  if (n == 0) return SF_OK;
  if (!d_sums || !d_counts) return ctx->fail(SF_E_ARG, "null sums or counts");
  if (!ctx->scratch_ptr) return ctx->fail(SF_E_DEVICE, "the scratch array has no address on this device");
  // the records alone must fit (checked before any device work; the tables, items, accumulators
  // and queue are checked as the plan takes them)
  int64_t pts = 0;
  for (int s = 0; s < n; ++s) pts += sl[s].N;
  const int G = sl[0].geometry == SF_GEOM_FLAT ? kFlat : kSphere;
  const int64_t rec_bytes = pts * (8 * rec_d64(G, sl[0].D) + (sl[0].precision != SF_PREC_FP64 ? 4 * rec_f32(G, sl[0].D) : 0));
  if (rec_bytes > (int64_t)(SF_NOALLOC_SCRATCH_BYTES / 2))
    return ctx->fail(SF_E_UNSUPPORTED,
                     "%lld points need %lld bytes of records, more than half of the allocation-free entry's "
                     "%llu MiB scratch; use sf_hist_batch_device", (long long)pts, (long long)rec_bytes,
                     (unsigned long long)(SF_NOALLOC_SCRATCH_BYTES >> 20));
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  std::vector<const double *> px(n), pu(n);
  std::vector<double> xmax(n, 0.0);
  for (int s = 0; s < n; ++s) {
    px[s] = sl[s].x;
    pu[s] = sl[s].u;
    if (sl[s].precision != SF_PREC_FP64 && sl[s].geometry == SF_GEOM_FLAT && sl[s].N > 0) {
      std::vector<double> hx((size_t)sl[s].W * sl[s].N);
      HIPCHK(hipMemcpyAsync(hx.data(), sl[s].x, sizeof(double) * hx.size(), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      double m = 0;
      for (double v : hx) m = std::max(m, fabs(v));
      xmax[s] = m;
    }
  }
  Plan &pl = ctx->noalloc;
  Arena arena{ctx->scratch_ptr, (size_t)SF_NOALLOC_SCRATCH_BYTES, 0};
  pl.arena = &arena;
  // no culling: its sort workspace would be another buffer
  rc = plan_build(ctx, pl, n, sl, px.data(), pu.data(), xmax.data(), st, false);
  pl.arena = nullptr;
  if (rc) return rc;
  for (int k = 0; k < 8; ++k) ctx->stats[k] = 0;
  ctx->stats[0] = pl.total_pairs;
  ctx->stats[2] = pl.n_items;
  ctx->stats[3] = pl.grid;
  ctx->stats[7] = pl.block;
  ctx->ev_pending = false;
  if (pl.total_pairs == 0) return SF_OK;
  unsigned long long *dc = (unsigned long long *)d_counts;
  rc = plan_launch(pl, pl.mixed, d_sums, dc, st, nullptr);
  if (rc) return rc;
  if (pl.mixed) {
    unsigned long long pc = 0;
    HIPCHK(hipMemcpyAsync(&pc, pl.pcount.p, sizeof(pc), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    ctx->stats[1] = (int64_t)pc;
    // the queue overflowed its share of the scratch: the merge was skipped, redo in binary64
    if ((int64_t)pc > pl.cap) rc = plan_launch(pl, false, d_sums, dc, st, nullptr);
  }
  return rc;
}

// ----------------------------------------------------------------------------------------------
// plans: build once, execute many times (optionally as a captured hipGraph)
// ----------------------------------------------------------------------------------------------
struct sf_plan {
  sf_ctx *ctx = nullptr;
  Plan pl;
  hipStream_t last_stream = nullptr;
  double *last_s = nullptr;
  unsigned long long *last_c = nullptr;
  // fused plan (sf_plan_create_multi): its outputs' problems and launch positions, and the last
  // execution's destinations (for the binary64 redo after a re-bin queue overflow)
  int n_out = 0;
  FusedSetup fs;
  std::vector<double *> last_ms;
  std::vector<unsigned long long *> last_mc;
};

extern "C" int sf_plan_create(sf_ctx *ctx, int32_t n, const sf_problem *sl, sf_plan **out) {
  if (!ctx || !out) return SF_E_ARG;
  *out = nullptr;
  ctx->err.clear();
  int rc = validate_batch(ctx, n, sl, false);
  if (rc) return rc;
  if (n == 0) return ctx->fail(SF_E_ARG, "a plan needs at least one slice");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  std::vector<const double *> px(n), pu(n);
  std::vector<double> xmax(n, 0.0);
  for (int s = 0; s < n; ++s) {
    px[s] = sl[s].x;
    pu[s] = sl[s].u;
    if (sl[s].precision != SF_PREC_FP64 && sl[s].geometry == SF_GEOM_FLAT && sl[s].N > 0) {
      // coordinate magnitude for the binary32 margin; later contents must stay within 2x of it
      std::vector<double> hx((size_t)sl[s].W * sl[s].N);
      HIPCHK(hipMemcpyAsync(hx.data(), sl[s].x, sizeof(double) * hx.size(), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      double m = 0;
      for (double v : hx) m = std::max(m, fabs(v));
      xmax[s] = 2.0 * m;
    }
  }
  sf_plan *plan = new sf_plan();
  plan->ctx = ctx;
  // no culling in plans: the positions behind x may change between executions
  rc = plan_build(ctx, plan->pl, n, sl, px.data(), pu.data(), xmax.data(), st, false);
  if (rc) {
    plan->pl.release();
    delete plan;
    return rc;
  }
  *out = plan;
  return SF_OK;
}

extern "C" int sf_plan_execute(sf_plan *plan, double *d_sums, uint64_t *d_counts, void *stream,
                               int32_t use_graph) {
  if (!plan) return SF_E_ARG;
  sf_ctx *ctx = plan->ctx;
  if (plan->n_out) return ctx->fail(SF_E_ARG, "a fused plan: use sf_plan_execute_multi");
  if (!d_sums || !d_counts) return ctx->fail(SF_E_ARG, "null sums or counts");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  Plan &pl = plan->pl;
  unsigned long long *dc = (unsigned long long *)d_counts;
  plan->last_stream = st;
  plan->last_s = d_sums;
  plan->last_c = dc;
  if (!use_graph) return plan_launch(pl, pl.mixed, d_sums, dc, st, nullptr);
  if (!pl.graph || pl.g_sums != d_sums || pl.g_counts != dc || pl.g_stream != st) {
    if (pl.graph) {
      HIPCHK(hipGraphExecDestroy(pl.graph));
      pl.graph = nullptr;
    }
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int rc = plan_launch(pl, pl.mixed, d_sums, dc, st, nullptr);
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return ctx->fail(SF_E_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
    const hipError_t e2 = hipGraphInstantiate(&pl.graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e2 != hipSuccess) {
      pl.graph = nullptr;
      return ctx->fail(SF_E_DEVICE, "graph instantiation failed: %s", hipGetErrorString(e2));
    }
    pl.g_sums = d_sums;
    pl.g_counts = dc;
    pl.g_stream = st;
  }
  HIPCHK(hipGraphLaunch(pl.graph, st));
  return SF_OK;
}

extern "C" int sf_plan_create_multi(sf_ctx *ctx, int32_t n, const sf_problem *sl, int32_t n_out,
                                    const sf_output *outs, sf_plan **out) {
  if (!ctx || !out) return SF_E_ARG;
  *out = nullptr;
  ctx->err.clear();
  if (n <= 0) return ctx->fail(SF_E_ARG, "a plan needs at least one slice");
  sf_plan *plan = new sf_plan();
  plan->ctx = ctx;
  plan->n_out = n_out;
  int rc = fused_setup(ctx, n, sl, n_out, outs, plan->fs);
  if (!rc && !plan->fs.separate.empty())
    rc = ctx->fail(SF_E_UNSUPPORTED,
                   "a fused plan needs a binary32 precision and outputs that fuse (output %d does not: "
                   "binary64, or FullVector{1}/{3}); make one plan per output instead",
                   plan->fs.separate[0]);
  if (rc) {
    delete plan;
    return rc;
  }
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  std::vector<const double *> px(n), pu(n);
  std::vector<double> xmax;
  for (int s = 0; s < n; ++s) {
    px[s] = sl[s].x;
    pu[s] = sl[s].u;
  }
  // later contents must stay within 2x of the coordinate magnitude read now
  rc = slice_xmax(ctx, n, sl, st, 2.0, xmax);
  // no culling in plans: the positions behind x may change between executions
  if (!rc) rc = plan_build(ctx, plan->pl, n, plan->fs.main().data(), px.data(), pu.data(), xmax.data(), st, false,
                           &plan->fs.ms);
  if (rc) {
    plan->pl.release();
    delete plan;
    return rc;
  }
  *out = plan;
  return SF_OK;
}

extern "C" int sf_plan_execute_multi(sf_plan *plan, double *const *d_sums, uint64_t *const *d_counts, void *stream,
                                     int32_t use_graph) {
  if (!plan) return SF_E_ARG;
  sf_ctx *ctx = plan->ctx;
  if (plan->n_out == 0) return ctx->fail(SF_E_ARG, "not a fused plan (use sf_plan_execute)");
  if (!d_sums || !d_counts) return ctx->fail(SF_E_ARG, "null output arrays");
  for (int k = 0; k < plan->n_out; ++k)
    if (!d_sums[k] || !d_counts[k]) return ctx->fail(SF_E_ARG, "output %d: null sums or counts", k);
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  Plan &pl = plan->pl;
  const FusedSetup &fs = plan->fs;
  bool moved = false;
  for (size_t j = 0; j < fs.ids.size(); ++j) {
    moved |= pl.mdst_s[j] != d_sums[fs.ids[j]] || pl.mdst_c[j] != (unsigned long long *)d_counts[fs.ids[j]];
    pl.mdst_s[j] = d_sums[fs.ids[j]];
    pl.mdst_c[j] = (unsigned long long *)d_counts[fs.ids[j]];
  }
  double *js = fs.joint >= 0 ? d_sums[fs.joint] : nullptr;
  unsigned long long *jc = fs.joint >= 0 ? (unsigned long long *)d_counts[fs.joint] : nullptr;
  plan->last_stream = st;
  plan->last_ms.assign(d_sums, d_sums + plan->n_out);
  plan->last_mc.assign((unsigned long long *const *)d_counts, (unsigned long long *const *)d_counts + plan->n_out);
  if (!use_graph) return plan_launch(pl, true, js, jc, st, nullptr);
  if (!pl.graph || moved || pl.g_sums != js || pl.g_counts != jc || pl.g_stream != st) {
    if (pl.graph) {
      HIPCHK(hipGraphExecDestroy(pl.graph));
      pl.graph = nullptr;
    }
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int rc = plan_launch(pl, true, js, jc, st, nullptr);
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return ctx->fail(SF_E_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
    const hipError_t e2 = hipGraphInstantiate(&pl.graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e2 != hipSuccess) {
      pl.graph = nullptr;
      return ctx->fail(SF_E_DEVICE, "graph instantiation failed: %s", hipGetErrorString(e2));
    }
    pl.g_sums = js;
    pl.g_counts = jc;
    pl.g_stream = st;
  }
  HIPCHK(hipGraphLaunch(pl.graph, st));
  return SF_OK;
}

extern "C" int sf_plan_sync(sf_plan *plan, int32_t *redone) {
  if (!plan) return SF_E_ARG;
  sf_ctx *ctx = plan->ctx;
  if (redone) *redone = 0;
  if (!plan->last_stream) return SF_OK;
  HIPCHK(hipSetDevice(ctx->device));
  Plan &pl = plan->pl;
  HIPCHK(hipStreamSynchronize(plan->last_stream));
  if (!pl.mixed || pl.total_pairs == 0) return SF_OK;
  int flag = 0;
  HIPCHK(hipMemcpy(&flag, pl.flag.p, sizeof(int), hipMemcpyDeviceToHost));
  if (flag && plan->n_out) {
    // a fused plan's last execute overflowed the re-bin queue and left every output untouched:
    // each output with the binary64 kernel
    for (int k = 0; k < plan->n_out; ++k) {
      std::vector<sf_problem> q = plan->fs.per[k];
      for (sf_problem &e : q) e.precision = SF_PREC_FP64;
      const int rc = sf_hist_batch_device(ctx, (int)q.size(), q.data(), plan->last_ms[k],
                                          (uint64_t *)plan->last_mc[k], plan->last_stream);
      if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(plan->last_stream));
    if (redone) *redone = 1;
  } else if (flag) {
    // the last execute overflowed the re-bin queue and left the outputs untouched
    const int rc = plan_launch(pl, false, plan->last_s, plan->last_c, plan->last_stream, nullptr);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(plan->last_stream));
    if (redone) *redone = 1;
  }
  return SF_OK;
}

extern "C" void sf_plan_destroy(sf_plan *plan) {
  if (!plan) return;
  (void)hipSetDevice(plan->ctx->device);
  if (plan->last_stream) (void)hipStreamSynchronize(plan->last_stream);
  plan->pl.release();
  delete plan;
}

// ----------------------------------------------------------------------------------------------
// autotuner
// ----------------------------------------------------------------------------------------------
extern "C" int sf_autotune(sf_ctx *ctx, int32_t n, const sf_problem *sl, int32_t out[6]) {
  if (!ctx) return SF_E_ARG;
  ctx->err.clear();
  int rc = validate_batch(ctx, n, sl, false);
  if (rc) return rc;
  if (n == 0) return ctx->fail(SF_E_ARG, "nothing to tune");
  HIPCHK(hipSetDevice(ctx->device));
  const sf_problem &p0 = sl[0];
  const bool J = p0.n_val > 0, mixed = p0.precision != SF_PREC_FP64;
  const bool fx = p0.precision == SF_PREC_MIXED_FIXED && J;
  const int ncell = out_cells(p0);
  DevBuf<double> ts;
  DevBuf<unsigned long long> tc;
  HIPCHK(ts.ensure((size_t)n * ncell));
  HIPCHK(tc.ensure((size_t)n * ncell));
  const Tuning saved = ctx->tune;
  std::vector<const double *> px(n), pu(n);
  std::vector<double> xmax(n, 0.0);
  for (int s = 0; s < n; ++s) {
    px[s] = sl[s].x;
    pu[s] = sl[s].u;
    xmax[s] = 1.0;  // timing only: margins do not change the work, results are discarded
  }
  // device time of one call with launch shape c (a warm-up call first)
  auto time_of = [&](const Tuning &c, double *t) {
    ctx->tune = c;
    for (int rep = 0; rep < 2; ++rep) {
      const int r = run_batch(ctx, n, sl, px.data(), pu.data(), xmax.data(), ts.p, tc.p, ctx->stream);
      if (r) return r;
      int64_t st[8];
      sf_last_stats(ctx, st);
      *t = (double)st[6];
    }
    return SF_OK;
  };
  // coordinate search: threads per workgroup x column variant, then the histogram's privatised
  // copies and (joint) slot layout, then the column-range length
  std::vector<Tuning> stage1;
  const int blocks[4] = {256, 512, 768, 1024};
  for (int b : blocks) {
    if (J && b < 512) continue;
    if (!J && b > 512) continue;
    if (J && b == 768 && !fx) continue;  // 768: the fixed-point cells' two-workgroup launch
    for (int v = (mixed ? 1 : 0); v <= (mixed ? 2 : 0); ++v) stage1.push_back(Tuning{b, 0, 0, v, 0, 0});
  }
  double best = 1e300;
  Tuning best_c = stage1.front();
  auto consider = [&](const Tuning &c) {
    double t = 0;
    const int r = time_of(c, &t);
    if (r == SF_E_UNSUPPORTED) return SF_OK;  // shape does not fit this device: skip it
    if (r) return r;
    if (t < best) {
      best = t;
      best_c = c;
    }
    return SF_OK;
  };
  for (const Tuning &c : stage1)
    if ((rc = consider(c))) break;
  if (!rc) {
    std::vector<Tuning> stage2;
    if (!J) {
      for (int cp : {4, 8, 32}) {
        Tuning c = best_c;
        c.copies = cp;
        stage2.push_back(c);
      }
    } else if (mixed) {
      for (int cp : {1, 2})
        for (int lay : {1, 2}) {
          if (cp == 1 && lay == 1) continue;  // the built-in choice, already timed
          Tuning c = best_c;
          c.copies = cp;
          c.layout = lay;
          stage2.push_back(c);
        }
    }
    for (const Tuning &c : stage2)
      if ((rc = consider(c))) break;
  }
  if (!rc) {
    const int b = best_c.block ? best_c.block : 256;
    for (int sc : {2, 8}) {
      Tuning c = best_c;
      c.j_chunk = sc * b;
      if ((rc = consider(c))) break;
    }
  }
  ctx->tune = saved;
  ts.release();
  tc.release();
  if (rc) return rc;
  int64_t total = 0;
  for (int s = 0; s < n; ++s) total += pairs_of(sl[s].N);
  ctx->tuned[tune_key(p0, total)] = best_c;
  if (out) {
    out[0] = best_c.block;
    out[1] = best_c.j_chunk;
    out[2] = best_c.bpc;
    out[3] = best_c.variant;
    out[4] = best_c.copies;
    out[5] = best_c.layout;
  }
  return SF_OK;
}
