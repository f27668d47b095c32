// sf_internal.h -- layouts shared by the host launcher (sf_api.cpp) and the kernels (sf_kernels.hip).
#pragma once
#include <stdint.h>

#if !defined(__HIPCC__)
#define __host__
#define __device__
#endif

namespace sfk {

enum Geom : int { kFlat = 0, kSphere = 1 };

// Per-point records (array of structures, 16-byte aligned) built on the device by the prep kernels.
//   binary64 record (exact path):
//     flat   D=2: x0 x1 u0 u1                      (4 doubles)
//     flat   D=3: x0 x1 x2 u0 u1 u2                (6 doubles)
//     sphere D  : P0 P1 P2 E0 E1 N0 N1 N2 u0 u1 [u2] pad -> 12 doubles; P = unit position,
//                 E = local east (E2 = 0), N = local north (HelperFunctions.jl:133-136)
//   binary32 record (mixed path), every binary64 input carried as a hi + lo pair of floats
//   (v = hi + lo to ~2^-48 relative) so that differences of nearby values stay accurate:
//     flat   D=2: xh0 xh1 xl0 xl1 uh0 uh1 ul0 ul1 |u| id 0 0                  (12 floats)
//     flat   D=3: xh0..2 xl0..2 uh0..2 ul0..2 |u| id 0 0                      (16 floats)
//     sphere D  : ph0..2 pl0..2 Uh0..2 Ul0..2 |u| urh url id                 (16 floats)
//                 U = u_e E + u_n N is the tangent velocity in R^3, ur the radial component
//                 (D = 3), P the unit position; |u| is an upper bound of the velocity norm;
//                 id is the point's index in the caller's order (int32 bits): binary32 records
//                 may be stored in a spatially sorted order (tile culling).
inline constexpr __host__ __device__ int rec_d64(int G, int D) { return (G == kFlat) ? ((D == 2) ? 4 : 6) : 12; }
inline constexpr __host__ __device__ int rec_f32(int G, int D) { return (G == kFlat && D == 2) ? 12 : 16; }
inline constexpr __host__ __device__ int rec_id(int G, int D) { return (G == kFlat) ? 4 * D + 1 : 15; }
template <int G, int D> struct Rec {
  static constexpr int d64 = rec_d64(G, D);
  static constexpr int f32 = rec_f32(G, D);
  static constexpr int id = rec_id(G, D);
};

// Distance-key cell table entry.  The positive key axis (flat: r^2; sphere: chord^2 for the
// binary32 path, central-angle distance for the binary64 path) is cut into cells at powers of two
// times (1 + m / 2^M); a key's cell is (bits >> (mantissa_bits - M)) - base, clamped.
struct Cell {
  float lo, hi;   // binary32 ambiguity interval inside this cell (lo > hi: none; -inf..inf: all)
  int32_t blo;    // binary32 path: bin count for keys below lo (keys above hi: blo + 1)
  int32_t bx;     // binary64 path: number of exact edges below the cell's lower bound
};

// Binary64 O(1) bin rule of a LinearBinEdges axis (optionally InfPadded); see lin_rule_bin.
struct LinRule {
  double f, l, inv, step;  // first_edge, last_edge, inv_step, step_val
  int64_t n;               // length of the range
  int32_t on;              // 0: search the edge vector instead
  int32_t padded;          // InfPaddedBinEdges around the range
};

struct Slice {
  LinRule drule;      // distance axis rule (flat point-field kernels: applied to sqrt(r^2))
  LinRule vrule;      // value axis rule
  int64_t N;          // points
  int64_t pt_off;     // first record of this slice in the record arrays
  int64_t out_off;    // first output cell
  int64_t base64;     // binary64 table: cell = (bits >> (52 - M64)) - base64
  int32_t M64, ncell64;
  int32_t M32, ncell32;
  int32_t base32;     // binary32 table: cell = (bits >> (23 - M32)) - base32
  int32_t tab_off;    // first Cell of this slice's table (binary64 keys)
  int32_t tab32_off;  // first Cell of this slice's binary32 (chord^2 / r^2) table
  int32_t dedge_off;  // exact distance-key edges (flat: e*e; sphere: e), n_dist + 1 doubles
  int32_t vedge_off;  // value edges, n_val + 1 doubles (and float copies at the same offset)
  int32_t v_uniform;  // value edges uniform enough for the guess-then-walk lookup
  int32_t v_negzero;  // count of -0.0 value edges (Julia isless places +0.0 above them)
  double v0, vinv;    // value guess t = (v - v0) * vinv
  float v0f, vinvf;   // binary32 value guess
  float vtc;          // -v0f * vinvf rounded to binary32: t = fmaf(v, vinvf, vtc)
  float vslack;       // binary32 value-axis slack in units of t (edge non-uniformity + rounding)
  float vabs;         // binary32 value-axis absolute slack of searched edges (LinearBinEdges rule)
  int32_t v_fx;       // value edges uniform to 1/64 of a bin: fixed-point joint accumulation allowed
  // fixed-point joint accumulation: a value v of value bin g (0-based) is stored as the integer
  // part q of Q = fmaf(v, vsf, c_g) - 2^23, c_g = fmaf(-g, fxc1, fxc0) (binary32; fxc1 = vwf vsf,
  // fxc0 = 2^23 + kFxBias - v0f vsf, so that Q ~ (v - (v0 + g w)) vsf + kFxBias); the flush
  // recovers sum v = (sum q + n (2^23 - c_g)) / vsf with the same c_g, exact but for the rounding
  // of each Q to an integer (half a unit, 2^-23 bin widths)
  float vwf, vsf;
  float fxc1, fxc0;
  // ... except values with |v| < vsmall (a quarter bin width), whose fixed-point error would not be
  // small relative to the value: they go to binary64 cells of the value bins gz .. gz + kZeroBins - 1
  float vsmall;
  int32_t gz;
  int32_t plane;      // single pass, 2D form: invariant (0..5) whose joint histogram this slice holds
  float xerr;         // flat: 2^-28 max|x| (binary32 value bound: reference's direction error)
  int32_t strip_n;    // B-strip launch: fields of the strip this slice leads (0: not a leader)
  // value edges log-spaced from edge vk0 on (binary32 lookup guesses the bin from log2 v):
  // log2(edge vk0 + m) ~ vl0f + m / vlinvf
  int32_t v_log, vk0;
  float vl0f, vlinvf;
};

struct Item {
  int32_t slice;
  int32_t i0;       // first row of the i-tile
  int32_t j0, j1;   // column range [j0, j1), j0 >= i0
};

struct Pending {      // a pair the binary32 kernel could not bin with certainty (binary32 record indices)
  int32_t slice, i, j;
  int32_t fields;     // B-strip: slices slice .. slice + fields - 1 share the pair (0: one slice)
};

struct Params {
  const Slice *slices;
  const Item *items;
  int32_t n_items;
  int32_t block_rows;      // i-tile rows (= threads per block)
  int32_t *counter;        // work-queue head (zeroed before launch)
  const double *rec64;     // binary64 records
  const float *rec32;      // binary32 records
  const Cell *cells;       // cell tables (binary64 tables and binary32 tables)
  const double *dedges;    // exact distance-key edges
  const double *vedges;    // value edges (binary64)
  const float *vedges32;   // value edges rounded to binary32 (non-uniform value lookup)
  double *sums;            // output sums (device)
  unsigned long long *counts;  // output counts (device)
  Pending *pending;        // re-bin queue (mixed kernel)
  unsigned long long *pending_count;  // [0] = appended, may exceed capacity
  int64_t pending_cap;
  unsigned long long *stats;  // [0] pairs visited by exact kernel recheck
  int32_t ncell_max;       // largest table of any slice (LDS sizing)
  int32_t tj;              // column records staged in LDS per chunk
  int32_t copies;          // 1D: per-lane histogram copies (power of two); joint: 1 or 2 (lane halves)
  int32_t jlayout;         // joint: LDS slot layout (joint_layout below); 1D: 0
  int32_t stage_bytes;     // LDS bytes of the column stage
  int32_t n_dist, n_val;
  int32_t op, nl, nt;
  int32_t order;           // operator order (value-error bound exponent)
  double radius;
  float vcoef;             // binary32 value error coefficient: |dv| <= vcoef * Umax^order * amp
  int32_t nmom;            // accumulated moments per bin: 1, or 4 for the single-pass invariants
  int32_t fx;              // joint histogram kept as 64-bit fixed-point words (kFx* below)
  int32_t qcap;            // binary32 kernel: re-bin queue entries staged in LDS per workgroup
  int32_t qw;              // ... of which each wave owns qcap / (block_rows / 64)
  int64_t flush_limit;     // pair updates a workgroup may add to its LDS histogram between flushes
  int32_t bl;              // flat pairs follow the batch-leading kernels (SF_SEM_BATCH)
  int32_t strip;           // B-strip launch: LDS histograms per workgroup (fields per strip), 0: off
  // fused outputs (one pair evaluation, several operators): n1d 1D histograms besides the joint
  // one (if any); 1D output k of slice s accumulates at msums / mcounts + (k n_slices + s) n_dist
  int32_t n1d;
  int32_t mcopies;         // lane-group copies of each fused 1D histogram (power of two)
  int32_t n_slices;
  int32_t mop[8], mnl[8], mnt[8];
  // binary32 value of fused 1D output k: X[a_k] * X[b_k] over the per-pair basis
  // X = {1, du_L, du_L^2, du_L^3, du_T, tn2, du_T^3, |du|^2, tn2 / (D - 1)} (fused_basis_index);
  // a_k and b_k are the 4-bit fields k of mab_a and mab_b
  uint32_t mab_a, mab_b;
  int32_t ja, jb;          // the joint operator's basis indices (fused plan with a joint output)
  double *msums;
  unsigned long long *mcounts;
};

// fused outputs: at most this many 1D operators per pair evaluation
inline constexpr int kMaxFused = 8;

// Basis indices (a, b) of a fused 1D operator, whose binary32 value op_value32 forms as the
// product X[a] * X[b] with the same operations (x * 1 and 1 * x are exact): du_L^nl * (1, du_T,
// tn2, du_T^3 for nt = 0..3); |du|^2 (S2, FullVector{2}); du_L |du|^2 (S3); tn2 / (D - 1) and
// du_L tn2 / (D - 1) (the component forms).  FullVector{1} and {3} take a square root: not
// fusable (returns false).
inline bool fused_basis_index(int op, int nl, int nt, int *a, int *b) {
  switch (op) {
    case 0: *a = nl; *b = nt == 0 ? 0 : 3 + nt; return nl >= 0 && nl <= 3 && nt >= 0 && nt <= 3;
    case 1: *a = 0; *b = 7; return true;
    case 2: *a = 1; *b = 7; return true;
    case 3: *a = 0; *b = 8; return true;
    case 4: *a = 1; *b = 8; return true;
    case 5: *a = 0; *b = 7; return nl == 2;
    default: return false;
  }
}

// Fixed-point joint cell: one 64-bit word, the sum over its updates of kFxUnit + q, where
// kFxUnit = 2^44 + kFxBase and q < 2^23 (value within [-1/64, 1 + 1/64] bin widths of a_g, scaled
// by 2^22, biased by 2^20).  With n updates the word is n kFxUnit + sum q and sum q < 2^43 <
// kFxUnit, so n = word / kFxUnit and sum q = word - n kFxUnit; the word stays below 2^64 while
// n <= kFxFlushLimit = 2^20 - 256 (2^20 kFxUnit alone would exceed it).
inline constexpr int kFxShift = 44;
inline constexpr float kFxScale = 4194304.0f;  // 2^22 per value-bin width
inline constexpr float kFxBias = 1048576.0f;   // 2^20
// the pair loop adds q as the bits of the binary32 number 2^23 + kFxBias + (value offset): that is
// kFxBase + q; one update adds kFxUnit + q to the word (count 1 in bits 44.., kFxBase below)
inline constexpr unsigned long long kFxBase = 0x4B000000ull;  // bits of 8388608.0f
inline constexpr unsigned long long kFxUnit = (1ull << 44) + kFxBase;
inline constexpr int64_t kFxFlushLimit = (1LL << 20) - 256;

// Single-pass output: (6, n_dist) per slice, rows S2 L2 T2 S3 L3 L1T2; the kernels accumulate the
// moments S2, L2, S3, L3 (rows 0, 1, 3, 4) and the count (row 0).
inline constexpr int kSpRows = 6;
inline constexpr __host__ __device__ int sp_row(int k) { return k < 2 ? k : k + 1; }

// LDS layout of the histogram kernels, shared by the host (sizing) and the device (carving).
struct SmemLayout {
  int32_t hs, hc, zs, zc, ms, mc, tab, dE, vE, vE32, stage, q, total;  // byte offsets
};

// fixed-point joint histograms: binary64 cells (sum, count) for values smaller than a quarter
// bin width, kZeroBins value bins around zero per distance bin
inline constexpr int kZeroBins = 5;

// LDS histogram slots (nmom binary64 sums, moment-major, and one 32-bit count each).  1D
// histograms keep one copy per lane group.  A joint histogram's slot of (distance bin d, value bin
// v), 0-based, is v * jv + d * jd within one copy of `slots` slots:
//   layout 1 (default): distance-bin rows of n_val slots at an odd stride (lanes of one wave
//                       usually share a distance bin but not a value bin, so their atomics then
//                       fall in different banks): jv = 1, jd = n_val | 1;
//   layout 2:           value-bin rows of n_dist slots at an odd stride: jd = 1, jv = n_dist | 1.
// With two copies (privatisation by lane halves) lanes 32..63 use the second copy.
struct JointLayout {
  int32_t jv, jd, slots;
};
inline __host__ __device__ JointLayout joint_layout(int n_dist, int n_val, int layout) {
  JointLayout L;
  if (layout == 2) {
    L.jv = n_dist | 1;
    L.jd = 1;
    L.slots = n_val * L.jv;
  } else {
    L.jv = 1;
    L.jd = n_val | 1;
    L.slots = n_dist * L.jd;
  }
  return L;
}
// single pass, 2D form (op SF_OP_SINGLE_PASS with value bins): every input slice runs as six
// plane slices, one per invariant, each a joint histogram written to every sixth output cell
inline __host__ __device__ bool sp2d(int op, int n_val) { return op == 6 && n_val > 0; }

inline constexpr int kDummySlots = 64;
inline constexpr int kZeroSetMax = 8;  // buffers cleared by one zero_multi launch
// largest binary32 key table (cells); the host's partition never exceeds it.  In LDS the table's
// (lo, hi) pairs start the region and its bin counts follow at the same 8-byte stride; for joint
// histograms they sit kCell32Max * 8 bytes on, so that one address serves both reads (the second
// takes a constant offset), for 1D ones right after the largest table of the launch (their LDS
// decides how many workgroups fit: 16 KB of table region cost them 12 %)
inline constexpr int kCell32Max = 1024;
// largest re-bin queue staged in LDS per workgroup (entries, split evenly between its waves);
// the host shrinks it where the LDS it takes would cost a workgroup per CU
inline constexpr int kQueueCap = 8192;
// LDS queue header: one fill counter per wave (at most 16 waves of 64 lanes)
inline constexpr int kQueueHead = 64;
// largest global re-bin queue (entries of 16 bytes: 4 GiB); a call needing more is redone with
// the binary64 kernel
inline constexpr int64_t kQueueMax = 1LL << 28;
// B-strip launches: velocity fields evaluated per pair geometry (rows keep each field's
// velocity record in registers)
inline constexpr int kStripMax = 4;

// Fused 1D outputs: the kernel evaluates and accumulates a fixed number NF >= n1d of them (the
// outputs past n1d are 1 * 1 into histograms that are never flushed), so its loop has no branch
// per output.  NF is 1, 2, 4, 6 or 8; kernel class (OPK) 9 .. 13 respectively.
inline __host__ __device__ int fused_pad(int n1d) {
  return n1d <= 2 ? n1d : (n1d <= 4 ? 4 : (n1d <= 6 ? 6 : 8));
}
inline __host__ __device__ int fused_opk(int n1d) {
  const int nf = fused_pad(n1d);
  return nf == 1 ? 9 : (nf == 2 ? 10 : (nf == 4 ? 11 : (nf == 6 ? 12 : 13)));
}

// copies: 1D lane-group copies, or the joint histogram's copies (1 or 2); jlayout: joint layout
inline __host__ __device__ SmemLayout smem_layout(int n_dist, int n_val, int ncell_max,
                                                  int stage_bytes, int copies, int nmom, int fx,
                                                  int qcap, int strip = 0, int jlayout = 1, int n1d = 0,
                                                  int mcopies = 1) {
  SmemLayout L;
  // + kDummySlots: one scratch slot per lane, target of unconditional atomics of lanes that have
  // nothing to add (never flushed); a B-strip launch keeps one histogram per field
  const int one = (n_val > 0) ? joint_layout(n_dist, n_val, jlayout).slots : n_dist;
  const int nh = one * copies * (strip > 1 ? strip : 1) + kDummySlots;
  int o = 0;
  L.hs = o; o += 8 * nh * nmom;
  L.hc = o; o += fx ? 0 : 4 * ((nh + 3) & ~3);  // fixed-point cells carry their count
  const int nz = fx ? kZeroBins * n_dist + kDummySlots : 0;
  L.zs = o; o += 8 * nz;
  L.zc = o; o += 4 * ((nz + 3) & ~3);
  // fused 1D outputs: fused_pad(n1d) histograms of n_dist x mcopies slots sharing one count per slot
  const int nm = n1d > 0 ? n_dist * mcopies + kDummySlots : 0;
  L.ms = o; o += 8 * nm * fused_pad(n1d);
  L.mc = o; o += 4 * ((nm + 3) & ~3);
  L.tab = o; o += 16 * ((n_val > 0 && ncell_max < kCell32Max) ? kCell32Max : ncell_max);
  L.dE = o; o += 8 * ((n_dist + 2) & ~1);
  L.vE = o; o += 8 * ((n_val + 2) & ~1);
  L.vE32 = o; o += 4 * ((n_val + 4) & ~3);
  L.stage = o; o += stage_bytes;
  // re-bin queue: the per-wave fill counters, then qcap (i, j) entries of 8 bytes
  L.q = o; o += qcap > 0 ? kQueueHead + 8 * qcap : 0;
  L.total = o;
  return L;
}

}  // namespace sfk
