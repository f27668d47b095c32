/* sfmi355x.h -- C ABI of libsfmi355x.so, the gfx950 pair-histogram backend.
 *
 * One call computes, for every unordered point pair (i < j) of a point field, the separation
 * distance bin and (optionally) the value bin of a structure-function operator, and accumulates
 * per-bin sums of the operator value and exact pair counts.  This is the computation behind the
 * reference's
 *     calculate_structure_function(sf, x, u, distance_bins[, value_bins]; distance_metric)
 * (src/Calculations/dispatch.jl:64-136) and its mutating form calculate_structure_function!
 * (dispatch.jl:346-373), as reached through the backend hook
 *     _dispatch_execution_backend(backend, shape::PointField{D}, sf, x, u, bins[, value_bins])
 * (dispatch.jl:510-594; mutating variants dispatch.jl:377-477).  A Julia backend type plugs in
 * there and ccalls these entry points (structurefunctions.jl_amd/julia/).
 *
 * Conventions (mirroring the reference):
 *   - arrays are Julia column-major: x is (W, N), u is (D, N); element [d, i] at d + W*i;
 *   - histograms are (n_dist) or (n_dist, n_val) column-major; sums are binary64, counts uint64;
 *   - outputs are ACCUMULATED (+=), like calculate_structure_function! (the allocating API zeroes
 *     first, src/Calculations/batch.jl:161-199);
 *   - every validation the reference performs before its O(N^2) loop (shapes.jl:67-105,
 *     serial.jl:111-120, HelperFunctions.jl:411-417, dispatch.jl:7-13) is done before any device
 *     work; failures return a negative code and sf_last_error() describes them; nothing aborts
 *     or throws across the ABI;
 *   - caller pointers are not retained after return;
 *   - a context serves one call at a time (like GPUSFWorkspace, src/Calculations/workspace.jl:44).
 */
#ifndef SFMI355X_H
#define SFMI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* return codes */
#define SF_OK 0
#define SF_E_ARG (-1)        /* ArgumentError in the reference                 */
#define SF_E_DIM (-2)        /* DimensionMismatch in the reference             */
#define SF_E_COUNT (-3)      /* count width cannot hold N(N-1)/2 (serial.jl:111-120) */
#define SF_E_DEVICE (-4)     /* HIP runtime failure                            */
#define SF_E_UNSUPPORTED (-5)

/* geometry: pair_geometry_for(distance_metric) (src/HelperFunctions.jl:390-393) */
#define SF_GEOM_FLAT 0        /* Distances.Euclidean()        -> FlatGeometry{D}, W = D          */
#define SF_GEOM_SPHERE_DEG 1  /* Distances.Haversine(radius)  -> SphericalGeometry, lon/lat deg  */
#define SF_GEOM_SPHERE_RAD 2  /* Distances.SphericalAngle()   -> SphericalGeometry, radians, R=1 */

/* operator kind (src/StructureFunctionTypes.jl) */
#define SF_OP_PROJECTED 0     /* ProjectedStructureFunctionType{nl, nt}, 0 <= nl + nt <= 3 (:43-89) */
#define SF_OP_S2 1            /* SecondOrderStructureFunctionType  ||du||^2       (:97,120)       */
#define SF_OP_S3 2            /* ThirdOrderStructureFunctionType   du_L ||du||^2  (:106,122)      */
#define SF_OP_T2COMP 3        /* TransverseComponentSecondOrder... ||du_t||^2/(D-1) (:137,147)    */
#define SF_OP_L1T2COMP 4      /* LongitudinalTransverseComponentThirdOrder...     (:145,150)      */
#define SF_OP_FULLVEC 5       /* FullVectorStructureFunctionType{nl} ||du||^nl, nl in 1..3 (:114) */
#define SF_OP_SINGLE_PASS 6   /* the six invariants S2, L2, T2, S3, L3, L1T2 in one pass; nl, nt
                                 ignored.
                                 n_val = 0 (calculate_structure_functions_single_pass[!],
                                 src/Calculations/serial_single_pass.jl:404-421, 602-700): output
                                 per slice (6, n_dist) column-major, rows in that order.  Rows 1,
                                 2, 4, 5 and count row 1 accumulate; then, over the whole output,
                                 row 3 = row 1 - row 2, row 6 = row 4 - row 5 and count rows 2..6 =
                                 count row 1 (_sp1d_derive_rows!, :193-219).
                                 n_val > 0, the 2D form (calculate_structure_functions_single_pass_2d[!],
                                 :712-728, 1161-1181): each invariant binned by distance and by its
                                 own value edges; val_edges holds six edge vectors back to back
                                 (6 * (n_val + 1) doubles, invariant t's at t * (n_val + 1)); output
                                 per slice (6, n_dist, n_val) column-major, every cell accumulated.
                                 Runs as six joint histograms, one per invariant.                 */

/* precision */
#define SF_PREC_FP64 0        /* binary64 pair arithmetic in the reference's order, bins certified
                                 from the exact keys and values (the reference's rules near every
                                 decision point): the drop-in (counts bit-exact)                 */
#define SF_PREC_MIXED 1       /* binary32 pair arithmetic, binary64 re-binning of pairs near any
                                 bin edge (counts bit-exact), binary64 accumulation              */
#define SF_PREC_MIXED_FIXED 2 /* as SF_PREC_MIXED, but a joint histogram whose value edges are
                                 uniform (to 1/64 of a bin) keeps each cell as one 64-bit
                                 fixed-point word (count + quantized sum) in LDS: counts still
                                 bit-exact; each pair's value enters its sum with an extra absolute
                                 error <= 2^-23 of its value bin's width.  Other shapes run as
                                 SF_PREC_MIXED.                                                  */

/* semantics: the reference evaluates a point field (PointField, x (W,N), u (D,N)) with its
 * point-field kernels and fields with auxiliary axes (u (D,N,B...), shared or per-slice x) with its
 * batch-leading kernels; on the flat geometry these differ in the last bits:
 *   SF_SEM_POINT: distance binned as r^2 against squared edges (squared digitize plan,
 *                 BinEdges.jl:410-558), values from _sf_raw (StructureFunctionTypes.jl:227-244),
 *                 single-pass du_L = dot(du, dx) * inv(sqrt(r^2)) (serial_single_pass.jl:245-253);
 *   SF_SEM_BATCH: distance r = sqrt(r^2) binned against the edges themselves, values from the
 *                 operator on rhat = dx / r, single-pass du_L = dot(du, rhat) (pair_frame,
 *                 pair_direction, pair_delta: HelperFunctions.jl:248-330; batch_leading.jl:105-330).
 * The sphere is evaluated the same way by both. */
#define SF_SEM_POINT 0
#define SF_SEM_BATCH 1

typedef struct sf_ctx sf_ctx;

/* A bin axis given as a LinearBinEdges (src/BinEdges.jl:94-131), optionally inside
 * InfPaddedBinEdges (:270-345).  The reference digitizes such an axis with an O(1) rule rather
 * than a search of the edge values:
 *   searchsortedfirst(v, x) = 1 if x <= first_edge; n + 1 if x > last_edge; otherwise
 *     idx = clamp(floor(muladd(x, inv_step, -first_edge * inv_step)) + 1, 1, n),
 *     idx + 1 if muladd(idx - 1, step_val, first_edge) < x, else idx  (BinEdges.jl:114-131),
 * and on the flat geometry's squared-distance plan (SquaredLinearPlan, :412-417, 506, 516) the
 * key is sqrt(r^2).  Its result can differ from a search of the edge values for queries within a
 * few ulps of an edge; passing the rule makes the counts those of the reference.  muladd is taken
 * as a fused multiply-add (what Julia emits on x86-64 with FMA and on AArch64).  The fields are
 * the LinearBinEdges fields; the edge vector of the problem must hold the same axis materialised
 * (collect(range), with -Inf / +Inf around it when padded) and is used for sizes and for the
 * binary32 kernels' certified binning.  A NaN query, which the reference rejects (floor(Int, NaN)
 * throws), is not counted. */
#define SF_RULE_LINEAR 1
typedef struct {
  int32_t kind;               /* SF_RULE_LINEAR                                                   */
  int32_t padded;             /* 1: InfPaddedBinEdges around the range (catch-all first/last bins) */
  int64_t n;                  /* length of the range (edges of the unpadded axis)                 */
  double first_edge, last_edge, inv_step, step_val;  /* LinearBinEdges fields (BinEdges.jl:94-100) */
} sf_bin_rule;

typedef struct {
  int32_t D;                  /* velocity components: 2 or 3                                     */
  int32_t W;                  /* coordinate rows of x: D (flat) or 2 (sphere)                     */
  int64_t N;                  /* points                                                           */
  const double *x;            /* (W, N)                                                           */
  const double *u;            /* (D, N)                                                           */
  int32_t geometry;           /* SF_GEOM_*                                                        */
  double radius;              /* Haversine radius (> 0); ignored unless SF_GEOM_SPHERE_DEG        */
  int32_t op, nl, nt;         /* SF_OP_* and its exponents                                        */
  int32_t n_dist;             /* distance bins; dist_edges has n_dist + 1 ascending entries,
                                 bins are (e[k-1], e[k]]                                          */
  const double *dist_edges;   /* host pointer                                                     */
  int32_t n_val;              /* 0 => 1D histogram; else value bins of the joint histogram        */
  const double *val_edges;    /* host pointer, n_val + 1 ascending entries (or NULL); six vectors
                                 back to back for the 2D single pass (SF_OP_SINGLE_PASS)          */
  int32_t precision;          /* SF_PREC_*                                                        */
  int32_t count_bits;         /* 32 or 64: width of the caller's count type (UInt32 default,
                                 dispatch.jl:69); calls whose N(N-1)/2 exceeds it are rejected    */
  int32_t semantics;          /* SF_SEM_*: which of the reference's pair loops the results follow */
  int32_t reserved;           /* 0                                                                */
  const sf_bin_rule *dist_rule;  /* NULL: distance bins are searched in dist_edges (BinEdges /
                                    LogBinEdges' physical edges); else the LinearBinEdges rule     */
  const sf_bin_rule *val_rule;   /* the same for the value axis (NULL for the 2D single pass)      */
} sf_problem;

/* Context: device, stream, persistent device workspace.
 * Replaces the per-call device setup of gpu_calculate_structure_function
 * (ext/StructureFunctionsKernelAbstractionsExt.jl:630-675) and GPUSFWorkspace
 * (ext/gpu/workspace.jl:25-56). */
int sf_create(int device, sf_ctx **out);
void sf_destroy(sf_ctx *ctx);
const char *sf_last_error(const sf_ctx *ctx);

/* Host-pointer entry.  x, u, sums, counts are host memory; synchronous.
 * Replaces gpu_calculate_structure_function[!] / gpu_calculate_structure_function_2d
 * (src/Calculations/gpu_stubs.jl:136-184; KA ext :385-401, :804-827, :1650-1667), i.e. the
 * PointField methods of _dispatch_execution_backend[!] (dispatch.jl:399-404, 537-544, 576-583). */
int sf_hist(sf_ctx *ctx, const sf_problem *p, double *sums, uint64_t *counts);

/* Many independent point fields ("slices", ragged N) in one persistent launch.  All slices must
 * share D, W, geometry, radius, operator, n_dist, n_val and precision; each slice has its own
 * edges.  Slice s accumulates into sums/counts[s * ncell .. (s+1) * ncell), ncell = n_dist * max(n_val,1)
 * (6 * n_dist * max(n_val, 1) for SF_OP_SINGLE_PASS).
 * Replaces gpu_calculate_structure_function_batch[!] (gpu_stubs.jl:191-213), generalised from
 * the reference's equal-N batches (batch_leading.jl:43-63) to ragged N. */
int sf_hist_batch(sf_ctx *ctx, int32_t n_slices, const sf_problem *slices, double *sums,
                  uint64_t *counts);

/* Device-pointer entries: x, u (inside each sf_problem), sums and counts are device memory on
 * the context's device; edges stay host pointers.  Work is enqueued on `stream` (a hipStream_t;
 * NULL = the context's own stream) and the call returns after enqueueing.  Device buffers must
 * stay alive until the stream reaches that point.  Counts/sums are accumulated atomically.
 * (The reference has no device-pointer entry; its GPU path stages host arrays each call,
 * KA ext :309-355.  This serves callers whose fields already live in HBM.) */
int sf_hist_device(sf_ctx *ctx, const sf_problem *p, double *d_sums, uint64_t *d_counts,
                   void *stream);
int sf_hist_batch_device(sf_ctx *ctx, int32_t n_slices, const sf_problem *slices, double *d_sums,
                         uint64_t *d_counts, void *stream);

/* Fused outputs: one evaluation of every pair feeds several operators of the same point fields.
 * The slices give the fields, geometry, distance bins, precision and semantics as for
 * sf_hist_batch_device (their op, nl, nt and n_val are ignored, their val_edges and val_rule too
 * unless a joint output asks for them); outs[k]
 * gives operator k and, for at most one of them, value bins (a joint histogram).  Output k of
 * slice s accumulates into d_sums[k] / d_counts[k] at s * ncell_k (ncell_k = n_dist, or
 * n_dist * n_val for the joint output).  Results equal n_out separate calls (counts bit-exact;
 * sums to each precision's bound); the binary32 precisions evaluate the pair geometry, the frame
 * and the increments once for all outputs, SF_PREC_FP64 runs the outputs one after the other.
 * At most SF_MAX_FUSED 1D outputs plus one joint output; SF_OP_SINGLE_PASS is not an output (it
 * is its own fused form).  No reference counterpart as one call: the reference evaluates several
 * moments per pair in its GPU histogram kernel (ext/gpu/sf_tiled.jl:7-13, 128-131) and shares a
 * pair's geometry across fields in its batch-leading loop (src/Calculations/batch_leading.jl:105-140);
 * each output is calculate_structure_function(outs[k] operator, x, u, distance_bins[, value_bins]). */
#define SF_MAX_FUSED 8
typedef struct {
  int32_t op, nl, nt;            /* operator (SF_OP_*, exponents)                                 */
  int32_t n_val;                 /* 0: 1D over the distance bins; else the joint histogram's value
                                    bins (n_val + 1 ascending edges in val_edges, host pointer)    */
  const double *val_edges;       /* NULL with n_val > 0: each slice's own val_edges and val_rule
                                    (per-slice value bins; each must hold n_val + 1 edges)        */
  const sf_bin_rule *val_rule;   /* NULL or the value axis' LinearBinEdges rule                   */
} sf_output;
int sf_hist_multi_device(sf_ctx *ctx, int32_t n_slices, const sf_problem *slices, int32_t n_out,
                         const sf_output *outs, double *const *d_sums, uint64_t *const *d_counts,
                         void *stream);
/* Host-pointer form of sf_hist_multi_device (x, u, sums and counts in host memory; synchronous). */
int sf_hist_multi(sf_ctx *ctx, int32_t n_slices, const sf_problem *slices, int32_t n_out,
                  const sf_output *outs, double *const *sums, uint64_t *const *counts);

/* Allocation-free variant of sf_hist_batch_device.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and results; the call makes no hipMalloc or
 * hipFree: every device buffer of the call (records, bin tables, work items, accumulators, the
 * re-bin queue) is carved from one static device array of SF_NOALLOC_SCRATCH_BYTES per device,
 * held until the process ends.  Limits: the records of all slices must fit in half of it (about
 * 400 000 points for the sphere in a binary32 precision, 700 000 in binary64) and the rest of the
 * call in the remainder, else SF_E_UNSUPPORTED is returned; the re-bin queue gets whatever is left,
 * and a call that overflows it is redone with the binary64 kernel (slower, same results); no tile
 * culling.  Every call on a device shares the array and leaves its contents behind (what must
 * start at zero is cleared by each call): calls that overlap in time on one device race, so
 * serialize them (one stream per device). */
#define SF_NOALLOC_SCRATCH_BYTES (128ull << 20)
int sf_hist_batch_device_noalloc(sf_ctx *ctx, int32_t n_slices, const sf_problem *slices,
                                 double *d_sums, uint64_t *d_counts, void *stream);

/* Autotuner: times candidate launch shapes on these slices (device pointers, as for
 * sf_hist_batch_device; outputs are discarded) -- threads per workgroup, column-record variant,
 * the histogram's privatisation (1D: 4..32 lane-group copies; joint: 1 or 2 copies by lane halves)
 * and joint slot layout, then the column-range length -- keeps the fastest in the context keyed by
 * the call's shape (geometry, D, joint, precision, operator class and order, 2D single-pass
 * planes, log2 of the pair count, n_dist, n_val), and writes it to out (block, j_chunk,
 * blocks_per_cu, variant, copies, layout; the arguments of sf_set_tuning).  Later calls with a
 * matching key and no sf_set_tuning override use it.  (The reference's nearest counterpart picks
 * a tile size from the problem shape, ext/cuda/kernels_2d.jl:171-187.) */
int sf_autotune(sf_ctx *ctx, int32_t n_slices, const sf_problem *slices, int32_t out[6]);

/* Plans: build once, execute many times.  A plan fixes the slices' shapes, bins, operator and
 * precision, and the device pointers x and u of every slice; the data behind those pointers may
 * change between executions (time slices updated in place).  sf_plan_execute enqueues the kernels
 * on `stream` and accumulates into d_sums / d_counts; with use_graph != 0 the kernel sequence is
 * captured into a hipGraph on the first execution for a given (outputs, stream) and replayed
 * afterwards.  sf_plan_sync waits for the last execution; if the mixed kernel's re-bin queue
 * overflowed (outputs then left untouched), it redoes that execution with the binary64 kernel and
 * sets *redone.  One plan serves one execution at a time.
 * (The reference re-stages and re-launches per call, KA ext :630-675; this is the persistent,
 * graph-captured batched path.) */
typedef struct sf_plan sf_plan;
int sf_plan_create(sf_ctx *ctx, int32_t n_slices, const sf_problem *slices, sf_plan **out);
int sf_plan_execute(sf_plan *plan, double *d_sums, uint64_t *d_counts, void *stream,
                    int32_t use_graph);
int sf_plan_sync(sf_plan *plan, int32_t *redone);
void sf_plan_destroy(sf_plan *plan);
/* Fused plan: the outputs of sf_hist_multi_device (at most SF_MAX_FUSED 1D operators and one
 * joint histogram) from one launch per execution, replayable as one hipGraph.  Binary32
 * precisions only, and every output must fuse (FullVector{1} and {3} do not): otherwise
 * SF_E_UNSUPPORTED (make one plan per output).  sf_plan_execute_multi takes one device sums /
 * counts pointer per output (laid out as for sf_hist_multi_device); sf_plan_sync and
 * sf_plan_destroy serve both kinds (after a re-bin queue overflow, sync redoes every output with
 * the binary64 kernel).  The reference's nearest counterpart is its batch-leading loop, which
 * shares a pair's geometry across fields (src/Calculations/batch_leading.jl:105-140). */
int sf_plan_create_multi(sf_ctx *ctx, int32_t n_slices, const sf_problem *slices, int32_t n_out,
                         const sf_output *outs, sf_plan **out);
int sf_plan_execute_multi(sf_plan *plan, double *const *d_sums, uint64_t *const *d_counts, void *stream,
                          int32_t use_graph);

/* Diagnostics of the last successful call on this context (for tests and benchmarks):
 * out[0] = pairs visited, out[1] = pairs re-binned in binary64 by the mixed kernel,
 * out[2] = work items, out[3] = workgroups launched, out[4] = histogram kernel time (ns),
 * out[5] = re-bin kernel time (ns), out[6] = device time from the histogram launch to the
 * final merge (ns),
 * out[7] = threads per workgroup.  Times come from HIP events recorded on the call's stream
 * around the launches; reading them waits for that stream.  Returns 0. */
int sf_last_stats(sf_ctx *ctx, int64_t out[8]);

/* The context's device, for roofline figures: out[0] = compute units, out[1] = peak shader clock
 * (kHz), out[2] = LDS bytes a workgroup may use, out[3] = lanes per wavefront.  No counterpart in
 * the reference (its GPU extensions report no device figures).  Returns 0. */
int sf_device_info(sf_ctx *ctx, int64_t out[4]);

/* Validation only, no device work and no context needed: returns the code sf_hist_batch would
 * return for these slices before launching, and writes the message into msg (if not NULL).
 * Mirrors the reference's pre-loop checks (shapes.jl:67-105, serial.jl:111-120,
 * HelperFunctions.jl:411-417). */
int sf_check(int32_t n_slices, const sf_problem *slices, char *msg, int32_t msg_len);

/* Device self-test of the deterministic binary64 functions the bit-exact path depends on:
 * fn 0: sincosd(a) -> out[2k], out[2k+1];  fn 1: sincos(a) -> out[2k], out[2k+1];
 * fn 2: atan2(a, b) -> out[k];  fn 3: sqrt(a) -> out[k];  fn 4: a / b -> out[k].
 * Host pointers; n values.  (Diagnostic; no reference counterpart.) */
int sf_eval_math(sf_ctx *ctx, int32_t fn, int64_t n, const double *a, const double *b, double *out);

/* Kernel tuning override (autotuner hook): threads per workgroup (= rows of an i-tile; a multiple
 * of 64 in 64..1024), column-range length per work item, workgroups per CU, the column-record
 * variant of the binary32 kernel (1: staged through LDS, 2: scalar loads, 3: B-strips -- 1D slices
 * of one batch that share positions run four fields per pair evaluation), the privatised copies of
 * the LDS histogram (1D: lane groups, a power of two up to 64; joint: 1 or 2, by lane halves) and
 * the joint histogram's slot layout (1: distance-bin rows of n_val slots at an odd stride; 2: rows
 * at a stride of 1 mod 32 slots).  Zero keeps the built-in choice.  Applies to subsequent calls on
 * this context.  (No reference counterpart; the reference's GPU path takes workgroup_size as a
 * keyword, KA ext :385-401.) */
int sf_set_tuning(sf_ctx *ctx, int32_t block, int32_t j_chunk, int32_t blocks_per_cu,
                  int32_t variant, int32_t copies, int32_t layout);

#ifdef __cplusplus
}
#endif

#endif /* SFMI355X_H */
