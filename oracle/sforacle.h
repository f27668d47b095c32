/* sforacle.h -- CPU oracle for the O(N^2) pair-histogram path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, in plain scalar C++, the reference's
 * SerialBackend pair loops (src/Calculations/serial.jl, src/Calculations/serial_2d.jl) and the
 * per-pair geometry, operators and binning they call.  It is the checker the gfx950 kernels are
 * compared against, and the CPU baseline timed by bench.py.  Product code never links it.
 *
 * Arrays are the reference's column-major layouts: x is (W, N), u is (D, N), element [d, i] at
 * d + W*i.  Outputs are accumulated (+=) like calculate_structure_function! (dispatch.jl:346-373).
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* geometry codes: pair_geometry_for(metric) (src/HelperFunctions.jl:390-393) */
#define SFO_GEOM_FLAT 0         /* Distances.Euclidean            -> FlatGeometry{D}            */
#define SFO_GEOM_SPHERE_DEG 1   /* Distances.Haversine(radius)    -> SphericalGeometry, degrees */
#define SFO_GEOM_SPHERE_RAD 2   /* Distances.SphericalAngle()     -> SphericalGeometry, radians */

/* operator codes: src/StructureFunctionTypes.jl */
#define SFO_OP_PROJECTED 0      /* ProjectedStructureFunctionType{nl, nt}              :43-89   */
#define SFO_OP_S2 1             /* SecondOrderStructureFunctionType                    :97,120  */
#define SFO_OP_S3 2             /* ThirdOrderStructureFunctionType                     :106,122 */
#define SFO_OP_T2COMP 3         /* TransverseComponentSecondOrderStructureFunctionType :137,147 */
#define SFO_OP_L1T2COMP 4       /* LongitudinalTransverseComponentThirdOrder...Type    :145,150 */
#define SFO_OP_FULLVEC 5        /* FullVectorStructureFunctionType{nl} (nl = NF)       :114,125 */

/* A LinearBinEdges axis (src/BinEdges.jl:94-131), optionally inside InfPaddedBinEdges
 * (:270-345): the fields of the reference's struct.  With a rule the axis is digitized by the
 * reference's O(1) rule (searchsortedfirst(::LinearBinEdges), :114-131) instead of a search of
 * the edge vector; the flat squared-distance plan applies it to sqrt(r^2) (SquaredLinearPlan,
 * :412-417, 506, 516). */
typedef struct {
  int32_t kind;           /* 1: LinearBinEdges                                         */
  int32_t padded;         /* InfPaddedBinEdges around the range                        */
  int64_t n;              /* length of the range                                       */
  double first_edge, last_edge, inv_step, step_val;
} sfo_bin_rule;

typedef struct {
  int32_t D;              /* velocity components, 2 or 3                               */
  int32_t W;              /* coordinate width: D (flat) or 2 (sphere lon, lat)         */
  int64_t N;              /* points                                                    */
  const double *x;        /* (W, N) column-major                                       */
  const double *u;        /* (D, N) column-major                                       */
  int32_t geometry;       /* SFO_GEOM_*                                                */
  double radius;          /* sphere radius (Haversine); 1 for SphericalAngle           */
  int32_t op, nl, nt;     /* SFO_OP_*; exponents for SFO_OP_PROJECTED / FULLVEC        */
  int32_t n_dist;         /* distance bins; dist_edges has n_dist + 1 entries          */
  const double *dist_edges;
  int32_t n_val;          /* 0 => 1D histogram; else value bins (joint histogram)      */
  const double *val_edges;
  int32_t batch_leading;  /* flat pairs as the batch-leading kernels evaluate them      */
  int32_t reserved;       /* (fields with auxiliary axes, batch_leading.jl:105-330):
                             r = sqrt(dx.dx) binned against the edges, values from the
                             operator on rhat = dx / r; 0: the point-field kernels     */
  const sfo_bin_rule *dist_rule;  /* NULL: search dist_edges                              */
  const sfo_bin_rule *val_rule;   /* NULL: search val_edges                               */
} sfo_problem;

/* digitize(x, rule) = searchsortedfirst - 1 for one query (tests of the rule itself) */
int64_t sfo_rule_digitize(const sfo_bin_rule *r, double x);

/* Histogram over outer rows i = i_begin, i_begin + i_step, ... < i_end (0-based), pairs (i, j>i).
 * nthreads > 1 splits those rows round-robin over threads with per-thread histograms reduced by
 * addition (ext/StructureFunctionsOhMyThreadsExt.jl:33-34,171-193).  compensated != 0 accumulates
 * each bin with Neumaier compensation (the reference adds naively, serial.jl:95-96).  abs_sums
 * (may be NULL) receives sum |value| per bin and mag_sums (may be NULL) sum ||du||^order, the
 * magnitude scale of the values (error bounds in tests).  Outputs have n_dist (1D) or
 * n_dist * n_val (joint, column-major) entries and are accumulated into.
 * Returns 0, or a negative code for invalid input (nothing is written then). */
int sfo_hist(const sfo_problem *p, int nthreads, int compensated, int64_t i_begin, int64_t i_end,
             int64_t i_step, double *sums, uint64_t *counts, double *abs_sums,
             double *mag_sums);

/* Single pass over the pairs accumulating the six invariants S2, L2, T2, S3, L3, L1T2
 * (calculate_structure_functions_single_pass, src/Calculations/serial_single_pass.jl:602-700).
 * p->op / value bins are ignored.  sums and counts are (6, n_dist) column-major; rows S2, L2, S3,
 * L3 and count row 1 accumulate, then T2 = S2 - L2, L1T2 = S3 - L3 and count rows 2..6 = count
 * row 1 are assigned over the whole buffer (_sp1d_derive_rows!, serial_single_pass.jl:208-219). */
int sfo_hist_single_pass(const sfo_problem *p, int nthreads, int64_t i_begin, int64_t i_end,
                         int64_t i_step, double *sums, uint64_t *counts);

/* Single pass, 2D (joint) form (calculate_structure_functions_single_pass_2d,
 * src/Calculations/serial_single_pass.jl:1161-1181): each counted pair's six invariants
 * (|du|^2, du_L^2, |du|^2 - du_L^2, du_L |du|^2, du_L^3, du_L (|du|^2 - du_L^2)) are binned by
 * distance and each by its own value edges.  p->n_val is the number of value bins; val_edges6
 * holds the six edge vectors back to back (n_val + 1 each); p->val_edges is ignored.  sums and
 * counts are (6, n_dist, n_val) column-major, accumulated into; abs_sums / mag_sums (may be NULL)
 * receive sum |value| and sum m^order (order 2 for the first three invariants, 3 else) with m the
 * magnitude scale of binary32 errors: ||du|| (flat), (||u_i|| + ||u_j||)(1 + c^2/sin(sigma)) on the
 * sphere, whose frame transport makes errors absolute in the velocities' size. */
int sfo_hist_single_pass_2d(const sfo_problem *p, int nthreads, int64_t i_begin, int64_t i_end,
                            int64_t i_step, const double *val_edges6, double *sums,
                            uint64_t *counts, double *abs_sums, double *mag_sums);

/* deterministic elementary functions (structurefunctions.jl_amd/csrc/sfmath.h), for tests */
void sfo_sincosd(double x, double *s, double *c);
void sfo_sincos(double x, double *s, double *c);
double sfo_atan2(double y, double x);

/* single-pair evaluation for tests: returns ok flag; writes distance, value, dbin, vbin */
int sfo_pair(const sfo_problem *p, int64_t i, int64_t j, double *distance, double *value,
             int64_t *dbin, int64_t *vbin);

#ifdef __cplusplus
}
#endif
