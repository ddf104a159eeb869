/*
 * halopair.h -- C ABI of the MI355X pair-counting engine (libhalopair.so).
 *
 * Drop-in replacement for the native engines behind
 * halotools.mock_observables.{npairs_3d, npairs_projected, npairs_xy_z,
 * marked_npairs_3d, marked_npairs_xy_z, npairs_per_object_3d,
 * npairs_jackknife_3d, npairs_jackknife_xy_z}.  The reference binds its engines as Cython `def`
 * functions that take a RectangularDoubleMesh plus the six coordinate arrays
 * and return a fresh NumPy array; here the mesh is built on the GPU and every
 * buffer is caller-owned device memory passed as a plain pointer.
 *
 * Conventions
 *   - All data pointers are device pointers (hipMalloc / torch tensors) unless
 *     marked "host".  The library never frees or retains a caller pointer after
 *     the call returns.
 *   - Every entry point only enqueues work on `stream`: no host<->device copy,
 *     no allocation, no synchronisation.  The calls can therefore be captured
 *     into a HIP graph.  Small host-side descriptions (bin edges, mock table)
 *     are carried into the workspace by kernel arguments.
 *   - Scratch memory is the caller's `workspace` of at least
 *     hp_workspace_size() bytes.  Two calls may share a workspace only if they
 *     are ordered on one stream.
 *   - Return value: 0 on success, negative on error; hp_last_error() then
 *     returns a thread-local, human-readable message.
 *   - Positions are FP64.  Counts are computed with the reference's exact FP64
 *     expression order (dx = (x1 - shift) - x2; dsq = dx*dx + dy*dy + dz*dz,
 *     no fused multiply-add), so integer counts are bit-identical to the
 *     reference engine on the same inputs.
 *   - Periodic boxes require every coordinate in [0, period]; a coordinate
 *     outside it (or non-finite) is reported through the `status` output.
 *
 * PC/ abbreviates halotools/mock_observables/pair_counters/ in the reference.
 */
#ifndef HALOPAIR_H
#define HALOPAIR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HP_ABI_VERSION 2
#define HP_MAX_BINS (1 << 20) /* edges per axis (and at most 2^26 histogram cells) */
#define HP_NOALLOC_SCRATCH_BYTES (64u << 20) /* fixed scratch of hp_npairs_3d_noalloc, per GPU */

/* Which reference engine a request replaces. */
enum hp_kind {
    HP_NPAIRS_3D = 0,        /* PC/cpairs/npairs_3d_engine.pyx:17            */
    HP_NPAIRS_PROJECTED = 1, /* PC/cpairs/npairs_projected_engine.pyx:17     */
    HP_NPAIRS_XY_Z = 2,      /* PC/cpairs/npairs_xy_z_engine.pyx:17          */
    HP_MARKED_NPAIRS_3D = 3, /* PC/marked_cpairs/marked_npairs_3d_engine.pyx:22 */
    HP_NPAIRS_PER_OBJECT_3D = 4, /* PC/cpairs/npairs_per_object_3d_engine.pyx:19 */
    HP_MARKED_NPAIRS_XY_Z = 5, /* PC/marked_cpairs/marked_npairs_xy_z_engine.pyx:22 */
    HP_NPAIRS_JACKKNIFE_3D = 6, /* PC/cpairs/npairs_jackknife_3d_engine.pyx:17 */
    HP_NPAIRS_JACKKNIFE_XY_Z = 7 /* PC/cpairs/npairs_jackknife_xy_z_engine.pyx:17 */
};

/* Bit flags of hp_request.flags */
enum hp_flags {
    HP_SAME_SAMPLES = 1, /* sample2 arrays are sample1's (auto-count)                 */
    HP_ROWS1 = 2,        /* marked id 1: write rows1 = d sums / d weights1 per point   */
    HP_ROWS2 = 4,        /* marked id 1: write rows2 = d sums / d weights2 per point   */
    HP_NO_SUMS = 8,      /* marked: skip the weighted sums (rows only)                */
    HP_ALSO_COUNTS = 16, /* marked: also produce the unweighted counts of npairs_3d   */
    HP_TANGENT = 32      /* marked, weight_func_id 1 with nw = 2 weights per point    */
                         /* (w, dw/dtheta): sums of w1*w2 and `tangents`, the sums of */
                         /* t1*w2 + w1*t2 = d(sums)/dtheta (forward-mode derivative)  */
};

/* One realisation ("mock") of a batched request.  A single call is nmocks=1. */
typedef struct hp_mock {
    int64_t n1;        /* points of sample 1 in this mock                          */
    int64_t n2;        /* points of sample 2 (ignored with HP_SAME_SAMPLES)        */
    double period[3];  /* periodic box lengths; for pbc=0 the enclosing box Lbox   */
    int32_t pbc;       /* 1: periodic boundary conditions                          */
    int32_t reserved;
    double cell[3];    /* fine-cell edge of the GPU mesh; <= 0 selects the default */
} hp_mock;

typedef struct hp_request {
    int32_t kind;              /* enum hp_kind                                         */
    int32_t flags;             /* enum hp_flags                                        */
    int32_t nmocks;            /* >= 1                                                 */
    int32_t wfunc;             /* marked: weight_func_id 1..17 (xy_z: 1..15)           */
    const hp_mock* mocks;      /* host [nmocks]; sample points of mock m follow those  */
                               /* of mock m-1 in the coordinate arrays                 */
    const double *x1, *y1, *z1; /* sample 1, n1 total: point i at x1[i * stride1] ...   */
    const double *x2, *y2, *z2; /* sample 2, n2 total (= sample 1 if SAME_SAMPLES)      */
    const double* w1;          /* marked: weights1, row-major (n1_total, nw)           */
    const double* w2;          /* marked: weights2, row-major (n2_total, nw)           */
    int32_t nw;                /* weights per point (id 1,2: 1; 3-11: 2; 12,13: 4; ...) */
    int32_t nr;                /* number of radial / rp edges                          */
    int32_t np;                /* xy_z: number of pi edges; projected: 1               */
    int32_t tile_waves;        /* pair-kernel workgroup: 1, 2, 4 or 8 wavefronts of 64 */
                               /* lane points; 0 = library default (8, fewer if the    */
                               /* histograms do not fit in LDS)                        */
    const double* r2;          /* host [nr]: rbins*rbins (or rp_bins*rp_bins)          */
    const double* p2;          /* host [np]: pi_bins*pi_bins, or {pi_max*pi_max}        */
    double search[3];          /* reserved, not read: the search lengths follow from   */
                               /* the edges (sqrt of the largest r2 / p2)              */
    /* outputs, device */
    int64_t* counts;           /* [nmocks][nr] cumulative (xy_z: [nmocks][nr][np])      */
    double* sums;              /* marked: [nmocks][nr] cumulative weighted sums         */
                               /* (marked xy_z: [nmocks][nr][np]; jackknife:           */
                               /* [nmocks][n_samples + 1][nr] or [..][nr][np])         */
    double* rows1;             /* HP_ROWS1: [n1_total][nr] per-bin (not cumulative)     */
    double* rows2;             /* HP_ROWS2: [n2_total][nr]                              */
    int64_t* evals;            /* optional [nmocks]: candidate pairs evaluated          */
    int32_t* status;           /* optional [1]: 0, or 1 when a point is outside its box */
    void* workspace;           /* device scratch                                       */
    size_t workspace_bytes;
    void* stream;              /* hipStream_t; NULL = default stream                   */
    void* pair_events[2];      /* optional hipEvent_t pair recorded on `stream` right  */
                               /* before and after the pair kernel(s) (profiling)      */
    void* build_events[2];     /* optional hipEvent_t pair recorded right before and   */
                               /* after the cell build (profiling)                     */
    double* tangents;          /* HP_TANGENT: [nmocks][nr] cumulative d(sums)/dtheta    */
    /* jackknife kinds: subsample tag (1..n_samples) of every point, device int32;  */
    /* w1 / w2 are then optional per-point weights (nw = 1), NULL meaning 1       */
    const int32_t* jtags1;
    const int32_t* jtags2;     /* ignored with HP_SAME_SAMPLES                        */
    int32_t n_samples;
    /* doubles between consecutive points of x1, y1, z1 (resp. x2, y2, z2): 0 or  */
    /* 1 for separate contiguous arrays, 3 for the columns of one row-major      */
    /* (N, 3) array (x1 = its start, y1 = x1 + 1, z1 = x1 + 2), read in place    */
    int32_t stride1;
    int32_t stride2;
    int32_t reserved2;
} hp_request;

/* Thread-local message describing the last error (empty string if none). */
const char* hp_last_error(void);

/* ABI version of the loaded library (HP_ABI_VERSION). */
int hp_abi_version(void);

/* Bytes of workspace the request needs (pointers in `req` are not read). */
size_t hp_workspace_size(const hp_request* req);

/* npairs_3d: cumulative int64 counts[m][k] of ordered pairs with
 * (x1-shift-x2)^2+... <= r2[k].  Replaces npairs_3d_engine
 * (PC/cpairs/npairs_3d_engine.pyx:17-184) and the RectangularDoubleMesh build
 * (PC/rectangular_mesh.py:228-374) that feeds it. */
int hp_npairs_3d(const hp_request* req);

/* npairs_3d without a workspace: no counterpart in the reference.  For callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): the same request and results
 * as hp_npairs_3d, with `workspace` NULL and `workspace_bytes` 0; the scratch is
 * a fixed HP_NOALLOC_SCRATCH_BYTES (64 MiB) array that every GPU the library
 * runs on holds from module load until the process ends.  A request fits when
 * hp_workspace_size(req) <= HP_NOALLOC_SCRATCH_BYTES (an auto-count of about
 * 720K points with 20 bins); a larger one returns an error.  The current HIP
 * device's array is used (the stream must belong to that device), and it keeps
 * the last call's contents: calls that overlap on one GPU race. */
int hp_npairs_3d_noalloc(const hp_request* req);

/* npairs_projected: counts[m][k] of pairs with dxy^2 <= r2[k] and
 * dz^2 <= p2[0].  Replaces PC/cpairs/npairs_projected_engine.pyx:17-192. */
int hp_npairs_projected(const hp_request* req);

/* npairs_xy_z: counts[m][k][g] of pairs with dxy^2 <= r2[k], dz^2 <= p2[g].
 * Replaces PC/cpairs/npairs_xy_z_engine.pyx:17-196. */
int hp_npairs_xy_z(const hp_request* req);

/* npairs_per_object_3d: counts[i][k] (int64, [n1_total][nr], input order of
 * sample 1) = number of sample-2 points j of i's mock with dsq_ij <= r2[k].
 * Replaces PC/cpairs/npairs_per_object_3d_engine.pyx:19-208 and its unsorting. */
int hp_npairs_per_object_3d(const hp_request* req);

/* marked_npairs_3d: sums[m][k] of wfunc(w1_i, w2_j) over pairs within r2[k].
 * Replaces PC/marked_cpairs/marked_npairs_3d_engine.pyx:22-218 together with
 * the marking functions of PC/marked_cpairs/marking_functions.pyx:14-217.
 * With wfunc == 1 and HP_ROWS1/HP_ROWS2 it also writes, for every point, the
 * per-bin derivative of the (non-cumulative) bin sums with respect to that
 * point's weight: rows1[i][s] = sum of w2_j over partners j in bin s.  This is
 * the Jacobian the backward pass contracts (no counterpart in the reference). */
int hp_marked_npairs_3d(const hp_request* req);

/* marked_npairs_xy_z: sums[m][k][g] of wfunc(w1_i, w2_j) over pairs with
 * dxy^2 <= r2[k] and dz^2 <= p2[g] (cumulative along both axes).  Replaces
 * PC/marked_cpairs/marked_npairs_xy_z_engine.pyx:22-228 with the marking
 * functions 1-15 it dispatches to (:231-263).  No flags beyond
 * HP_SAME_SAMPLES. */
int hp_marked_npairs_xy_z(const hp_request* req);

/* npairs_jackknife_3d: sums[m][s][k] = sum over pairs within r2[k] of the
 * reference's jackknife weight of subsample s (0: w1*w2 for every pair; s >= 1:
 * w1*w2 when neither point carries tag s, w1*w2/2 when exactly one does, 0
 * when both do).  Replaces PC/cpairs/npairs_jackknife_3d_engine.pyx:17-236
 * (jweight :239-287).  npairs_jackknife_xy_z: the same over (rp, pi) cells,
 * sums[m][s][k][g]; replaces PC/cpairs/npairs_jackknife_xy_z_engine.pyx:17-246.
 * Sums of unit weights (the default) are exact; with other weights each
 * subsample row is row 0 minus the weight that subsample removes, summed with
 * atomics (last-bit order effects).  A tag outside [1, n_samples] sets bit 2
 * of the status word. */
int hp_npairs_jackknife_3d(const hp_request* req);
int hp_npairs_jackknife_xy_z(const hp_request* req);

/* Backward of the product-weighted marked sums (no counterpart in the
 * reference): grad[p] = sum_s rows[p][s] * gsuffix[mock(p)][s], where
 * gsuffix[m][s] = sum_{k>=s} dL/dsums[m][k].  `offsets` (host, nmocks+1)
 * delimits each mock's points (rows has offsets[nmocks] rows); rows, gsuffix
 * and grad are device arrays. */
int hp_marked_backward(const double* rows, int32_t nr, const double* gsuffix,
                       const int64_t* offsets, int32_t nmocks, double* grad, void* stream);

/* Per-mock sums (no counterpart in the reference; the batched model-fitting
 * step folds per-point gradients into per-mock parameters with it):
 * out[m] = sum of values[offsets[m] .. offsets[m+1]) in a fixed order (no
 * atomics: bitwise reproducible).  `offsets` is host memory; values, out and
 * scratch are device arrays; scratch holds hp_segment_sum_scratch(offsets,
 * nmocks) doubles (< 0 on error). */
int64_t hp_segment_sum_scratch(const int64_t* offsets, int32_t nmocks);

/* Zheng07 central weights and their logMmin tangents in one pass (no
 * counterpart among the reference's pair counters: the batched model-fitting
 * step's weight model, empirical_models/occupation_models/
 * zheng07_components.py:149-153): for every point i of mock m = mock_id[i],
 * u = (log_mass[i] - logMmin[m]) / sigma, wt[2i] = 0.5 * (1 + erf(u)),
 * wt[2i+1] = -exp(-u*u) / (sqrt(pi) * sigma).  log_mass, mock_id, logMmin
 * and wt are device arrays; the call only enqueues work on `stream`. */
int hp_zheng07_tangent(const double* log_mass, const int64_t* mock_id, const double* logMmin, double sigma,
                       int64_t n, double* wt, void* stream);
int hp_segment_sum(const double* values, const int64_t* offsets, int32_t nmocks, double* out,
                   double* scratch, int64_t scratch_elems, void* stream);

/* Diagnostics (host only, no GPU needed; used by the CPU test-suite). */
/* The kernels' bucketed bin lookup: out[i] = number of suffix-minimum edges of r2 below d[i];
 * info = {shift, base, table entries, multi}. */
int hp_debug_bin_lookup(const double* r2, int32_t K, const double* d, int64_t n, int32_t* out, int32_t* info);
/* Hilbert index of cell (x, y, z) on a 2^bits grid: the lane-group order of the pair kernels. */
uint32_t hp_debug_hilbert(uint32_t x, uint32_t y, uint32_t z, int32_t bits);

#ifdef __cplusplus
}
#endif

#endif /* HALOPAIR_H */
