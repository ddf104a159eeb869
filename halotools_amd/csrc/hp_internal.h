// hp_internal.h -- device-side layout shared by the HIP translation units of
// libhalopair.so.  Not part of the public ABI (see include/halopair.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/halopair.h"

namespace hp {

constexpr int kWave = 64;            // CDNA wavefront width
constexpr int kGroup = 64;           // lane-side points per wave work item (Hilbert order)
constexpr int kTableMax = 512;       // bucket table entries per axis
constexpr int kRadixBits = 8;        // digit width of the LSD counting sort
constexpr int kRadixDigits = 1 << kRadixBits;
constexpr int kSortTile = 2048;      // keys per counting-sort tile (one workgroup of 4 waves)
constexpr int kSortTileLarge = 8192; // ... for sorts of at least kSortLargeMin keys (8 waves)
constexpr int64_t kSortLargeMin = 4 << 20;

// Per-mock geometry, resident in the workspace.  "side" 0 is sample 1, side 1
// is sample 2.  Cells of the GPU mesh are shared by both sides.
struct alignas(64) MockGeom {
    double period[3];     // box lengths (PBC period, or enclosing box)
    double cell[3];       // fine-cell edge per dimension
    int32_t ndiv[3];      // cells per dimension
    int32_t pbc;
    int32_t mbits;        // bits per dimension of the lane-ordering (Hilbert) grid
    int32_t pad;
    int64_t p_off[2];     // first point of this mock on each side
    int64_t n[2];         // points of this mock on each side
    int64_t cell_off;     // first global cell id of this mock
    int64_t key_off;      // first global Hilbert key of this mock
    int64_t group_off[2]; // first lane group of this mock, per lane side
    int64_t ngroups[2];
    float search[3];      // search half-widths (window), with safety margin
    float eps;            // absolute safety margin of the f32 window planning
    float r2xy;           // squared planar search radius (with margin)
    int32_t shape;        // 0: sphere (3-D), 1: cylinder (rp, pi)
    double inv_cell[3];   // 1 / cell: the key pass multiplies (the windows carry a margin)
    double inv_hcell[3];  // 2^mbits / period: cells per unit length of the lane-ordering grid
};

// Bucketed bin lookup for an ascending edge array e[0..K-1], e[K] = +inf.
// bucket(d) = min(sat(hi32(d) - (base << shift)) >> shift, nb-1); with at most
// one edge inside any bucket (multi == 0) the bin start is
// table[b].k + (table[b].edge < d); otherwise a short scan finishes the job.
struct BinSpec {
    int32_t nbins;   // K
    int32_t shift;
    int32_t base;    // >= 0
    uint32_t basesh; // base << shift
    int32_t nb;      // table entries
    int32_t multi;   // 1 if some bucket holds more than one edge
    int64_t edges_off;  // byte offset of e[0..K+1] in the workspace
    int64_t table_off;  // byte offset of TabEntry table[nb]
};

// One bucket of the table: k = number of edges strictly below the bucket's
// smallest value, edge = e[k] (the only edge that can lie inside the bucket).
struct alignas(16) TabEntry {
    double edge;
    uint32_t k;
    uint32_t pad;
};

// Counting-sort scratch of one sort
struct SortBufs {
    uint32_t* keys[2];
    uint32_t* vals[2];
    uint32_t* tile_counts;   // [digits][ntiles]
    uint32_t* digit_totals;  // [digits]
};

// Tiles of one side's sort: tiles never cross from one mock into the next.
struct SortTiles {
    const uint32_t* first = nullptr;  // [nmocks + 1]: first tile of each mock, then the tile count
    int tile = 0;                     // keys per tile (kSortTile or kSortTileLarge)
    int ntiles = 0;
};

enum Order : int { kHilbert = 0, kRowMajor = 1 };

// Sorted copy of one side in one order.
struct SortedSide {
    double* x;
    double* y;
    double* z;
    double* w;          // (n, nw) row-major, or nullptr
    uint32_t* perm;     // sorted position -> original index
    uint32_t* cell_start;  // row-major order only: [total_cells + 1]
    int64_t xs;         // doubles between consecutive points of x, y, z (4: the sorted 32-byte rows)
};

// One side's coordinates (and up to four weights) packed as fixed-size rows
// in input order: written by the side's first key pass (which reads the
// coordinates anyway) and read by the row-major gather, so that gather reads
// one row per point at a random position instead of one element from each
// of the x, y, z (and weight) arrays.  The fourth double of a row holds, in
// its low word, the point's position in the Hilbert order (written by the
// Hilbert build's last radix pass when a symmetric auto-count needs it).
struct PackedRows {
    double* p = nullptr;  // [n][stride], or nullptr: no packed rows
    int stride = 4;       // doubles per row: 4 (x, y, z, slot) or 8 (x, y, z, slot, w0..w3)
    int nw = 0;           // weights packed from the fifth double on (0: gathered from w)
};

// ------------------------------------------------------------------ launchers
// cellbuild.hip
int launch_store_bytes(void* dst, const void* src, size_t nbytes, hipStream_t s);
int launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s);
int launch_group_mock(const MockGeom* g, int nmocks, int side, int32_t* group_mock,
                      int64_t ngroups_total, hipStream_t s);
// The side's key pass (the first build of the side) writes its packed rows
// (`rows`, when rows.p is set and keys_in is null; a Hilbert key pass puts
// each point's Hilbert key in its row).  The row-major build's sorted copy
// is one 32-byte row per point (out.x = x of row 0, out.y = out.x + 1,
// out.z = out.x + 2).  Symmetric auto-counts also need, for every point of
// the row-major order, its place in the Hilbert order: with want_hkey the
// gather reads the key through its permutation and writes (Hilbert key << 32)
// | input index into the row's fourth word, which orders the points of a mock
// as the stable Hilbert sort does.  When one side is sorted in both
// orders, the Hilbert build also writes the row-major keys (keys_row_out) and
// the row-major build starts from them (keys_in).  Null pointers skip each of
// these.
// Keys per sort tile for a sort of n keys (HALOPAIR_SORT_TILE = 2048 or 8192
// forces one size, for tests and measurements).
int sort_tile(int64_t n);
int build_sorted_side(const MockGeom* geom, int nmocks, int side, Order order,
                      const double* x, const double* y, const double* z, const double* w,
                      int nw, int64_t n, int64_t total_cells, int key_bits,
                      SortBufs sb, SortTiles st, SortedSide out, int32_t* status, hipStream_t s, PackedRows rows,
                      bool want_hkey = false, uint32_t* keys_row_out = nullptr,
                      const uint32_t* keys_in = nullptr, int64_t xs = 1);

// pairs.hip
struct PairLaunch {
    int mode;              // kModeCount3D ...
    int swap;              // 1: lanes are sample 2
    int also_count;        // marked: also accumulate unweighted counts
    int wfunc;             // marked weight function id
    int waves;             // wavefronts per workgroup (1, 2, 4, 8); 0 = default
    const MockGeom* geom;
    const int32_t* group_mock;
    int64_t ngroups;
    SortedSide lane;       // lane side: input-order x/y/z/w plus the Hilbert permutation
    SortedSide strm;       // row-major sorted stream side
    int nw;
    BinSpec br;            // radial / rp bins
    BinSpec bp;            // pi bins (xy_z) or pi_max (projected, nbins = 1)
    const void* ws;        // workspace base (edges / tables)
    uint64_t* pcount;      // [nmocks][crep][nbins_count], zeroed, accumulated with atomics
    int crep;              // count replicas per mock (spread the atomics over cache lines)
    double* psum;          // [ngroups][nr]
    double* ptan;          // tangent mode: [ngroups][nr]
    int64_t* orows;        // per-object mode: [n_lane_total][nr] cumulative counts (input order)
    double* rows;          // [n_lane_total][nr] (original order) or nullptr
    double* corr;          // auto-counts: rows2 - rows1 corrections from seam pairs, or nullptr
    uint64_t* pevals;      // [ngroups]
    // symmetric auto-counts: the Hilbert order keys of the stream points, the
    // fourth word of each 32-byte stream row ((Hilbert key << 32) | input
    // index; hkey[4 * j]), or null; each unshifted unordered pair is then
    // evaluated once and counted twice.  hrows/hrow_stride: the lane side's
    // packed rows (the Hilbert key of each point in its fourth word), for the
    // key of the point that follows a lane group
    const uint64_t* hkey;
    const double* hrows;
    int hrow_stride;
    // jackknife modes: per-mock subsample tables [nmocks][nsamp + 1][nbins]
    // (row 0: all pairs; row s: weight lost when subsample s is removed),
    // accumulated with atomics; lane and stream weights are (w, tag) rows
    double* jacc;
    int nsamp;
    int jslots;            // LDS rows per wave for subsample tags (<= kJackSlots)
    // persistent launches: one work-queue counter per XCD (8 words, zeroed
    // by launch_pairs before every launch)
    uint32_t* queue;
};
enum Mode : int { kModeCount3D = 0, kModeProj = 1, kModeXYZ = 2, kModeWSum = 3, kModeJac = 4, kModeTan = 5,
                  kModeObj = 6, kModeWXYZ = 7, kModeJack = 8, kModeJackXY = 9 };
constexpr int kJackSlots = 16;  // most subsample tags a wave keeps in LDS
int launch_pairs(const PairLaunch& L, hipStream_t s);

// reduce.hip
int launch_prefix_counts(const uint64_t* acc, int nr, int np, int nmocks, int crep, int64_t* counts, hipStream_t s);
int64_t segment_sum_scratch(const int64_t* offsets, int nmocks);
int launch_segment_sum(const double* v, const int64_t* offsets, int nmocks, double* out, double* scratch,
                       hipStream_t s);
int launch_rows_dot(const double* rows, int nr, const double* gsuffix, const int64_t* offsets,
                    int nmocks, double* grad, hipStream_t s);
int launch_add_rows(const double* a, const double* b, double* out, int64_t n, hipStream_t s);
int launch_pack_jack(const double* w, const int32_t* tags, int64_t n, int nsamp, double* out, int32_t* status,
                     hipStream_t s);
int launch_jack_finish(double* acc, int nr, int np, int nmocks, int nsamp, double* out, hipStream_t s);
int launch_zheng07_tangent(const double* logm, const int64_t* mock_id, const double* logMmin, double sigma,
                           int64_t n, double* wt, hipStream_t s);

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ int64_t mock_of_point(const MockGeom* __restrict__ g, int nmocks, int side,
                                                 int64_t p) {
    int lo = 0, hi = nmocks - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (g[mid].p_off[side] <= p) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Index of cell (x, y, z) along a 3-D Hilbert curve over a 2^b grid, b <= 10.
// Method: J. Skilling, "Programming the Hilbert curve", AIP Conf. Proc. 707,
// 381 (2004) -- AxesToTranspose, then interleave the transposed bits.  Unlike
// a Morton code the curve never jumps: consecutive cells share a face, so any
// run of consecutive points stays compact.
__host__ __device__ __forceinline__ uint32_t hilbert3(uint32_t x, uint32_t y, uint32_t z, int b) {
    if (b <= 0) return 0u;
    uint32_t X[3] = {x, y, z};
    const uint32_t M = 1u << (b - 1);
    for (uint32_t Q = M; Q > 1; Q >>= 1) {
        const uint32_t P = Q - 1;
        for (int i = 0; i < 3; i++) {
            if (X[i] & Q) {
                X[0] ^= P;
            } else {
                const uint32_t t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    X[1] ^= X[0];
    X[2] ^= X[1];
    uint32_t t = 0;
    for (uint32_t Q = M; Q > 1; Q >>= 1)
        if (X[2] & Q) t ^= Q - 1;
    X[0] ^= t;
    X[1] ^= t;
    X[2] ^= t;
    uint32_t h = 0;
    for (int k = b - 1; k >= 0; k--)
        h = (h << 3) | (((X[0] >> k) & 1u) << 2) | (((X[1] >> k) & 1u) << 1) | ((X[2] >> k) & 1u);
    return h;
}

}  // namespace hp
