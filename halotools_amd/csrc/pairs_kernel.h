// pairs.hip -- cell-pair tile kernels (replace the j-loops of
// PC/cpairs/npairs_3d_engine.pyx:103-182, PC/cpairs/npairs_projected_engine.pyx:103-190,
// PC/cpairs/npairs_xy_z_engine.pyx:103-194,
// PC/marked_cpairs/marked_npairs_3d_engine.pyx:103-216 and
// PC/marked_cpairs/marked_npairs_xy_z_engine.pyx:103-226).
//
// Work decomposition (MI355X-first, not the reference's cell loop):
//   * one wavefront = one "group" of 64 lane points that are consecutive in the
//     Hilbert-sorted copy of the lane side, hence spatially compact;
//   * the wave reduces its points' bounding box, walks the (x, y) columns of
//     fine cells within the search radius of that box, and for each column
//     visits one contiguous z-run of the row-major sorted stream side (split
//     at the periodic seam);
//   * stream points are wave-uniform, so they arrive through scalar loads and
//     sit in SGPRs; every lane evaluates its own point against them;
//   * the bin start of each pair comes from a bucketed edge table in LDS and
//     one exact FP64 comparison; histograms are per lane in LDS (bin-major,
//     bank-conflict free) and are reduced once per group in a fixed order.
//
// Exactness: every pair is evaluated with the reference's arithmetic,
// x1tmp = x1 - x2shift (x2shift in {-L, 0, +L}, chosen by which periodic image
// of the stream cell is visited), dx = x1tmp - x2, dsq = dx*dx + dy*dy + dz*dz,
// with FP contraction off, and binned with the same "<=" semantics.  A pair
// within the search radius is visited exactly once with its minimum image;
// other images of it lie at least L - rmax >= 2 rmax away and bin nowhere.
#pragma once

#include <algorithm>

#include "hp_internal.h"

#pragma clang fp contract(off)

namespace hp {

// ------------------------------------------------------------- marking functions
// PC/marked_cpairs/marking_functions.pyx:14-217 (ids 1-17).
__device__ __forceinline__ double wfunc_eval(int id, const double* a, const double* b) {
    switch (id) {
        case 1: return a[0] * b[0];
        case 2: return a[0] + b[0];
        case 3: return a[0] == b[0] ? a[1] * b[1] : 0.0;
        case 4: return a[0] != b[0] ? a[1] * b[1] : 0.0;
        case 5: return b[0] > a[0] ? a[1] * b[1] : 0.0;
        case 6: return b[0] < a[0] ? a[1] * b[1] : 0.0;
        case 7: return b[0] > (a[0] + a[1]) ? b[1] : 0.0;
        case 8: return b[0] < (a[0] + a[1]) ? b[1] : 0.0;  // '+' as in the reference (:104-115)
        case 9: return fabs(a[0] - b[0]) < a[1] ? b[1] : 0.0;
        case 10: return fabs(a[0] - b[0]) > a[1] ? b[1] : 0.0;
        case 11: return b[0] > a[0] * a[1] ? b[1] : 0.0;
        case 12: return a[0] * b[0] * (a[1] * b[1] + a[2] * b[2] + a[3] * b[3]);
        case 13: {
            const double c = a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
            return a[0] * b[0] * (a[1] * b[1] + a[2] * b[2] + a[3] * b[3]) * c;
        }
        case 14: return a[0] * b[0] * (a[1] * b[1] + a[2] * b[2]);
        case 15: {
            const double c = a[1] * b[1] + a[2] * b[2];
            return a[0] * b[0] * (a[1] * b[1] + a[2] * b[2]) * c;
        }
        case 16: {
            if (a[4] != b[4]) return 0.0;
            const double c = a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
            return a[0] * b[0] * (a[1] * b[1] + a[2] * b[2] + a[3] * b[3]) * c;
        }
        case 17: {
            if (a[4] == b[4]) return 0.0;
            const double c = a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
            return a[0] * b[0] * (a[1] * b[1] + a[2] * b[2] + a[3] * b[3]) * c;
        }
        default: return 0.0;
    }
}

// ------------------------------------------------------------- helpers
__device__ __forceinline__ float f32_down(double v) {
    float f = (float)v;
    if ((double)f > v) f = nextafterf(f, -__builtin_inff());
    return f;
}
__device__ __forceinline__ float f32_up(double v) {
    float f = (float)v;
    if ((double)f < v) f = nextafterf(f, __builtin_inff());
    return f;
}

__device__ __forceinline__ float wave_min(float v) {
    for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
// butterfly: every lane ends with bitwise the same, order-fixed sum
__device__ __forceinline__ double wave_sum_f64(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ float unif(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// LDS copy of a bucket entry: the exact edge to compare with, and the byte
// offset of the histogram row the pair lands in when it does not exceed it.
struct alignas(16) LdsEntry {
    double edge;
    uint32_t addr;
    uint32_t k;
};

// Row offset of start bin s within a launch that covers bins [k0, k0 + kc):
// bins outside it go to the trash row kc (a later or earlier launch counts
// them).  Launches over a range of bins (k0 > 0 or kc < all bins) take the
// scanning lookup (MULTI), which maps each pair's bin through this.
__host__ __device__ __forceinline__ uint32_t chunk_addr(uint32_t s, uint32_t k0, uint32_t kc, uint32_t step) {
    return (s >= k0 && s < k0 + kc) ? (s - k0) * step : kc * step;
}
static_assert(sizeof(LdsEntry) == 16, "bucket entries are read as one 16-byte LDS load");

// Byte offset of the first bin a value d falls in (bin k iff d <= e[k]):
// bucket from the high word of d, one exact FP64 comparison with the only
// edge that can lie inside the bucket (MULTI == 0), or a scan otherwise.
template <int MULTI>
__device__ __forceinline__ uint32_t bin_addr(double d, const LdsEntry* __restrict__ tab, uint32_t shift,
                                             uint32_t basesh, uint32_t nbm1, uint32_t step,
                                             const double* __restrict__ gedges, uint32_t k0, uint32_t kc) {
    const uint32_t hi = (uint32_t)__double2hiint(d);
    const uint32_t b = min(__builtin_elementwise_sub_sat(hi, basesh) >> shift, nbm1);
    const LdsEntry e = tab[b];
    if (!MULTI) return e.addr + (e.edge < d ? step : 0u);
    uint32_t k = e.k;
    while (gedges[k] < d) ++k;
    return chunk_addr(k, k0, kc, step);
}

#ifndef HP_BATCH
#define HP_BATCH 8
#endif
constexpr int kBatch = HP_BATCH;  // stream points per scalar-load wait
#ifndef HP_SPLIT
#define HP_SPLIT 1
#endif
#ifndef HP_SPLIT_CHUNK
#define HP_SPLIT_CHUNK 2
#endif
#ifndef HP_SPLIT_XY
#define HP_SPLIT_XY 1
#endif
#ifndef HP_COUNT_WAVES
#define HP_COUNT_WAVES 7  // waves per SIMD the count kernels' registers are held to
#endif
#ifndef HP_PERSIST
#define HP_PERSIST 1  // persistent workgroups (FP64-histogram modes) taking lane groups from work queues
#endif

// Read-only global data seen through the constant address space: uniform
// loads from it always become scalar loads (SGPR results, no VGPRs).
typedef const __attribute__((address_space(4))) char cchar;
__device__ __forceinline__ double sload_f64(const void* base, uint32_t byte_off) {
    cchar* const p = (cchar*)(const char*)base + byte_off;  // address-space cast
    return *(const __attribute__((address_space(4))) double*)p;
}

__device__ __forceinline__ double sload_f64_at(const double* base, int64_t i) {
    return ((const __attribute__((address_space(4))) double*)(const char*)base)[i];
}
__device__ __forceinline__ uint32_t sload_u32_at(const uint32_t* base, int64_t i) {
    return ((const __attribute__((address_space(4))) uint32_t*)(const char*)base)[i];
}

struct PairArgs {
    PairLaunch L;
    int wave_bytes;     // histogram bytes per wave
    int cnt_off;        // marked + counts: byte offset of the count rows
    int shared_count;   // 1: one histogram per wave instead of per lane
    int smem;           // dynamic LDS bytes
    int ring_off;       // byte offset of the survivor ring within a wave's LDS
    int tab_bytes;      // LDS bytes of the bucket tables, ahead of the wave areas
    int jrow;           // jackknife: bytes of one subsample row of the wave's table
    int jr0_off;        // 3-D jackknife: byte offset of the all-pairs rows in 16 lane columns
    int k0, kc;         // this launch histograms the radial / rp start bins [k0, k0 + kc)
    int grid;           // persistent launch: compute units of the device
    uint64_t flush_at;  // shared u32 histograms: flush once this many increments may have landed
};

// Survivor ring per wave: up to kBatch-1 leftover entries plus one 64-point chunk.
constexpr int kRing = (kBatch - 1 + 64 + 3) & ~3;
static_assert(kRing >= kBatch - 1 + 64, "survivor ring too small for the batch size");

// Occupancy the register allocation is held to (waves per SIMD): the count
// kernels are limited by VGPRs otherwise, the FP64-histogram kernels by LDS.
template <int MODE, int VAR = 0>
constexpr int kWavesPerSimd =
    (MODE == kModeJac) ? 1
                       // product marks (compiled in) need fewer registers than the generic marking functions
                       : (((MODE == kModeWSum || MODE == kModeWXYZ) && VAR) ? 5
                          : ((MODE == kModeWSum || MODE == kModeTan || MODE == kModeWXYZ || MODE == kModeJack ||
                              MODE == kModeJackXY)
                                 ? 4
                                 : (MODE == kModeXYZ ? 5 : (MODE == kModeObj ? 6 : HP_COUNT_WAVES))));

// per-object counts: one u32 row set per lane (each lane owns its point's counts)

// Modes whose workgroups are persistent (measured on MI355X: the FP64-histogram
// passes gain 2-7 %; the integer count passes lose 8-26 %, their register
// budget leaving no room for the work loop's state without spilling).
#ifndef HP_SPREAD
#define HP_SPREAD 1  // spread lane order for the modes with wave-shared histogram rows
#endif
#ifndef HP_QUEUES
#define HP_QUEUES 8  // work queues of a persistent launch (one per XCD)
#endif
#ifndef HP_XCD_CHUNK
#define HP_XCD_CHUNK 64  // non-persistent modes: blocks per XCD run (0: blocks in dispatch order)
#endif
template <int MODE>
constexpr bool kPersist = HP_PERSIST && (MODE == kModeWSum || MODE == kModeJac || MODE == kModeTan ||
                                         MODE == kModeWXYZ || MODE == kModeJack || MODE == kModeJackXY);

// VAR selects a variant of the mode: the Jacobian pass with the lanes on
// sample 2 (kModeJac), or the product marking function w1[0] * w2[0] with one
// weight per point compiled in (kModeWSum, kModeWXYZ).
template <int MODE, int VAR, int ALSO, int MULTI, int WAVES>
__global__ void __launch_bounds__(WAVES * 64, (kWavesPerSimd<MODE, VAR>)) k_pairs(const PairArgs A) {
    constexpr int SWAP = (MODE == kModeJac) ? VAR : 0;
    constexpr bool PROD = (MODE == kModeWSum || MODE == kModeWXYZ) && VAR;
    // jackknife: FP64 tables of (w1 w2) per bin, one row for all pairs and one
    // per subsample tag the wave meets (weights are (w, tag) rows)
    constexpr bool JK = (MODE == kModeJack || MODE == kModeJackXY);
    constexpr bool F64 = (MODE == kModeWSum || MODE == kModeJac || MODE == kModeTan || MODE == kModeWXYZ || JK);
    // 2-D (rp, pi) binning: one histogram per wave, [Kr + 1][Kp + 1] cells
    // (u32 counts, or FP64 weighted sums), the last row / column being trash
    constexpr bool XY = (MODE == kModeXYZ || MODE == kModeWXYZ || MODE == kModeJackXY);
    // Modes whose histogram rows are shared by every lane of the wave (the
    // (rp, pi) histograms, the jackknife loss rows) hand out a full group's
    // points to the lanes in a spread order: lane l holds the point at
    // position 4 (l mod 16) + l / 16 of the group's Hilbert run, so the 16 or
    // 32 lanes that one LDS pass serves hold points from across the whole
    // group, which fall into different (rp, pi) cells more often than
    // neighbours along the curve do (fewer lanes of a pass on one address).
    // Partial groups keep the identity order.  The symmetric self block reads
    // stream point k as the point lane k holds, so pairs keep their lane order.
    constexpr bool SPREAD = HP_SPREAD && (XY || MODE == kModeJack);
    auto spread_pos = [&](int l, int n) __attribute__((always_inline)) {
        return (SPREAD && n == kGroup) ? (((l & 15) << 2) | (l >> 4)) : l;
    };
    constexpr uint32_t cell_bytes = (MODE == kModeWXYZ || MODE == kModeJackXY) ? 8u : 4u;
    // LDS: the bucket tables (sized to this request's buckets), then one area per wave
    extern __shared__ __align__(16) unsigned char lds[];
    LdsEntry* const s_tr = reinterpret_cast<LdsEntry*>(lds);
    LdsEntry* const s_tp = s_tr + A.L.br.nb;
    unsigned char* const smem = lds + A.tab_bytes;
    const PairLaunch& L = A.L;
    // Kr: all radial / rp bins (strides of the outputs); kc of them, from k0
    // on, are histogrammed by this launch (the LDS rows); the rest go to the trash
    const int Kr = L.br.nbins, Kp = L.bp.nbins;
    const int k0 = A.k0, kc = A.kc;
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;

    // histogram geometry: rows of one bin; (Kr+1) rows, the last one a trash
    // row for pairs beyond the largest edge (no branch in the pair loop)
    // Sums that need no per-lane factor (generic marks, tangent sums) share
    // FP64 rows among lanes l, l+16, l+32, l+48: an 8-byte LDS operation is
    // served 16 lanes per pass, so lanes of one column never collide within
    // a pass, and the rows take a quarter of the LDS (more waves per CU).
    // Their count rows are shared by lanes l and l+32 the same way.
    // The count rows of the count modes are shared by lanes l and l+32 (a
    // 4-byte LDS operation is served 32 lanes per pass).
    constexpr bool SHARE4 = (MODE == kModeWSum || MODE == kModeTan);
    constexpr bool SHARE2 = (MODE == kModeCount3D || MODE == kModeProj);
    const bool shared = (MODE == kModeXYZ) || (!F64 && A.shared_count);  // shared u32 counts
    // The forward-mode pass keeps a bin's W row, dW/dtheta row and count row in
    // one 384-byte block (128 B each), so the second and third atomics of a
    // pair address the first one's block at fixed offsets (immediate offsets
    // of the LDS instructions, no address arithmetic).
    constexpr bool TANI = (MODE == kModeTan);
    constexpr uint32_t kTanRow = 128u, kTanCnt = 256u;
    // The marked (rp, pi) pass never writes its trash row or column (pairs
    // outside (rp_max, pi_max) skip the atomics), so its table has Kr x Kp
    // cells only: 20 waves per CU fit in the LDS instead of 16.
    constexpr bool WXC = (MODE == kModeWXYZ);
    const uint32_t step_r =
        XY ? (uint32_t)(WXC ? Kp : Kp + 1) * cell_bytes
           : (JK ? 8u
                 : (shared ? 4u
                           : (TANI ? (ALSO ? 384u : 256u)
                                   : ((SHARE4 || SHARE2) ? 128u : (MODE == kModeObj ? 256u : 512u)))));
    const uint32_t lane_off =
        (shared || XY || JK) ? 0u
               : (SHARE4 ? (uint32_t)(lane & 15) * 8u
                         : (SHARE2 ? (uint32_t)(lane & 31) * 4u : (uint32_t)lane * (MODE == kModeObj ? 4u : 8u)));
    const uint32_t cnt_col = SHARE4 ? (uint32_t)(lane & 31) * 4u : 0u;  // SHARE4 count rows
    {
        const char* ws = static_cast<const char*>(L.ws);
        const TabEntry* gtr = reinterpret_cast<const TabEntry*>(ws + L.br.table_off);
        for (int i = threadIdx.x; i < L.br.nb; i += blockDim.x) {
            const TabEntry e = gtr[i];
            s_tr[i].edge = e.edge;
            s_tr[i].addr = e.k * step_r;  // the non-MULTI variants run with k0 = 0, kc = Kr
            s_tr[i].k = e.k;
        }
        if (XY) {
            const TabEntry* gtp = reinterpret_cast<const TabEntry*>(ws + L.bp.table_off);
            for (int i = threadIdx.x; i < L.bp.nb; i += blockDim.x) {
                const TabEntry e = gtp[i];
                s_tp[i].edge = e.edge;
                s_tp[i].addr = e.k * cell_bytes;
                s_tp[i].k = e.k;
            }
        }
        if (!kPersist<MODE>) {
            uint32_t* h = reinterpret_cast<uint32_t*>(smem);
            const int nwords = (WAVES * A.wave_bytes) / 4;
            for (int i = threadIdx.x; i < nwords; i += blockDim.x) h[i] = 0u;
        }
    }
    __syncthreads();  // no block-level synchronisation below

    unsigned char* const wsm = smem + wid * A.wave_bytes;  // this wave's histogram rows
    unsigned char* const hl = wsm + lane_off;               // ... at this lane's column

    // Work distribution.  Persistent: the grid holds as many workgroups as
    // fit on the device at once, and each wave takes lane groups one at a
    // time from eight queues, one per XCD (workgroups are dispatched to the
    // XCDs round-robin, so blockIdx % 8 is the XCD): queue x hands out the
    // x-th contiguous eighth of the (Hilbert-ordered) groups, so an XCD's L2
    // sees one region; a wave whose queue is empty moves on to the next
    // XCD's.  Every wave leaves once all eight are exhausted (one shared
    // queue measured 4-13 % slower).  Other modes launch one wave per lane
    // group; XCD x takes runs of HP_XCD_CHUNK consecutive blocks (see below),
    // which cut the L2 fetches of the 1M-point count from 1.8 GB to 0.65 GB
    // and the 64-mock batch count by 3.5 % (one contiguous eighth per XCD
    // fetched less still but ran 8 % slower on that batch: its XCDs finished
    // unevenly).
    int qn = 0;
    const int xq = (int)(blockIdx.x & 7u);
    for (;;) {
        int64_t g = -1;
        if (kPersist<MODE>) {
            while (qn < HP_QUEUES) {
                const int q = (xq + qn) % HP_QUEUES;
                const int64_t lo = L.ngroups * q / HP_QUEUES, hi = L.ngroups * (q + 1) / HP_QUEUES;
                uint32_t i = 0;
                if (lane == 0) i = __hip_atomic_fetch_add(L.queue + q, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                i = (uint32_t)uni((int)i);
                if (lo + (int64_t)i < hi) {
                    g = lo + (int64_t)i;
                    break;
                }
                qn++;
            }
            if (g < 0) break;
            // this wave's histogram rows start from zero for every group
            uint32_t* h = reinterpret_cast<uint32_t*>(wsm);
            for (int i = lane; i < A.wave_bytes / 4; i += 64) h[i] = 0u;
            __builtin_amdgcn_wave_barrier();
        } else {
            int64_t b = blockIdx.x;
#if HP_XCD_CHUNK > 0
            {
                // chunked XCD interleave: XCD x (= b % 8) runs its j-th block
                // (j = b / 8) on block (j / C) * 8C + x * C + j % C, so each
                // XCD works through runs of C consecutive blocks (a compact
                // region of the Hilbert-ordered groups) while every XCD still
                // gets runs spread over the whole range; blocks past the last
                // full 8C super-run keep their own index (a bijection)
                constexpr int64_t C = HP_XCD_CHUNK;
                const int64_t full = ((int64_t)gridDim.x / (8 * C)) * (8 * C);
                if (b < full) b = ((b >> 3) / C) * (8 * C) + (b & 7) * C + ((b >> 3) % C);
            }
#endif
            g = b * WAVES + wid;
            if (g >= L.ngroups) break;
        }
        const double* gedges_r = reinterpret_cast<const double*>(static_cast<const char*>(L.ws) + L.br.edges_off);
        const double* gedges_p = reinterpret_cast<const double*>(static_cast<const char*>(L.ws) + L.bp.edges_off);

        const int m = uni(L.group_mock[g]);
        const MockGeom& G = L.geom[m];
        const int side = SWAP;
        const int64_t gl = g - G.group_off[side];
        const int64_t p0 = G.p_off[side] + gl * kGroup;
        const int64_t left = G.n[side] - gl * kGroup;
        const int cnt = uni((int)(left < kGroup ? left : kGroup));
        // position in the group's Hilbert run of the point this lane holds
        const int lpos = spread_pos(lane, cnt);
        const bool act = lpos < cnt;
        const int64_t p = p0 + (act ? lpos : 0);
        // lane points are read in input order through the Hilbert permutation
        const int64_t q = (int64_t)L.lane.perm[p];
        // inactive lanes carry NaN: every comparison fails, so they land in the trash row
        const double qnan = __builtin_nan("");
        const double px = act ? L.lane.x[q * L.lane.xs] : qnan;
        const double py = act ? L.lane.y[q * L.lane.xs] : qnan;
        const double pz = act ? L.lane.z[q * L.lane.xs] : qnan;

        const double Lx = G.period[0], Ly = G.period[1], Lz = G.period[2];
        const int nx = G.ndiv[0], ny = G.ndiv[1], nz = G.ndiv[2];
        const int pbc = G.pbc;
        const int64_t cell_off = G.cell_off;

        // ---- bounding box of the group (f32, rounded outwards), grown by the
        // eps margin (1e-5 of the box size; the f32 sums of the window and
        // filter tests below round by ~1e-7 of it)
        const float eps = G.eps;
        const float INF = __builtin_inff();
        const float bxlo = unif(wave_min(act ? f32_down(px) : INF) - eps);
        const float bylo = unif(wave_min(act ? f32_down(py) : INF) - eps);
        const float bzlo = unif(wave_min(act ? f32_down(pz) : INF) - eps);
        const float bxhi = unif(wave_max(act ? f32_up(px) : -INF) + eps);
        const float byhi = unif(wave_max(act ? f32_up(py) : -INF) + eps);
        const float bzhi = unif(wave_max(act ? f32_up(pz) : -INF) + eps);
        const float Rx = G.search[0], Ry = G.search[1], Rz = G.search[2];
        const float cxf = (float)G.cell[0], cyf = (float)G.cell[1], czf = (float)G.cell[2];
        const float r2xy = G.r2xy;
        const int shape = uni(G.shape);

        int jx0 = (int)floorf((bxlo - Rx) / cxf), jx1 = (int)floorf((bxhi + Rx) / cxf);
        int jy0 = (int)floorf((bylo - Ry) / cyf), jy1 = (int)floorf((byhi + Ry) / cyf);
        if (pbc) {
            jx0 = max(jx0, -nx); jx1 = min(jx1, 2 * nx - 1);
            jy0 = max(jy0, -ny); jy1 = min(jy1, 2 * ny - 1);
        } else {
            jx0 = max(jx0, 0); jx1 = min(jx1, nx - 1);
            jy0 = max(jy0, 0); jy1 = min(jy1, ny - 1);
        }
        jx0 = uni(jx0); jx1 = uni(jx1); jy0 = uni(jy0); jy1 = uni(jy1);

        // lane-side weights
        double lw[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        const int nw = L.nw;
        if (F64) {
            for (int c = 0; c < nw && c < 5; c++) lw[c] = L.lane.w[q * nw + c];
        }
        // stream side: one 32-byte row per point (x, y, z, order key)
        const double* __restrict__ SR = L.strm.x;
        const double* __restrict__ SW = L.strm.w;
        const uint32_t* __restrict__ CS = L.strm.cell_start;
        const int wfunc = L.wfunc;
        const double pimax2 = (MODE == kModeProj) ? gedges_p[0] : 0.0;
        const uint32_t shr = L.br.shift, bshr = L.br.basesh, nbr1 = L.br.nb - 1;
        const uint32_t shp = L.bp.shift, bshp = L.bp.basesh, nbp1 = L.bp.nb - 1;
        const uint32_t trash_r = (uint32_t)kc * step_r;  // projected: pairs beyond pi_max
        const uint32_t trash_p = (uint32_t)Kp * cell_bytes;  // xy_z: beyond the last pi edge

        uint64_t evals = 0;
        const int shared_bins = (MODE == kModeXYZ) ? kc * Kp : kc;  // this launch's bins
        const int row_len = (MODE == kModeXYZ) ? Kp + 1 : 1;
        // per-mock count accumulator: integer adds are exact in any order, so the
        // groups add their histograms with atomics (no per-group partials)
        const int all_bins = (MODE == kModeXYZ) ? Kr * Kp : Kr;
        uint64_t* const cacc =
            L.pcount ? L.pcount + ((int64_t)m * L.crep + (int64_t)(gl % L.crep)) * all_bins +
                           (int64_t)k0 * ((MODE == kModeXYZ) ? Kp : 1)
                     : nullptr;
        uint32_t* const ch32 = reinterpret_cast<uint32_t*>(wsm + A.cnt_off);
        uint64_t bound = 0;

        // Pairs are processed in batches: the stream points of a batch arrive with
        // one scalar-load wait; for each chunk of HP_SPLIT_CHUNK slots the
        // distances are formed, then the chunk's bucket-table reads are issued
        // back to back (one LDS wait), then the compares; the fire-and-forget LDS
        // atomics of the batch follow.  (Chunks of 2 keep the count kernel within
        // its register budget; 2 instead of 1 slot per wait: -4 % kernel time.)
        auto sep = [&](double qx, double qy, double qz, double shx, double shy, double shz, int sxs, int sys, int szs,
                       double ax, double ay, double az, double& dx, double& dy, double& dz) __attribute__((always_inline)) {
            if (!SWAP) {
                dx = ax - qx;
                dy = ay - qy;
                dz = az - qz;
            } else {
                if (sxs) qx = qx - shx;
                if (sys) qy = qy - shy;
                if (szs) qz = qz - shz;
                dx = qx - px;
                dy = qy - py;
                dz = qz - pz;
            }
        };
        // histogram byte offset (within this wave's rows, before the lane offset)
        auto addr_of = [&](double dx, double dy, double dz) __attribute__((always_inline)) -> uint32_t {
            if (MODE == kModeProj) {
                const double dxy = dx * dx + dy * dy;
                const double dz2 = dz * dz;
                const uint32_t a = bin_addr<MULTI>(dxy, s_tr, shr, bshr, nbr1, step_r, gedges_r, k0, kc);
                return (dz2 <= pimax2) ? a : trash_r;
            } else if (XY) {
                // pairs outside (rp_max, pi_max) are marked trash_r and skip the
                // atomics: in one shared per-wave histogram they would all pile
                // onto the same trash cells
                const double dxy = dx * dx + dy * dy;
                const double dz2 = dz * dz;
                const uint32_t ar = bin_addr<MULTI>(dxy, s_tr, shr, bshr, nbr1, step_r, gedges_r, k0, kc);
                const uint32_t ap = bin_addr<MULTI>(dz2, s_tp, shp, bshp, nbp1, cell_bytes, gedges_p, 0u, (uint32_t)Kp);
                return (ar < trash_r && ap < trash_p) ? ar + ap : trash_r;
            } else {
                const double dsq = dx * dx + dy * dy + dz * dz;
                return bin_addr<MULTI>(dsq, s_tr, shr, bshr, nbr1, step_r, gedges_r, k0, kc);
            }
        };
        // stream points of this group's mock, addressed by 32-bit offsets from
        // the mock's first point (the ring holds mock-relative indices)
        const int64_t sbase = G.p_off[SWAP ? 0 : 1];
        const double* const SRm = SR + sbase * 4;
        const double* const SWm = F64 ? SW + sbase * nw : nullptr;
        // Unweighted auto-count: an unshifted pair has dx_ji = -dx_ij exactly
        // (IEEE subtraction is antisymmetric), so both ordered pairs bin alike.
        // Unshifted stream points are then taken only when they follow the whole
        // group in the lane (Hilbert) order and count twice; the group's own
        // points are paired separately below; shifted images (whose two orders
        // may round apart) are evaluated from both sides, once each.
        const uint64_t* __restrict__ HK = L.hkey;
        const bool sym = (!F64 || MODE == kModeTan || JK) && !SWAP && HK != nullptr;

        // ---- jackknife tables.  Row 0 sums w1 w2 of every pair; the row of tag t
        // sums what removing subsample t takes away from it: w1 w2 when both
        // points carry t, half of it for each point that does when their tags
        // differ.  Subsample s of the result is then row 0 minus the row of s.
        // Tags get an LDS row in the order the wave meets them (the lane points'
        // tags first); once the rows run out, a tag's share goes straight to the
        // mock's table in HBM.
        const int nbj = XY ? Kr * Kp : Kr;
        int stag[kJackSlots];  // tag of LDS row r + 1 (wave-uniform)
        int ns = 0;            // rows in use
        int lslot = 0;         // this lane's row (0: none)
        int ltag = 0;
        if (JK) {
#pragma unroll
            for (int j = 0; j < kJackSlots; j++) stag[j] = 0;
            ltag = (int)lw[1];
            uint64_t pend = (uint64_t)__ballot(act && ltag >= 1 && ltag <= L.nsamp);
            while (pend) {
                const int t = __builtin_amdgcn_readlane(ltag, __builtin_ctzll(pend));
                const uint64_t mm = (uint64_t)__ballot(act && ltag == t);
                if (ns < L.jslots) {
#pragma unroll
                    for (int j = 0; j < kJackSlots; j++)
                        if (j == ns) stag[j] = t;
                    ns++;
                    if (act && ltag == t) lslot = ns;
                }
                pend &= ~mm;
            }
        }
        int jlast_t = -1, jlast_s = 0;
        // row of a (wave-uniform) tag, assigning a free row on first sight
        auto slot_of = [&](int t) __attribute__((always_inline)) -> int {
            if (t == jlast_t) return jlast_s;
            int sl = 0;
#pragma unroll
            for (int j = 0; j < kJackSlots; j++)
                if (j < ns && stag[j] == t) sl = j + 1;
            if (sl == 0 && ns < L.jslots && t >= 1 && t <= L.nsamp) {
#pragma unroll
                for (int j = 0; j < kJackSlots; j++)
                    if (j == ns) stag[j] = t;
                ns++;
                sl = ns;
            }
            jlast_t = t;
            jlast_s = sl;
            return sl;
        };
        auto jk_global = [&](int tag, uint32_t a, double val) __attribute__((always_inline)) {
            if (tag < 1 || tag > L.nsamp) return;
            const uint32_t c = a >> 3;
            const int64_t b = XY ? (int64_t)(k0 + c / (uint32_t)(Kp + 1)) * Kp + (c % (uint32_t)(Kp + 1))
                                 : (int64_t)(k0 + c);
            __hip_atomic_fetch_add(L.jacc + ((int64_t)m * (L.nsamp + 1) + tag) * nbj + b, val, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        };
        // one pair in cell a: v = w1 w2 (times 2 for an unshifted pair under symmetry)
        // A group whose points all carry one tag t (the common case for spatial
        // subsamples; 3-D bins only) keeps the all-pairs row in 16 lane columns,
        // where no two lanes of one LDS pass collide, and records only the shares
        // its cross-tag pairs give their partners' tags (half of each, in their
        // rows; what goes straight to HBM is also summed in row 1, t's own row,
        // unused otherwise): the loss of t is then the all-pairs row minus them.
        bool jut = false;
        if (JK && !XY) jut = (ns == 1) && ((uint64_t)__ballot(act && lslot != 1) == 0);
        const uint32_t r0col = (uint32_t)(lane & 15) * 8u;
        auto jk_add = [&](uint32_t a, double v, int ts, int ss) __attribute__((always_inline)) {
            if (!XY && jut) {
                __hip_atomic_fetch_add(reinterpret_cast<double*>(wsm + A.jr0_off + a * 16u + r0col), v,
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (ltag != ts) {
                    const double sv = 0.5 * v;
                    if (ss > 0) {
                        __hip_atomic_fetch_add(reinterpret_cast<double*>(wsm + ss * A.jrow + a), sv, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                    } else {
                        jk_global(ts, a, sv);
                        __hip_atomic_fetch_add(reinterpret_cast<double*>(wsm + A.jrow + a), sv, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                    }
                }
                return;
            }
            // all-pairs row: in 3-D its 16 lane columns (no two lanes of one LDS
            // pass on one address), summed at the flush; (rp, pi) keeps one row
            if (!XY)
                __hip_atomic_fetch_add(reinterpret_cast<double*>(wsm + A.jr0_off + a * 16u + r0col), v,
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            else
                __hip_atomic_fetch_add(reinterpret_cast<double*>(wsm + a), v, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
            const bool same_t = (ltag == ts);
            const double lv = same_t ? v : 0.5 * v;
            if (lslot > 0)
                __hip_atomic_fetch_add(reinterpret_cast<double*>(wsm + lslot * A.jrow + a), lv, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
            else
                jk_global(ltag, a, lv);
            if (!same_t) {
                const double sv = 0.5 * v;
                if (ss > 0)
                    __hip_atomic_fetch_add(reinterpret_cast<double*>(wsm + ss * A.jrow + a), sv, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                else
                    jk_global(ts, a, sv);
            }
        };
        // symmetric auto-counts: a stream point follows the group in the Hilbert
        // order iff its order key, (Hilbert key << 32) | input index, is at
        // least that of the point at Hilbert position p0 + cnt (none after the
        // mock's last group)
        uint64_t hend = ~0ull;
        if (sym && p0 + cnt < G.p_off[0] + G.n[0]) {
            const uint32_t pn = sload_u32_at(L.lane.perm, p0 + cnt);
            const uint32_t hk = sload_u32_at(reinterpret_cast<const uint32_t*>(L.hrows),
                                             (int64_t)pn * (2 * L.hrow_stride) + 6);
            hend = ((uint64_t)hk << 32) | (uint64_t)pn;
        }
        // r: mock-relative stream index
        // inc: count-mode increment (2 for an unshifted pair evaluated once for both orders)
        auto accumulate = [&](uint32_t a, uint32_t r, uint32_t inc, bool) __attribute__((always_inline)) {
            if (!F64) {
                __hip_atomic_fetch_add(reinterpret_cast<uint32_t*>(hl + a), inc, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
            } else {
                double v;
                if (MODE == kModeJac) {
                    v = sload_f64(SWm, r * 8u);
                } else {
                    double sw[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
                    for (int c = 0; c < 5; c++)
                        if (c < nw) sw[c] = sload_f64(SWm, (r * (uint32_t)nw + (uint32_t)c) * 8u);
                    v = wfunc_eval(wfunc, lw, sw);
                }
                __hip_atomic_fetch_add(reinterpret_cast<double*>(hl + a), v, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                if (ALSO) {
                    const uint32_t ca = SHARE4 ? a + cnt_col : ((a + lane_off) >> 1);
                    __hip_atomic_fetch_add(reinterpret_cast<uint32_t*>(wsm + A.cnt_off + ca), 1u, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            }
        };

        // ---- survivor ring: stream indices that passed the point filter, all
        // with the same periodic image (the ring is drained when the image changes)
        constexpr int B = kBatch;
        uint32_t* const ring = reinterpret_cast<uint32_t*>(wsm + A.ring_off);
        int qlen = 0;
        int rimg = -1;
        uint32_t rinc = 1u;  // count increment of the ring's entries
        bool rcorr = false;  // Jacobian auto-count in a shifted image: also the reverse orientation
        bool rhalf = false;  // tangent mode under symmetry, shifted image
        // tangent mode: the lane's (w, t), halved for a shifted image under symmetry
        double rl0 = lw[0], rl1 = lw[1];
        double rshx = 0.0, rshy = 0.0, rshz = 0.0;
        int rsxs = 0, rsys = 0, rszs = 0;
        double rax = px, ray = py, raz = pz;
        // evaluate ring[q0, q0 + nv), nv <= B, against the 64 lane points: the
        // indices arrive with one LDS read, the points with scalar loads at
        // 32-bit offsets (one shift serves x, y and z)
        auto consume = [&](int q0, int nv) __attribute__((always_inline)) {
            const int v = (int)ring[q0 + min(lane, nv - 1)];
            uint32_t off[B];
#pragma unroll
            for (int u = 0; u < B; u++) off[u] = (uint32_t)__builtin_amdgcn_readlane(v, u) * 8u;
            double bx[B], by[B], bz[B];
#pragma unroll
            for (int u = 0; u < B; u++) {
                // the point's 32-byte row: off * 4 bytes from the mock's first row
                bx[u] = sload_f64(SRm, off[u] * 4u);
                by[u] = sload_f64(SRm, off[u] * 4u + 8u);
                bz[u] = sload_f64(SRm, off[u] * 4u + 16u);
            }
            // the Jacobian and product-mark passes issue the batch's weights
            // with its coordinates (one scalar-load wait for both); the
            // tangent and jackknife passes load their two words per point
            // after the binning, where they hold fewer live scalar registers
            // (the tangent pass ran 2 % slower with them issued here)
            double bw[B], bt[B];
            if (MODE == kModeJac || PROD) {
#pragma unroll
                for (int u = 0; u < B; u++) bw[u] = sload_f64(SWm, off[u]);
            }
            evals += (uint64_t)cnt * (uint64_t)nv;
            uint32_t a[B];
            if (HP_SPLIT && (!XY || HP_SPLIT_XY) && !MULTI) {
                // per chunk of C slots: the distances first, then all of their
                // bucket reads, then the compares -- one LDS wait per chunk
                // instead of one per slot (C bounds the live registers)
                constexpr int C = (HP_SPLIT_CHUNK < B) ? HP_SPLIT_CHUNK : B;
#pragma unroll
                for (int h = 0; h < B; h += C) {
                    double dq[C], dz2[C];
#pragma unroll
                    for (int c = 0; c < C; c++) {
                        const int u = h + c;
                        double dx, dy, dz;
                        sep(bx[u], by[u], bz[u], rshx, rshy, rshz, rsxs, rsys, rszs, rax, ray, raz, dx, dy, dz);
                        if (MODE == kModeProj || XY) {
                            dq[c] = dx * dx + dy * dy;
                            dz2[c] = dz * dz;
                        } else {
                            dq[c] = dx * dx + dy * dy + dz * dz;
                        }
                    }
                    uint4 ent[C], enp[C];
#pragma unroll
                    for (int c = 0; c < C; c++) {
                        const uint32_t hi = (uint32_t)__double2hiint(dq[c]);
                        const uint32_t b = min(__builtin_elementwise_sub_sat(hi, bshr) >> shr, nbr1);
                        ent[c] = *reinterpret_cast<const uint4*>(s_tr + b);
                        if (XY) {
                            const uint32_t hp = (uint32_t)__double2hiint(dz2[c]);
                            const uint32_t bp = min(__builtin_elementwise_sub_sat(hp, bshp) >> shp, nbp1);
                            enp[c] = *reinterpret_cast<const uint4*>(s_tp + bp);
                        }
                    }
#pragma unroll
                    for (int c = 0; c < C; c++) {
                        const int u = h + c;
                        const double edge = __hiloint2double((int)ent[c].y, (int)ent[c].x);
                        uint32_t v = ent[c].z + (edge < dq[c] ? step_r : 0u);
                        if (MODE == kModeProj) v = (dz2[c] <= pimax2) ? v : trash_r;
                        if (XY) {
                            // outside (rp_max, pi_max): trash_r, which skips the atomics
                            const double ep = __hiloint2double((int)enp[c].y, (int)enp[c].x);
                            const uint32_t ap = enp[c].z + (ep < dz2[c] ? cell_bytes : 0u);
                            v = (v < trash_r && ap < trash_p) ? v + ap : trash_r;
                        }
                        a[u] = (u < nv) ? v : trash_r;
                    }
                }
            } else if (nv == B) {
#pragma unroll
                for (int u = 0; u < B; u++) {
                    double dx, dy, dz;
                    sep(bx[u], by[u], bz[u], rshx, rshy, rshz, rsxs, rsys, rszs, rax, ray, raz, dx, dy, dz);
                    a[u] = addr_of(dx, dy, dz);
                }
            } else {
                // short batch: the repeated slots go to the trash row
#pragma unroll
                for (int u = 0; u < B; u++) {
                    double dx, dy, dz;
                    sep(bx[u], by[u], bz[u], rshx, rshy, rshz, rsxs, rsys, rszs, rax, ray, raz, dx, dy, dz);
                    a[u] = (u < nv) ? addr_of(dx, dy, dz) : trash_r;
                }
            }
            if (MODE == kModeJac && !SWAP && rcorr) {
                // Auto-count with one set of weights: the sample-2 rows need the
                // swapped orientation of each pair, (x2 + x2shift) - x1, which
                // across the periodic seam may round differently from the forward
                // one; record the rare disagreements so rows2 = rows1 + corr (the
                // swapped pass's rows, exactly for sums of small integers and to
                // the rounding of the row sums otherwise: corr is added with
                // atomics from many waves, in no fixed order).
                const uint32_t* __restrict__ SP = L.strm.perm;
#pragma unroll
                for (int u = 0; u < B; u++) {
                    if (u >= nv) break;
                    const double rx = rsxs ? (bx[u] + rshx) - px : px - bx[u];
                    const double ry = rsys ? (by[u] + rshy) - py : py - by[u];
                    const double rz = rszs ? (bz[u] + rshz) - pz : pz - bz[u];
                    const uint32_t ar = addr_of(rx, ry, rz);
                    if (ar != a[u]) {
                        const uint32_t sf = a[u] / step_r, sr = ar / step_r;
                        double* crow = L.corr + (int64_t)sload_u32_at(SP, sbase + (off[u] >> 3)) * Kr + k0;
                        if (sf < (uint32_t)kc)
                            __hip_atomic_fetch_add(crow + sf, lw[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        if (sr < (uint32_t)kc)
                            __hip_atomic_fetch_add(crow + sr, -lw[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                }
            }
            if (MODE == kModeTan || MODE == kModeJac || JK) {
                if (MODE == kModeTan || JK) {
#pragma unroll
                    for (int u = 0; u < B; u++) {
                        bw[u] = sload_f64(SWm, off[u] * 2u);
                        bt[u] = sload_f64(SWm, off[u] * 2u + 8u);
                    }
                }
                int jt[B], js[B];
                if (JK) {
                    // tags and rows of the batch's stream points (wave-uniform;
                    // a row is only taken when some lane pairs with the point)
#pragma unroll
                    for (int u = 0; u < B; u++) {
                        jt[u] = uni((int)bt[u]);
                        js[u] = ((uint64_t)__ballot(a[u] != trash_r) != 0) ? slot_of(jt[u]) : 0;
                    }
                }
#pragma unroll
                for (int u = 0; u < B; u++) {
                    // the tangent pass's three atomics are worth a branch; the
                    // Jacobian's one is not (it adds into the trash row instead)
                    if (a[u] == trash_r && (JK || MODE == kModeTan)) continue;
                    if (JK) {
                        jk_add(a[u], rl0 * bw[u], jt[u], js[u]);
                    } else if (MODE == kModeTan) {
                        // pair terms w_i w_j and t_i w_j + w_i t_j (t = dw/dtheta)
                        const double v = rl0 * bw[u];
                        // one fused multiply-add (a tangent term, not a distance:
                        // within the sums' 1e-12 tolerance, deterministic)
                        const double tv = __builtin_fma(rl0, bt[u], rl1 * bw[u]);
                        __hip_atomic_fetch_add(reinterpret_cast<double*>(hl + a[u]), v, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                        __hip_atomic_fetch_add(reinterpret_cast<double*>(hl + kTanRow + a[u]), tv,
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        if (ALSO)
                            __hip_atomic_fetch_add(
                                reinterpret_cast<uint32_t*>(wsm + kTanCnt + a[u] + cnt_col), rinc,
                                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    } else {
                        __hip_atomic_fetch_add(reinterpret_cast<double*>(hl + a[u]), bw[u], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                        if (ALSO)
                            __hip_atomic_fetch_add(reinterpret_cast<uint32_t*>(wsm + A.cnt_off + ((a[u] + lane_off) >> 1)),
                                                   1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    }
                }
            } else if (PROD) {
                // product marks: the batch's stream weights came with its coordinates
#pragma unroll
                for (int u = 0; u < B; u++) {
                    if (XY && a[u] == trash_r) continue;
                    __hip_atomic_fetch_add(reinterpret_cast<double*>(hl + a[u]), lw[0] * bw[u], __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                    if (ALSO)
                        __hip_atomic_fetch_add(reinterpret_cast<uint32_t*>(wsm + A.cnt_off + a[u] + cnt_col), 1u,
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            } else {
#pragma unroll
                for (int u = 0; u < B; u++)
                    if (!XY || a[u] != trash_r) accumulate(a[u], off[u] >> 3, rinc, false);
            }
        };
        auto drain = [&]() __attribute__((always_inline)) {
            for (int q = 0; q < qlen; q += B) consume(q, min(B, qlen - q));
            qlen = 0;
        };
        // Shared (per-wave) uint32 histograms: each stream point adds at most
        // 2 (an unshifted pair counted for both orders) for each of the 64
        // lanes.  Before the points of another chunk could make a bin wrap,
        // the histogram is flushed into the group's uint64 accumulator; the
        // survivors still waiting in the ring land in the emptied histogram.
        auto guard_shared = [&](uint64_t add) __attribute__((always_inline)) {
            bound += add;
            if (bound >= A.flush_at) {
                for (int s = lane; s < shared_bins; s += 64) {
                    const int src = (MODE == kModeXYZ) ? (s / Kp) * row_len + s % Kp : s;
                    if (ch32[src])
                        __hip_atomic_fetch_add(cacc + s, (uint64_t)ch32[src], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
                }
                for (int s = lane; s < (kc + 1) * row_len; s += 64) ch32[s] = 0u;
                __builtin_amdgcn_wave_barrier();
                bound = add + (uint64_t)qlen * 128u;
            }
        };

        for (int jx = jx0; jx <= jx1; jx++) {
            const float gx0 = fmaxf(0.f, fmaxf((float)jx * cxf - bxhi, bxlo - (float)(jx + 1) * cxf));
            const int tx = jx < 0 ? -1 : (jx >= nx ? 1 : 0);
            const int wx = jx - tx * nx;
            for (int jy = jy0; jy <= jy1; jy++) {
                const float gy0 = fmaxf(0.f, fmaxf((float)jy * cyf - byhi, bylo - (float)(jy + 1) * cyf));
                const float d2 = gx0 * gx0 + gy0 * gy0;
                if (d2 > r2xy) continue;
                const float dzr = (shape == 0) ? sqrtf(fmaxf(r2xy - d2, 0.f)) + eps : Rz;
                int jz0 = (int)floorf((bzlo - dzr) / czf), jz1 = (int)floorf((bzhi + dzr) / czf);
                if (pbc) { jz0 = max(jz0, -nz); jz1 = min(jz1, 2 * nz - 1); }
                else { jz0 = max(jz0, 0); jz1 = min(jz1, nz - 1); }
                jz0 = uni(jz0); jz1 = uni(jz1);
                if (jz0 > jz1) continue;
                const int ty = jy < 0 ? -1 : (jy >= ny ? 1 : 0);
                const int wy = jy - ty * ny;
                const int64_t colbase = cell_off + ((int64_t)wx * ny + wy) * nz;
                // up to three periodic pieces of the z-run
                for (int piece = -1; piece <= 1; piece++) {
                    int u0, u1;
                    if (piece == -1) { u0 = jz0; u1 = min(jz1, -1); }
                    else if (piece == 0) { u0 = max(jz0, 0); u1 = min(jz1, nz - 1); }
                    else { u0 = max(jz0, nz); u1 = jz1; }
                    if (u0 > u1) continue;
                    const int tz = piece;
                    const int w0 = u0 - tz * nz, w1 = u1 - tz * nz;
                    const int jb = uni((int)CS[colbase + w0]);
                    const int je = uni((int)CS[colbase + w1 + 1]);
                    if (jb >= je) continue;
                    // x2shift of the reference: the shift that brings the sample-2
                    // point next to the sample-1 point.  Normal orientation visits
                    // sample-2 images (x2shift = t*L); swapped orientation visits
                    // sample-1 images (x2shift = -t*L).
                    const int sxs = SWAP ? -tx : tx, sys = SWAP ? -ty : ty, szs = SWAP ? -tz : tz;
                    const double shx = sxs < 0 ? -Lx : (sxs > 0 ? Lx : 0.0);
                    const double shy = sys < 0 ? -Ly : (sys > 0 ? Ly : 0.0);
                    const double shz = szs < 0 ? -Lz : (szs > 0 ? Lz : 0.0);
                    double ax = px, ay = py, az = pz;  // normal: x1tmp of the lane point
                    if (!SWAP) {
                        if (sxs) ax = px - shx;
                        if (sys) ay = py - shy;
                        if (szs) az = pz - shz;
                    }
                    const int j = jb;
                    // Point filter: 64 stream points at a time (one per lane, one
                    // vector load) are tested against the group's box grown by the
                    // search length; survivors go to the ring and are evaluated
                    // against all 64 lane points in batches of B.  The image of
                    // stream point q visited here sits at q + t*L for either
                    // orientation.  f32 with the box rounded outwards and grown by
                    // the eps margin keeps every in-range pair.
                    const int img = (sxs + 1) * 9 + (sys + 1) * 3 + (szs + 1);
                    if (img != rimg) {
                        drain();
                        rimg = img;
                        rshx = shx; rshy = shy; rshz = shz;
                        rsxs = sxs; rsys = sys; rszs = szs;
                        rax = ax; ray = ay; raz = az;
                        rinc = (sym && img == 13) ? 2u : 1u;
                        rcorr = MODE == kModeJac && !SWAP && L.corr != nullptr && img != 13;
                        rhalf = MODE == kModeTan && sym && img != 13;
                        if (JK) rl0 = (sym && img == 13) ? 2.0 * lw[0] : lw[0];
                        if (MODE == kModeTan) {
                            rl0 = rhalf ? 0.5 * lw[0] : lw[0];
                            rl1 = rhalf ? 0.5 * lw[1] : lw[1];
                        }
                    }
                    const bool after_group = sym && img == 13;
                    const float ftx = (float)(tx * Lx), fty = (float)(ty * Ly), ftz = (float)(tz * Lz);
                    for (int c = j; c < je; c += 64) {
                        if (shared) guard_shared((uint64_t)cnt * 128u);
                        const int jj = c + lane;
                        bool ok = false;
                        if (jj < je) {
                            // the stream point's row in one 32-byte load
                            const double4 sr = reinterpret_cast<const double4*>(SR)[jj];
                            const float fx = (float)sr.x + ftx;
                            const float fy = (float)sr.y + fty;
                            const float fz = (float)sr.z + ftz;
                            const float gx = fmaxf(0.f, fmaxf(fx - bxhi, bxlo - fx));
                            const float gy = fmaxf(0.f, fmaxf(fy - byhi, bylo - fy));
                            const float gz = fmaxf(0.f, fmaxf(fz - bzhi, bzlo - fz));
                            const float gxy = gx * gx + gy * gy;
                            ok = (shape == 0) ? (gxy + gz * gz <= r2xy) : (gxy <= r2xy && gz <= Rz);
                            if (after_group) ok = ok && (uint64_t)__double_as_longlong(sr.w) >= hend;
                        }
                        const uint64_t m = (uint64_t)__ballot(ok);
                        if (ok) {
                            const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                            ring[qlen + (int)rank] = (uint32_t)(jj - sbase);
                        }
                        // lanes read each other's ring slots: keep LDS accesses in order
                        __builtin_amdgcn_wave_barrier();
                        qlen += __builtin_popcountll(m);
                        int qh = 0;
                        for (; qlen - qh >= B; qh += B) consume(qh, B);
                        if (qh > 0) {
                            // move the < B leftovers to the front of the ring
                            const int r = qlen - qh;
                            if (r > 0) {
                                const uint32_t t = ring[qh + min(lane, r - 1)];
                                if (lane < r) ring[lane] = t;
                                __builtin_amdgcn_wave_barrier();
                            }
                            qlen = r;
                        }
                    }
                }
            }
        }
        drain();

        if (sym) {
            if (shared) guard_shared((uint64_t)cnt * 128u);
            // the group's own points as stream points, unshifted: lane i with
            // point k counts 2 when k follows i in the lane order, 1 for k == i
            // (the self pair), 0 when k precedes i (counted from the other side)
            for (int kb = 0; kb < cnt; kb += B) {
                double bx[B], by[B], bz[B], bw[B], bt[B];
#pragma unroll
                for (int u = 0; u < B; u++) {
                    const int64_t qk =
                        (int64_t)sload_u32_at(L.lane.perm, p0 + spread_pos(min(kb + u, cnt - 1), cnt));
                    bx[u] = sload_f64_at(L.lane.x, qk * L.lane.xs);
                    by[u] = sload_f64_at(L.lane.y, qk * L.lane.xs);
                    bz[u] = sload_f64_at(L.lane.z, qk * L.lane.xs);
                    if (MODE == kModeTan || JK) {
                        bw[u] = sload_f64_at(L.lane.w, 2 * qk);
                        bt[u] = sload_f64_at(L.lane.w, 2 * qk + 1);
                    }
                }
                evals += (uint64_t)cnt * (uint64_t)min(B, cnt - kb);
                uint32_t a[B];
#pragma unroll
                for (int u = 0; u < B; u++) {
                    double dx, dy, dz;
                    sep(bx[u], by[u], bz[u], 0.0, 0.0, 0.0, 0, 0, 0, px, py, pz, dx, dy, dz);
                    const int k = kb + u;
                    a[u] = (k < cnt && k >= lane) ? addr_of(dx, dy, dz) : trash_r;
                }
                if (JK) {
#pragma unroll
                    for (int u = 0; u < B; u++) {
                        const int ts = uni((int)bt[u]);
                        const int ss = ((uint64_t)__ballot(a[u] != trash_r) != 0) ? slot_of(ts) : 0;
                        if (a[u] != trash_r) jk_add(a[u], (((kb + u > lane) ? 2.0 : 1.0) * lw[0]) * bw[u], ts, ss);
                    }
                    continue;
                }
#pragma unroll
                for (int u = 0; u < B; u++) {
                    const uint32_t inc = (kb + u > lane) ? 2u : 1u;
                    if (MODE == kModeTan && a[u] != trash_r) {
                        // rows are doubled at the end: the self pair adds half
                        const double h = (kb + u > lane) ? 1.0 : 0.5;
                        const double wj = bw[u] * h, tj = bt[u] * h;
                        const double v = lw[0] * wj;
                        const double tv = __builtin_fma(lw[0], tj, lw[1] * wj);
                        __hip_atomic_fetch_add(reinterpret_cast<double*>(hl + a[u]), v, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                        __hip_atomic_fetch_add(reinterpret_cast<double*>(hl + kTanRow + a[u]), tv,
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        if (ALSO) {
                            __hip_atomic_fetch_add(
                                reinterpret_cast<uint32_t*>(wsm + kTanCnt + a[u] + cnt_col), inc,
                                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        }
                    } else if (MODE != kModeTan && (!XY || a[u] != trash_r)) {
                        accumulate(a[u], 0u, inc, false);
                    }
                }
            }
        }

        // ---- group epilogue (the trash row kc is dropped); LDS row s is bin k0 + s
        if (shared) {
            for (int s = lane; s < shared_bins; s += 64) {
                const int src = (MODE == kModeXYZ) ? (s / Kp) * row_len + s % Kp : s;
                if (ch32[src])
                    __hip_atomic_fetch_add(cacc + s, (uint64_t)ch32[src], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        } else if (MODE == kModeCount3D || MODE == kModeProj || ALSO) {
            for (int s = 0; s < kc; s++) {
                const uint32_t c = TANI ? (lane < 32 ? ch32[s * 96 + lane] : 0u)  // 384-byte blocks
                                        : ((SHARE4 || SHARE2) ? (lane < 32 ? ch32[s * 32 + lane] : 0u)
                                                              : ch32[s * 64 + lane]);
                const uint64_t t = wave_sum_u64((uint64_t)c);
                if (lane == s % 64 && t) __hip_atomic_fetch_add(cacc + s, t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        const double* fh = reinterpret_cast<const double*>(wsm);
        if (MODE == kModeWXYZ && L.psum) {
            // the group's (rp, pi) cells without the trash row and column
            const int nc = kc * Kp;
            double* dst = L.psum + g * (int64_t)Kr * Kp + (int64_t)k0 * Kp;
            for (int s = lane; s < nc; s += 64) dst[s] = WXC ? fh[s] : fh[(s / Kp) * (Kp + 1) + s % Kp];
        }
        if (MODE == kModeWSum) {
            if (L.psum) {
                for (int s = 0; s < kc; s++) {
                    const double t = wave_sum_f64(lane < 16 ? fh[s * 16 + lane] : 0.0);
                    if (lane == 0) L.psum[g * Kr + k0 + s] = t;
                }
            }
        }
        if (MODE == kModeObj && act) {
            // the lane's own counts, cumulative over bins, at its input position
            // cumulative across launches: an earlier launch wrote bins < k0
            int64_t* row = L.orows + q * Kr;
            int64_t run = k0 > 0 ? row[k0 - 1] : 0;
            for (int s = 0; s < kc; s++) {
                run += (int64_t)ch32[s * 64 + lane];
                row[k0 + s] = run;
            }
        }
        if (JK && !XY && jut) {
            // uniform-tag group: row 0 = the 16 lane columns; t's loss = row 0
            // minus the shares recorded in rows 1..ns; partners' rows as they are
            double* dst = L.jacc + (int64_t)m * (L.nsamp + 1) * nbj + k0;
            for (int k = lane; k < kc; k += 64) {
                const double* r0 = reinterpret_cast<const double*>(wsm + A.jr0_off + k * 128);
                double tot = 0.0;
                for (int c = 0; c < 16; c++) tot += r0[c];
                double xs = 0.0;
                for (int r = 1; r <= ns; r++) xs += reinterpret_cast<const double*>(wsm + r * A.jrow)[k];
                if (tot != 0.0) __hip_atomic_fetch_add(dst + k, tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (tot - xs != 0.0)
                    __hip_atomic_fetch_add(dst + (int64_t)ltag * nbj + k, tot - xs, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                for (int r = 2; r <= ns; r++) {
                    int tag = 0;
#pragma unroll
                    for (int j = 0; j < kJackSlots; j++)
                        if (j == r - 1) tag = stag[j];
                    const double val = reinterpret_cast<const double*>(wsm + r * A.jrow)[k];
                    if (val != 0.0)
                        __hip_atomic_fetch_add(dst + (int64_t)tag * nbj + k, val, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        } else if (JK) {
            // the wave's rows into the mock's table (row 0, then each tag's row)
            const int cells = XY ? (kc + 1) * (Kp + 1) : kc + 1;
            for (int r = 0; r <= ns; r++) {
                int tag = 0;
#pragma unroll
                for (int j = 0; j < kJackSlots; j++)
                    if (j == r - 1) tag = stag[j];
                const double* row = reinterpret_cast<const double*>(wsm + r * A.jrow);
                double* dst = L.jacc + ((int64_t)m * (L.nsamp + 1) + tag) * nbj;
                for (int c = lane; c < cells; c += 64) {
                    double val = row[c];
                    if (!XY && r == 0) {
                        // 3-D all-pairs row: the sum of its 16 lane columns
                        const double* r0 = reinterpret_cast<const double*>(wsm + A.jr0_off + c * 128);
                        val = 0.0;
                        for (int k = 0; k < 16; k++) val += r0[k];
                    }
                    if (val == 0.0) continue;
                    int64_t b = c;
                    if (XY) {
                        const int kr = c / (Kp + 1), kp = c % (Kp + 1);
                        if (kr >= kc || kp >= Kp) continue;
                        b = (int64_t)(k0 + kr) * Kp + kp;
                    } else if (c >= kc) {
                        continue;
                    } else {
                        b = k0 + c;
                    }
                    __hip_atomic_fetch_add(dst + b, val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
        if (MODE == kModeTan) {
            // rows of w_i w_j and of t_i w_j + w_i t_j, doubled under symmetry
            const double* th = reinterpret_cast<const double*>(wsm + kTanRow);
            const double sc = sym ? 2.0 : 1.0;
            for (int s = 0; s < kc; s++) {
                const int rs = ALSO ? 48 : 32;  // doubles from one bin's W row to the next
                const double ws = wave_sum_f64(lane < 16 ? fh[s * rs + lane] : 0.0);
                const double ts = wave_sum_f64(lane < 16 ? th[s * rs + lane] : 0.0);
                if (lane == 0) {
                    if (L.psum) L.psum[g * Kr + k0 + s] = sc * ws;
                    L.ptan[g * Kr + k0 + s] = sc * ts;
                }
            }
        }
        if (MODE == kModeJac) {
            if (L.rows && act) {
                double* row = L.rows + q * Kr + k0;
                for (int s = 0; s < kc; s++) row[s] = fh[s * 64 + lane];
            }
            if (L.psum && !SWAP) {
                for (int s = 0; s < kc; s++) {
                    const double t = wave_sum_f64(act ? lw[0] * fh[s * 64 + lane] : 0.0);
                    if (lane == 0) L.psum[g * Kr + k0 + s] = t;
                }
            }
        }
        if (L.pevals && lane == 0) L.pevals[g] = evals;
        if (!kPersist<MODE>) break;
    }  // work loop
}

template <int MODE, int VAR, int ALSO, int MULTI, int WAVES>
static int launch_t(const PairArgs& A, hipStream_t s) {
    int64_t nblocks = (A.L.ngroups + WAVES - 1) / WAVES;
    if (nblocks <= 0) return 0;
    if (kPersist<MODE>) {
        // as many workgroups as the device holds at once (registers, LDS, waves)
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pairs<MODE, VAR, ALSO, MULTI, WAVES>, WAVES * 64,
                                                         (size_t)A.smem) != hipSuccess || per_cu < 1)
            per_cu = 1;
        nblocks = std::min<int64_t>(nblocks, (int64_t)A.grid * per_cu);
        if (!A.L.queue || launch_fill_u32(A.L.queue, 0u, 8, s)) return -1;  // fresh work queues
    }
    hipLaunchKernelGGL((k_pairs<MODE, VAR, ALSO, MULTI, WAVES>), dim3((unsigned)nblocks), dim3(WAVES * 64),
                       (size_t)A.smem, s, A);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int MODE, int VAR, int ALSO, int MULTI>
static int launch_w(const PairArgs& A, int waves, hipStream_t s) {
    if (waves == 8) return launch_t<MODE, VAR, ALSO, MULTI, 8>(A, s);
    if (waves == 4) return launch_t<MODE, VAR, ALSO, MULTI, 4>(A, s);
    if (waves == 2) return launch_t<MODE, VAR, ALSO, MULTI, 2>(A, s);
    return launch_t<MODE, VAR, ALSO, MULTI, 1>(A, s);
}

template <int MODE, int VAR, int ALSO>
static int launch_m(const PairArgs& A, int waves, int multi, hipStream_t s) {
    if (multi) return launch_w<MODE, VAR, ALSO, 1>(A, waves, s);
    return launch_w<MODE, VAR, ALSO, 0>(A, waves, s);
}

// One launch of mode MODE with the workgroup shape and variant flags chosen
// by launch_pairs (pairs.hip); `var` is the kernel's VAR (swapped Jacobian
// pass, or compiled-in product marks).  Each mode is instantiated in its own
// translation unit (pairs_mode_*.hip) so the variants compile in parallel.
template <int MODE>
int launch_mode(const PairArgs& A, int waves, int multi, int also, int var, hipStream_t s) {
    if constexpr (MODE == kModeJac || MODE == kModeWSum) {
        if (var)
            return also ? launch_m<MODE, 1, 1>(A, waves, multi, s) : launch_m<MODE, 1, 0>(A, waves, multi, s);
        return also ? launch_m<MODE, 0, 1>(A, waves, multi, s) : launch_m<MODE, 0, 0>(A, waves, multi, s);
    }
    if constexpr (MODE == kModeWXYZ)
        return var ? launch_m<MODE, 1, 0>(A, waves, multi, s) : launch_m<MODE, 0, 0>(A, waves, multi, s);
    if constexpr (MODE == kModeTan)
        return also ? launch_m<MODE, 0, 1>(A, waves, multi, s) : launch_m<MODE, 0, 0>(A, waves, multi, s);
    else
        return launch_m<MODE, 0, 0>(A, waves, multi, s);
}

}  // namespace hp
