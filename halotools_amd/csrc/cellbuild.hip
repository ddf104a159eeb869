// cellbuild.hip -- GPU cell-list build (replaces RectangularMesh,
// PC/rectangular_mesh.py:181-225, whose argsort + searchsorted bucket the
// points of each sample into a uniform grid).
//
// The build is a stable LSD counting sort of 32-bit cell keys (8-bit digits,
// tiles of 2048 or 8192 keys), so the order of points inside a cell is
// their input order and every downstream floating-point sum is run-to-run
// deterministic.  Two key flavours exist:
//   row-major  (ix*ny + iy)*nz + iz  -- stream side; the cells of one (ix, iy)
//              column are contiguous, so a column's z-range is one segment;
//   Hilbert    3-D Hilbert index    -- lane side; 64 consecutive points form a
//              spatially compact group (the curve has no jumps).
// Keys are local to a mock; the sort's tiles keep mocks apart, so one sort
// handles a whole batch of realisations with the key bits of one mock.
#include <cstdlib>
#include <cstring>

#include "hp_internal.h"

namespace hp {

// ----------------------------------------------------------------- utilities
struct ByteChunk {
    uint32_t* dst;
    uint32_t nwords;
    uint32_t words[508];  // 2032 bytes of payload per launch
};

__global__ void k_store_words(const ByteChunk c) {
    for (uint32_t i = threadIdx.x; i < c.nwords; i += blockDim.x) c.dst[i] = c.words[i];
}

int launch_store_bytes(void* dst, const void* src, size_t nbytes, hipStream_t s) {
    // Carries a host blob into device memory through kernel arguments: no
    // memcpy API, so the upload is legal inside HIP-graph capture.
    const size_t nwords_total = (nbytes + 3) / 4;
    const uint8_t* bytes = static_cast<const uint8_t*>(src);
    for (size_t w0 = 0; w0 < nwords_total; w0 += 508) {
        ByteChunk c;
        const size_t nw = nwords_total - w0 < 508 ? nwords_total - w0 : 508;
        c.dst = static_cast<uint32_t*>(dst) + w0;
        c.nwords = static_cast<uint32_t>(nw);
        const size_t b0 = w0 * 4;
        const size_t nb = (b0 + nw * 4 <= nbytes) ? nw * 4 : nbytes - b0;
        std::memset(c.words, 0, sizeof(c.words));
        std::memcpy(c.words, bytes + b0, nb);
        hipLaunchKernelGGL(k_store_words, dim3(1), dim3(256), 0, s, c);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

int launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s) {
    if (n <= 0) return 0;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s, p, v, n);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

__global__ void k_group_mock(const MockGeom* __restrict__ g, int nmocks, int side, int32_t* __restrict__ gm,
                             int64_t ngroups) {
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= ngroups) return;
    int lo = 0, hi = nmocks - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (g[mid].group_off[side] <= q) lo = mid; else hi = mid - 1;
    }
    gm[q] = lo;
}

int launch_group_mock(const MockGeom* g, int nmocks, int side, int32_t* gm, int64_t ngroups, hipStream_t s) {
    if (ngroups <= 0) return 0;
    hipLaunchKernelGGL(k_group_mock, dim3((unsigned)((ngroups + 255) / 256)), dim3(256), 0, s, g, nmocks, side,
                       gm, ngroups);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ----------------------------------------------------------------- cell keys
// Cell index of a coordinate: floor(v / cell), clamped into [0, n-1] so that
// v == period lands in the last cell (as PC/rectangular_mesh.py:19-22 does).
// The quotient is formed as v * (1 / cell): the pair kernels plan their
// windows with a safety margin, so the rounding of the quotient (a point
// within rounding of a cell face may land in the cell next to it) does not
// affect which pairs are counted.
__device__ __forceinline__ int cell_index(double v, double inv_cell, int n, bool& bad, double period, int pbc) {
    if (!(v == v) || v > 1.0e300 || v < -1.0e300) { bad = true; return 0; }
    if (pbc && (v < 0.0 || v > period)) bad = true;
    const double f = floor(v * inv_cell);
    if (!(f >= 0.0)) return 0;
    if (f >= (double)n) return n - 1;
    return (int)f;
}

// keys_row: when the same side is also sorted in row-major order, the
// Hilbert pass writes those keys too (one read of the coordinates for both).
// rows: the point's packed row (coordinates, and weights when they fit).
__global__ void k_keys(const MockGeom* __restrict__ geom, int nmocks, int side, int order,
                       const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
                       int64_t n, uint32_t* __restrict__ keys, int32_t* __restrict__ status,
                       uint32_t* __restrict__ keys_row, const double* __restrict__ w, int nw, PackedRows rows,
                       int64_t xs) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= n) return;
    // the mock of the point: a wave's 64 consecutive points nearly always lie
    // in one mock, found by two wave-uniform searches (scalar loads); a wave
    // that straddles mocks searches per lane
    const int pf = __builtin_amdgcn_readfirstlane((int)p);
    const int64_t mf = mock_of_point(geom, nmocks, side, pf);
    const int64_t ml = mock_of_point(geom, nmocks, side, min((int64_t)pf + 63, n - 1));
    const int64_t m = (mf == ml) ? mf : mock_of_point(geom, nmocks, side, p);
    const MockGeom& G = geom[m];
    const double px = x[p * xs], py = y[p * xs], pz = z[p * xs];
    bool bad = false;
    const int ix = cell_index(px, G.inv_cell[0], G.ndiv[0], bad, G.period[0], G.pbc);
    const int iy = cell_index(py, G.inv_cell[1], G.ndiv[1], bad, G.period[1], G.pbc);
    const int iz = cell_index(pz, G.inv_cell[2], G.ndiv[2], bad, G.period[2], G.pbc);
    // keys are local to the mock: the sort keeps mocks apart by their tiles
    uint32_t key;
    const uint32_t rkey = (uint32_t)(((int64_t)ix * G.ndiv[1] + iy) * G.ndiv[2] + iz);
    if (order == kRowMajor) {
        key = rkey;
    } else {
        if (keys_row) keys_row[p] = rkey;
        // lane order: a power-of-two grid of 2^mbits cells per dimension, so the
        // Hilbert curve covers the box exactly (no passes through empty padding)
        const int hb = G.mbits;
        const int hn = 1 << hb;
        bool dummy = false;
        const int hx = cell_index(px, G.inv_hcell[0], hn, dummy, G.period[0], 0);
        const int hy = cell_index(py, G.inv_hcell[1], hn, dummy, G.period[1], 0);
        const int hz = cell_index(pz, G.inv_hcell[2], hn, dummy, G.period[2], 0);
        key = hilbert3((uint32_t)hx, (uint32_t)hy, (uint32_t)hz, hb);
    }
    keys[p] = key;
    if (rows.p) {
        // the row's fourth word: the Hilbert key (Hilbert order), which with the
        // point's index fixes its place in the stable Hilbert sort
        double* r = rows.p + p * rows.stride;
        const double slot = __hiloint2double(0, order == kRowMajor ? 0 : (int)key);
        reinterpret_cast<double4*>(r)[0] = make_double4(px, py, pz, slot);
        if (rows.nw > 0) {
            double4 wr = make_double4(0.0, 0.0, 0.0, 0.0);
            wr.x = w[p * nw];
            if (rows.nw > 1) wr.y = w[p * nw + 1];
            if (rows.nw > 2) wr.z = w[p * nw + 2];
            if (rows.nw > 3) wr.w = w[p * nw + 3];
            reinterpret_cast<double4*>(r)[1] = wr;
        }
    }
    if (bad && status) atomicOr(status, 1);
}

// ----------------------------------------------------- LSD counting sort
// Tiles of TILE keys: per-tile digit counts, an exclusive scan of each digit's
// counts over the tiles, then a scatter of every tile in digit order.  Large
// sorts take bigger tiles: the per-tile digit counts are read and written at
// 256 scattered words per tile, and a tile's keys leave in 256 runs, so 8192
// keys per tile (runs of ~32 keys, full cache lines) move a fraction of the
// bytes of 2048-key tiles (runs of ~8 keys) at tens of millions of points.
//
// Mocks are sorted side by side: the points of a mock are contiguous, its
// tiles never cross into the next mock (tile_first[m] = first tile of mock m,
// tile_first[nmocks] = tiles in all), and a mock's keys go to its own range of
// positions, so the keys only need the bits of the largest mock (one or two
// passes fewer than keys offset by mock for a batch of mocks).
struct TileSpan {
    int64_t start;  // first key of the tile
    int tn;         // keys in the tile
    int m;          // its mock
};

__device__ __forceinline__ TileSpan tile_span(const uint32_t* __restrict__ tile_first, const MockGeom* __restrict__ geom,
                                              int nmocks, int side, int t, int tile) {
    int lo = 0, hi = nmocks - 1;  // the last mock whose first tile is <= t (skips empty mocks)
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((int)tile_first[mid] <= t) lo = mid; else hi = mid - 1;
    }
    TileSpan r;
    r.m = lo;
    const int64_t k = (int64_t)(t - (int)tile_first[lo]) * tile;
    r.start = geom[lo].p_off[side] + k;
    r.tn = (int)min((int64_t)tile, geom[lo].n[side] - k);
    return r;
}

template <int TILE>
__global__ void __launch_bounds__(256) k_radix_hist(const uint32_t* __restrict__ keys, const MockGeom* __restrict__ geom,
                                                    const uint32_t* __restrict__ tile_first, int nmocks, int side,
                                                    int shift, uint32_t* __restrict__ tile_counts, int ntiles) {
    __shared__ uint32_t h[kRadixDigits];
    h[threadIdx.x] = 0;
    __syncthreads();
    const TileSpan ts = tile_span(tile_first, geom, nmocks, side, (int)blockIdx.x, TILE);
    if (ts.tn == TILE && (ts.start & 3) == 0) {
        // full, aligned tile: 16-byte loads
        const uint4* k4 = reinterpret_cast<const uint4*>(keys + ts.start);
        for (int i = threadIdx.x; i < TILE / 4; i += 256) {
            const uint4 k = k4[i];
            atomicAdd(&h[(k.x >> shift) & (kRadixDigits - 1)], 1u);
            atomicAdd(&h[(k.y >> shift) & (kRadixDigits - 1)], 1u);
            atomicAdd(&h[(k.z >> shift) & (kRadixDigits - 1)], 1u);
            atomicAdd(&h[(k.w >> shift) & (kRadixDigits - 1)], 1u);
        }
    } else {
        for (int i = threadIdx.x; i < ts.tn; i += 256)
            atomicAdd(&h[(keys[ts.start + i] >> shift) & (kRadixDigits - 1)], 1u);
    }
    __syncthreads();
    tile_counts[(int64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of each digit's row of per-tile counts; row totals out.
__global__ void __launch_bounds__(256) k_radix_scan(uint32_t* __restrict__ tile_counts, int ntiles,
                                                    uint32_t* __restrict__ digit_totals) {
    __shared__ uint32_t part[256];
    uint32_t* row = tile_counts + (int64_t)blockIdx.x * ntiles;
    const int per = (ntiles + 255) / 256;
    const int b0 = threadIdx.x * per;
    const int b1 = min(b0 + per, ntiles);
    uint32_t sum = 0;
    for (int i = b0; i < b1; i++) sum += row[i];
    part[threadIdx.x] = sum;
    __syncthreads();
    // Hillis-Steele inclusive scan of 256 partial sums
    for (int off = 1; off < 256; off <<= 1) {
        uint32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0;
    for (int i = b0; i < b1; i++) {
        const uint32_t c = row[i];
        row[i] = run;
        run += c;
    }
    if (threadIdx.x == 255) digit_totals[blockIdx.x] = part[255];
}

// Inclusive Hillis-Steele scan of v over the first 256 threads of the block
// (every thread of the block calls it; threads >= 256 pass 0 and get 0).
template <int NT>
__device__ __forceinline__ uint32_t scan_digits(uint32_t* scan, uint32_t v) {
    const int d = threadIdx.x;
    if (d < kRadixDigits) scan[d] = v;
    __syncthreads();
    for (int off = 1; off < kRadixDigits; off <<= 1) {
        const uint32_t u = (d < kRadixDigits && d >= off) ? scan[d - off] : 0u;
        __syncthreads();
        if (d < kRadixDigits) scan[d] += u;
        __syncthreads();
    }
    return d < kRadixDigits ? scan[d] : 0u;
}

// One workgroup of WAVES wavefronts per tile of TILE keys.  Each wave ranks
// its TILE / WAVES keys stably (a lane's rank among lanes with the same digit
// comes from eight ballots; per-wave digit counters carry it across rounds),
// the tile is staged in LDS in digit order, and the staged keys go out in
// runs of equal digits, so consecutive threads write consecutive addresses.
// A mock's keys go to its own positions: position p_off[m] + (keys of the
// mock with smaller digits) + (keys of its earlier tiles with the same digit)
// + rank in the tile, all from the exclusive prefix G[d][t] of the digit
// counts over every tile.  The last row-major pass writes global cell ids
// (key + the mock's first cell) for the cell-start table.
template <int TILE, int WAVES>
__global__ void __launch_bounds__(WAVES * 64) k_radix_scatter(const uint32_t* __restrict__ kin,
                                                              const uint32_t* __restrict__ vin,
                                                              uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                              const MockGeom* __restrict__ geom,
                                                              const uint32_t* __restrict__ tile_first, int nmocks,
                                                              int side, int shift, int add_cell_off,
                                                              const uint32_t* __restrict__ tile_counts,
                                                              const uint32_t* __restrict__ digit_totals, int ntiles) {
    constexpr int NT = WAVES * 64;
    constexpr int kPerWave = TILE / WAVES;
    constexpr int kRounds = kPerWave / 64;
    static_assert(NT >= kRadixDigits && kPerWave % 64 == 0, "tile shape");
    __shared__ uint32_t skey[TILE];
    __shared__ uint32_t sval[TILE];
    __shared__ uint32_t wcnt[WAVES][kRadixDigits];  // per-wave digit counts, then offsets
    __shared__ uint32_t toff[kRadixDigits];         // first staged position of each digit
    __shared__ uint32_t gbase[kRadixDigits];        // global position of the tile's first key per digit
    __shared__ uint32_t scan[kRadixDigits];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = tid >> 6;
    const int tile = blockIdx.x;
    const TileSpan ts = tile_span(tile_first, geom, nmocks, side, tile, TILE);
    const int64_t base = ts.start;
    const int tn = ts.tn;
    const uint32_t koff = add_cell_off ? (uint32_t)geom[ts.m].cell_off : 0u;

    for (int i = tid; i < WAVES * kRadixDigits; i += NT) (&wcnt[0][0])[i] = 0u;
    __syncthreads();

    // ---- stable ranks within each wave
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t kk[kRounds], vv[kRounds], rk[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; r++) {
        const int li = wv * kPerWave + r * 64 + lane;  // position in the tile
        const bool valid = li < tn;
        const uint32_t key = valid ? kin[base + li] : 0u;
        const uint32_t val = valid ? (vin ? vin[base + li] : (uint32_t)(base + li)) : 0u;
        const uint32_t d = (key >> shift) & (kRadixDigits - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < kRadixBits; bit++) {
            const bool on = (d >> bit) & 1u;
            const uint64_t bm = __ballot(on);
            peers &= on ? bm : ~bm;
        }
        const uint32_t before = (uint32_t)__popcll(peers & lt);
        const uint32_t cnt = valid ? wcnt[wv][d] : 0u;
        __builtin_amdgcn_wave_barrier();
        if (valid && before == 0) wcnt[wv][d] = cnt + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
        kk[r] = key;
        vv[r] = val;
        rk[r] = valid ? cnt + before : 0xffffffffu;
    }
    __syncthreads();

    // ---- digit offsets in the staged tile: digits in order, waves in order
    const int d = tid;  // the first 256 threads own one digit each
    uint32_t tot = 0;
    if (d < kRadixDigits)
        for (int w = 0; w < WAVES; w++) tot += wcnt[w][d];
    const uint32_t incl = scan_digits<NT>(scan, tot);
    if (d < kRadixDigits) {
        const uint32_t start = incl - tot;
        toff[d] = start;
        uint32_t run = start;
        for (int w = 0; w < WAVES; w++) {
            const uint32_t c = wcnt[w][d];
            wcnt[w][d] = run;
            run += c;
        }
    }
    // global base: the mock's first position, the mock's keys of smaller
    // digits (an exclusive scan of the mock's digit totals), then this digit's
    // keys of the mock's earlier tiles
    const int f0 = (int)tile_first[ts.m], f1 = (int)tile_first[ts.m + 1];
    uint32_t t = 0, g0 = 0;
    if (d < kRadixDigits) {
        const uint32_t* row = tile_counts + (int64_t)d * ntiles;
        g0 = row[f0];
        t = (f1 < ntiles ? row[f1] : digit_totals[d]) - g0;
    }
    __syncthreads();
    const uint32_t tincl = scan_digits<NT>(scan, t);
    if (d < kRadixDigits)
        gbase[d] = (uint32_t)geom[ts.m].p_off[side] + tincl - t + (tile_counts[(int64_t)d * ntiles + tile] - g0);
    __syncthreads();

    // ---- stage in digit order
#pragma unroll
    for (int r = 0; r < kRounds; r++) {
        if (rk[r] == 0xffffffffu) continue;
        const uint32_t dg = (kk[r] >> shift) & (kRadixDigits - 1);
        const uint32_t pos = wcnt[wv][dg] + rk[r];
        skey[pos] = kk[r];
        sval[pos] = vv[r];
    }
    __syncthreads();

    // ---- write out: runs of equal digits land on consecutive addresses
    for (int i = tid; i < tn; i += NT) {
        const uint32_t key = skey[i];
        const uint32_t dg = (key >> shift) & (kRadixDigits - 1);
        const uint32_t pos = gbase[dg] + (uint32_t)i - toff[dg];
        kout[pos] = key + koff;
        vout[pos] = sval[i];
    }
}

// Sorted copies in the new order: one 32-byte row per point (x, y, z and, with
// want_hkey, the point's Hilbert order key (Hilbert key << 32) | input index --
// the stable Hilbert sort orders the points of a mock by exactly this value --
// else zero), and the weight rows, read through each point's packed input row
// (x, y, z, the Hilbert key, and the weights when they fit in the row).
// With `keys` (the sorted global cell ids) the same pass writes the cell
// starts: cell_start[c] = first sorted position whose key >= c, for c in
// [0, total_cells]; thread q = n writes the trailing empty cells.
__global__ void k_gather(const uint32_t* __restrict__ perm, int64_t n, PackedRows rows, const double* __restrict__ w,
                         int nw, double* __restrict__ orow, double* __restrict__ ow, int want_hkey,
                         const uint32_t* __restrict__ keys, int64_t total_cells, uint32_t* __restrict__ cell_start) {
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q > n) return;
    if (keys) {
        const int64_t kcur = (q < n) ? (int64_t)keys[q] : total_cells;
        const int64_t kprev = (q > 0) ? (int64_t)keys[q - 1] : -1;
        for (int64_t c = kprev + 1; c <= kcur && c <= total_cells; c++) cell_start[c] = (uint32_t)q;
    }
    if (q == n) return;
    const int64_t p = perm[q];
    const double* r = rows.p + p * rows.stride;
    const double4 c = reinterpret_cast<const double4*>(r)[0];
    const uint64_t hk = want_hkey ? (((uint64_t)(uint32_t)__double2loint(c.w) << 32) | (uint64_t)(uint32_t)p) : 0ull;
    // bit copies (the order key is not a floating-point value)
    reinterpret_cast<ulonglong4*>(orow)[q] =
        make_ulonglong4((unsigned long long)__double_as_longlong(c.x), (unsigned long long)__double_as_longlong(c.y),
                        (unsigned long long)__double_as_longlong(c.z), (unsigned long long)hk);
    if (ow) {
        if (rows.nw > 0) {
            const double4 wr = reinterpret_cast<const double4*>(r)[1];
            ow[q * nw] = wr.x;
            if (nw > 1) ow[q * nw + 1] = wr.y;
            if (nw > 2) ow[q * nw + 2] = wr.z;
            if (nw > 3) ow[q * nw + 3] = wr.w;
        } else {
            for (int k = 0; k < nw; k++) ow[q * nw + k] = w[p * nw + k];
        }
    }
}

static inline int check(const char*) { return hipGetLastError() == hipSuccess ? 0 : -1; }

// Keys per counting-sort tile for a sort of n keys: 8192 from kSortLargeMin
// keys on (HALOPAIR_SORT_TILE=2048 or 8192 forces one size, for tests and
// measurements).
int sort_tile(int64_t n) {
    if (const char* e = std::getenv("HALOPAIR_SORT_TILE")) {
        const int v = std::atoi(e);
        if (v == kSortTile || v == kSortTileLarge) return v;
    }
    return n >= kSortLargeMin ? kSortTileLarge : kSortTile;
}

int build_sorted_side(const MockGeom* geom, int nmocks, int side, Order order, const double* x, const double* y,
                      const double* z, const double* w, int nw, int64_t n, int64_t total_cells, int key_bits,
                      SortBufs sb, SortTiles st, SortedSide out, int32_t* status, hipStream_t s, PackedRows rows,
                      bool want_hkey, uint32_t* keys_row_out, const uint32_t* keys_in, int64_t xs) {
    if (n <= 0) {
        if (order == kRowMajor) return launch_fill_u32(out.cell_start, 0u, total_cells + 1, s);
        return 0;
    }
    const unsigned nb = (unsigned)((n + 255) / 256);
    // keys_in: this order's keys were written by the other order's key pass
    if (!keys_in)
        hipLaunchKernelGGL(k_keys, dim3(nb), dim3(256), 0, s, geom, nmocks, side, (int)order, x, y, z, n, sb.keys[0],
                           status, order == kHilbert ? keys_row_out : (uint32_t*)nullptr, w, nw, rows, xs);
    const int tile = st.tile, ntiles = st.ntiles;
    if (tile != kSortTile && tile != kSortTileLarge) return -1;
    if (ntiles <= 0) return -1;
    const int passes = key_bits <= 0 ? 1 : (key_bits + kRadixBits - 1) / kRadixBits;
    int cur = 0;
    for (int pass = 0; pass < passes; pass++) {
        const int shift = pass * kRadixBits;
        const uint32_t* kin = (pass == 0 && keys_in) ? keys_in : sb.keys[cur];
        if (tile == kSortTileLarge)
            hipLaunchKernelGGL(k_radix_hist<kSortTileLarge>, dim3(ntiles), dim3(256), 0, s, kin, geom, st.first,
                               nmocks, side, shift, sb.tile_counts, ntiles);
        else
            hipLaunchKernelGGL(k_radix_hist<kSortTile>, dim3(ntiles), dim3(256), 0, s, kin, geom, st.first, nmocks,
                               side, shift, sb.tile_counts, ntiles);
        hipLaunchKernelGGL(k_radix_scan, dim3(kRadixDigits), dim3(256), 0, s, sb.tile_counts, ntiles,
                           sb.digit_totals);
        // the last pass writes the permutation straight into out.perm, and
        // (row-major) the global cell ids for the cell-start table
        const bool last = pass == passes - 1;
        uint32_t* vdst = last ? out.perm : sb.vals[cur ^ 1];
        const uint32_t* vsrc = pass == 0 ? (const uint32_t*)nullptr : sb.vals[cur];
        const int add_off = (last && order == kRowMajor) ? 1 : 0;
        if (tile == kSortTileLarge)
            hipLaunchKernelGGL((k_radix_scatter<kSortTileLarge, 8>), dim3(ntiles), dim3(512), 0, s, kin, vsrc,
                               sb.keys[cur ^ 1], vdst, geom, st.first, nmocks, side, shift, add_off, sb.tile_counts,
                               sb.digit_totals, ntiles);
        else
            hipLaunchKernelGGL((k_radix_scatter<kSortTile, 4>), dim3(ntiles), dim3(256), 0, s, kin, vsrc,
                               sb.keys[cur ^ 1], vdst, geom, st.first, nmocks, side, shift, add_off, sb.tile_counts,
                               sb.digit_totals, ntiles);
        cur ^= 1;
    }
    if (check("sort")) return -1;
    // The lane side (Hilbert order) is read once per group through the
    // permutation; only the stream side needs sorted copies.
    if (order == kRowMajor) {
        // the gather of the packed rows and the cell starts in one pass
        if (!rows.p) return -1;  // the row-major gather reads the packed rows
        hipLaunchKernelGGL(k_gather, dim3((unsigned)((n + 1 + 255) / 256)), dim3(256), 0, s, out.perm, n, rows, w,
                           nw, out.x, out.w, want_hkey ? 1 : 0, sb.keys[cur], total_cells, out.cell_start);
    }
    return check("gather");
}

}  // namespace hp
