// capi.hip -- extern "C" entry points of libhalopair.so (include/halopair.h).
//
// Host-side planning mirrors what the reference does in Python before it calls
// an engine (PC/npairs_3d.py:111-150): square the bins (rbins*rbins,
// PC/cpairs/npairs_3d_engine.pyx:46), choose a mesh, and run the engine over
// every cell.  The GPU mesh is its own (fine cells sized from the search
// radius, not the reference's <= 50 cells per dimension); counts do not depend
// on the mesh because every pair is evaluated with its minimum-image shift.
#include <math.h>

#include <cstring>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "hp_internal.h"

namespace hp {
int launch_reduce_sums_s(const double* partial, int nr, int np, const MockGeom* geom, int nmocks, int side,
                         int64_t max_groups, double* scratch, double* sums, hipStream_t s);
int launch_reduce_evals_s(const uint64_t* pevals, const MockGeom* geom, int nmocks, int side, int64_t max_groups,
                          uint64_t* scratch, int64_t* evals, hipStream_t s);
int64_t reduce_scratch_elems(int64_t max_groups_per_mock, int nmocks, int nb);
}  // namespace hp

using namespace hp;

static thread_local std::string g_err;

static int fail(const std::string& msg) {
    g_err = msg;
    return -1;
}

// ------------------------------------------------------------ bin tables
static uint32_t hiword(double d) {
    uint64_t u;
    std::memcpy(&u, &d, 8);
    return (uint32_t)(u >> 32);
}
static double from_hiword(uint32_t h) {
    uint64_t u = (uint64_t)h << 32;
    double d;
    std::memcpy(&d, &u, 8);
    return d;
}

// Edges as the kernels see them: the reference increments counts[k] for every
// k from K-1 down while dsq <= r2[k] (PC/cpairs/npairs_3d_engine.pyx:178-182),
// so a pair lands in bin k iff dsq <= min(r2[k..K-1]).  The suffix minimum is
// ascending, and the first bin of a pair is #{k : e[k] < dsq}.
static std::vector<double> suffix_min(const double* r2, int K) {
    std::vector<double> e((size_t)K + 2, INFINITY);
    double run = INFINITY;
    for (int k = K - 1; k >= 0; k--) {
        run = std::min(run, r2[k]);
        e[(size_t)k] = run;
    }
    return e;
}

struct Table {
    BinSpec spec{};
    std::vector<double> e;       // K + 2 entries
    std::vector<TabEntry> t;
};

static Table make_table(const std::vector<double>& e_in, int K) {
    Table T;
    T.e = e_in;
    T.spec.nbins = K;
    double first_pos = INFINITY;
    for (int k = 0; k < K; k++)
        if (T.e[k] > 0.0 && T.e[k] < first_pos) first_pos = T.e[k];
    const double last = K > 0 ? T.e[K - 1] : 0.0;
    int shift = 20, base = 0, nb = 2;
    if (first_pos < INFINITY && last < INFINITY) {
        for (shift = 0; shift <= 31; shift++) {
            base = std::max((int)(hiword(first_pos) >> shift) - 1, 0);
            nb = (int)(hiword(last) >> shift) - base + 2;
            if (nb <= kTableMax) break;
        }
    }
    T.spec.shift = shift;
    T.spec.base = base;
    T.spec.basesh = (uint32_t)base << shift;
    T.spec.nb = nb;
    T.t.assign((size_t)nb, TabEntry{0.0, 0u, 0u});
    auto lo = [&](int b) -> double {
        if (b <= 0) return 0.0;
        return from_hiword((uint32_t)((int64_t)(b + base) << shift));
    };
    // e[0..K-1] is ascending: the edges below a bucket and inside it are
    // found by binary search (O(nb log K) for any number of edges)
    int multi = 0;
    const double* e0 = T.e.data();
    for (int b = 0; b < nb; b++) {
        const double l = lo(b);
        const double h = (b + 1 < nb) ? lo(b + 1) : INFINITY;
        const int below = (int)(std::lower_bound(e0, e0 + K, l) - e0);
        const int upto = (b + 1 == nb) ? K : (int)(std::lower_bound(e0, e0 + K, h) - e0);
        T.t[(size_t)b].k = (uint32_t)below;
        T.t[(size_t)b].edge = T.e[(size_t)below];
        if (upto - below > 1) multi = 1;
    }
    T.spec.multi = multi;
    return T;
}

static int lookup_host(const Table& T, double d) {
    const uint32_t hi = hiword(d);
    const uint32_t t = hi > T.spec.basesh ? hi - T.spec.basesh : 0u;
    const uint32_t b = std::min(t >> T.spec.shift, (uint32_t)(T.spec.nb - 1));
    const TabEntry& e = T.t[b];
    if (!T.spec.multi) return (int)e.k + (e.edge < d ? 1 : 0);
    uint32_t k = e.k;
    while (T.e[(size_t)k] < d) ++k;
    return (int)k;
}

// ------------------------------------------------------------ plan
struct Side {
    bool needed = false;
    size_t x = 0, y = 0, z = 0, w = 0, perm = 0, cs = 0;
};

struct Plan {
    int kind = 0, flags = 0, M = 0;
    bool same = false, marked = false, rows1 = false, rows2 = false, sums = false, counts = false;
    int nw = 0;
    int64_t n[2] = {0, 0};
    int64_t ngroups[2] = {0, 0};
    int64_t max_groups[2] = {0, 0};
    int64_t total_cells = 0, total_keys = 0;
    int bits_row = 1, bits_hilbert = 1;  // key bits of the largest mock (keys are local to a mock)
    int tile[2] = {0, 0}, ntiles[2] = {0, 0};  // sort tiles of each side
    std::vector<uint32_t> tfirst[2];           // first tile of each mock, then the tile count
    size_t o_tfirst[2] = {0, 0};
    std::vector<MockGeom> geom;
    Table tr, tp;
    int Kr = 0, Kp = 0, nb_count = 0;
    // workspace layout (byte offsets)
    size_t o_geom = 0, o_er = 0, o_ep = 0, o_tr = 0, o_tp = 0;
    size_t o_gm[2] = {0, 0};
    Side sorted[2][2];  // [data side][order]
    size_t o_keys[2] = {0, 0}, o_vals[2] = {0, 0}, o_tiles = 0, o_digits = 0;
    size_t o_pcount = 0, o_psum = 0, o_pevals = 0, o_red = 0, o_status = 0;
    bool elim = false;
    size_t o_corr = 0, o_rows_tmp = 0;
    size_t o_rows[2] = {0, 0};       // packed coordinate rows of each side with a row-major order
    int row_stride[2] = {4, 4}, row_nw[2] = {0, 0};
    size_t o_keys_row = 0;           // row-major keys of a side sorted in both orders
    bool sym = false;  // auto-count (unweighted or tangent): symmetric evaluation
    int crep = 1;
    bool tangent = false;
    bool obj = false;  // per-object counts
    bool mxyz = false;  // marked_npairs_xy_z: 2-D weighted sums
    bool jack = false;  // jackknife kinds: (w, tag) rows, subsample tables
    int nsamp = 0, jslots = 0;
    size_t o_pk[2] = {0, 0}, o_jacc = 0;
    bool weighted() const { return marked || jack; }
    size_t o_ptan = 0;
    size_t o_queue = 0;
    size_t total = 0;
    std::vector<unsigned char> consts;  // bytes [0, o_gm[0])
};

static size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

constexpr double kCellTarget = 4096.0;  // most stream points per default fine cell
constexpr int kHilbertMaxBits = 8;   // lane-ordering grid <= 256^3 cells per mock
constexpr int kCountReplicas = 256;  // count accumulator copies per mock

static int bits_for(uint64_t range) {
    int b = 1;
    while (b < 32 && ((uint64_t)1 << b) < range) b++;
    return b;
}

static int make_plan(const hp_request* r, int kind, Plan& P) {
    if (!r) return fail("null request");
    P.kind = kind;
    P.flags = r->flags;
    P.M = r->nmocks;
    if (P.M < 1) return fail("nmocks must be >= 1");
    if (!r->mocks) return fail("mocks must not be null");
    P.same = (r->flags & HP_SAME_SAMPLES) != 0;
    P.mxyz = (kind == HP_MARKED_NPAIRS_XY_Z);
    P.marked = (kind == HP_MARKED_NPAIRS_3D) || P.mxyz;
    P.jack = (kind == HP_NPAIRS_JACKKNIFE_3D || kind == HP_NPAIRS_JACKKNIFE_XY_Z);
    P.obj = (kind == HP_NPAIRS_PER_OBJECT_3D);
    P.rows1 = P.marked && (r->flags & HP_ROWS1);
    P.rows2 = P.marked && (r->flags & HP_ROWS2);
    P.sums = P.marked && !(r->flags & HP_NO_SUMS);
    P.counts = (!P.marked && !P.obj && !P.jack) || (P.marked && (r->flags & HP_ALSO_COUNTS));
    P.tangent = P.marked && (r->flags & HP_TANGENT);
    P.nw = P.marked ? r->nw : (P.jack ? 2 : 0);
    if (P.jack) {
        if (r->flags & ~HP_SAME_SAMPLES) return fail("jackknife counts take no flags besides HP_SAME_SAMPLES");
        if (r->n_samples < 1) return fail("n_samples must be >= 1");
        if ((r->w1 || r->w2) && r->nw != 1) return fail("jackknife weights are one per point (nw = 1)");
        P.nsamp = r->n_samples;
        // LDS rows per wave for the tags a group meets; the pair launch caps
        // them further to ~24 KB of table per wave
        P.jslots = std::min(kJackSlots, P.nsamp);
    }
    if (r->tile_waves != 0 && r->tile_waves != 1 && r->tile_waves != 2 && r->tile_waves != 4 && r->tile_waves != 8)
        return fail("tile_waves must be 0, 1, 2, 4 or 8");
    if (P.marked) {
        if (P.mxyz && (r->flags & ~HP_SAME_SAMPLES))
            return fail("marked_npairs_xy_z takes no flags besides HP_SAME_SAMPLES");
        if (P.mxyz && (r->wfunc < 1 || r->wfunc > 15)) return fail("marked_npairs_xy_z: weight_func_id must be in 1..15");
        if (r->wfunc < 1 || r->wfunc > 17) return fail("weight_func_id must be in 1..17");
        if (P.nw < 1 || P.nw > 5) return fail("weights per point must be in 1..5");
        if ((P.rows1 || P.rows2) && (r->wfunc != 1 || P.nw != 1))
            return fail("per-point rows (HP_ROWS1/2) need weight_func_id 1 with one weight per point");
        if (P.tangent && (r->wfunc != 1 || P.nw != 2 || P.rows1 || P.rows2))
            return fail("HP_TANGENT needs weight_func_id 1 with (w, dw/dtheta) per point and no rows");
    }
    P.Kr = r->nr;
    if (P.Kr < 1 || P.Kr > HP_MAX_BINS) return fail("number of radial bins out of range");
    if (!r->r2) return fail("r2 must not be null");
    const bool xyz = (kind == HP_NPAIRS_XY_Z) || P.mxyz || (kind == HP_NPAIRS_JACKKNIFE_XY_Z);  // 2-D histogram
    if (xyz || kind == HP_NPAIRS_PROJECTED) {
        P.Kp = r->np;
        if (P.Kp < 1 || P.Kp > HP_MAX_BINS || !r->p2) return fail("pi bins out of range");
        if (kind == HP_NPAIRS_PROJECTED && P.Kp != 1) return fail("npairs_projected takes one pi_max");
    }
    // any number of radial / rp bins: histograms that do not fit in one
    // launch's LDS are counted over several launches of bin ranges
    if ((int64_t)P.Kr * std::max(P.Kp, 1) > ((int64_t)1 << 26)) return fail("more than 2^26 histogram cells");

    std::vector<double> er = suffix_min(r->r2, P.Kr);
    P.tr = make_table(er, P.Kr);
    double Rxy = sqrt(er[(size_t)P.Kr - 1]);
    double Rz = Rxy;
    int shape = 0;
    if (xyz) {
        std::vector<double> ep = suffix_min(r->p2, P.Kp);
        P.tp = make_table(ep, P.Kp);
        Rz = sqrt(ep[(size_t)P.Kp - 1]);
        shape = 1;
    } else if (kind == HP_NPAIRS_PROJECTED) {
        P.tp.e.assign(3, INFINITY);
        P.tp.e[0] = r->p2[0];
        P.tp.spec.nbins = 1;
        P.tp.spec.nb = 1;
        P.tp.t.assign(1, TabEntry{INFINITY, 0u, 0u});
        Rz = sqrt(r->p2[0]);
        shape = 1;
    } else {
        P.tp.e.assign(3, INFINITY);
        P.tp.spec.nbins = 0;
        P.tp.spec.nb = 1;
        P.tp.t.assign(1, TabEntry{INFINITY, 0u, 0u});
    }
    if (!(Rxy == Rxy) || !(Rz == Rz)) return fail("bins must not be NaN");
    P.nb_count = xyz ? P.Kr * P.Kp : P.Kr;

    // ---- per-mock geometry
    P.geom.resize((size_t)P.M);
    int64_t p_off[2] = {0, 0}, g_off[2] = {0, 0}, cell_off = 0, key_off = 0;
    int64_t max_cells = 1, max_hkeys = 1;
    for (int m = 0; m < P.M; m++) {
        const hp_mock& hm = r->mocks[m];
        MockGeom& G = P.geom[(size_t)m];
        std::memset(&G, 0, sizeof(G));
        const int64_t nm[2] = {hm.n1, P.same ? hm.n1 : hm.n2};
        if (nm[0] < 0 || nm[1] < 0) return fail("negative point count");
        double Lmax = 0.0;
        int64_t ncells = 1;
        // default fine-cell edge: the search length.  The pair kernel culls
        // stream points one by one against each lane group's box, so coarse
        // cells cost only a cheap filter pass, while finer cells add segments
        // (scalar round trips) per group; kCellTarget stream points per cell
        // is the floor for very dense samples
        const double volume = hm.period[0] * hm.period[1] * hm.period[2];
        const double density = volume > 0.0 ? (double)nm[1] / volume : 0.0;
        const double natural = density > 0.0 ? cbrt(kCellTarget / density) : INFINITY;
        for (int d = 0; d < 3; d++) {
            const double L = hm.period[d];
            if (!(L > 0.0) || !(L < INFINITY)) return fail("period / box length must be positive and finite");
            const double R = (d < 2) ? Rxy : Rz;
            if (hm.pbc && R > L / 3.0 * (1.0 + 1e-12))
                return fail("search length exceeds period/3 (the reference's mesh limit)");
            double b = hm.cell[d] > 0.0 ? hm.cell[d] : std::min(std::max(natural, R / 3.0), R);
            if (!(b > 0.0)) b = L;  // zero search radius
            double nd = floor(L / b);
            if (nd < 1.0) nd = 1.0;
            if (nd > 1024.0) nd = 1024.0;
            G.ndiv[d] = (int32_t)nd;
            G.cell[d] = L / (double)G.ndiv[d];
            G.inv_cell[d] = 1.0 / G.cell[d];
            G.period[d] = L;
            Lmax = std::max(Lmax, L);
            ncells *= G.ndiv[d];
        }
        G.pbc = hm.pbc ? 1 : 0;
        // lane-ordering grid: about one point per Hilbert cell, so a run of 64
        // consecutive points follows the curve closely and stays compact
        const double nlane = (double)std::max(nm[0], nm[1]);
        int mb = 1;
        while (mb < kHilbertMaxBits && (double)(1 << mb) < 1.25 * cbrt(nlane)) mb++;
        G.mbits = mb;
        for (int d = 0; d < 3; d++) G.inv_hcell[d] = (double)(1 << mb) / G.period[d];
        const float eps = (float)(1e-5 * Lmax);
        G.eps = eps;
        G.search[0] = (float)(Rxy * (1.0 + 1e-6)) + eps;
        G.search[1] = G.search[0];
        G.search[2] = (float)(Rz * (1.0 + 1e-6)) + eps;
        G.r2xy = G.search[0] * G.search[0] * (1.0f + 1e-6f);
        G.shape = shape;
        for (int s = 0; s < 2; s++) {
            G.p_off[s] = p_off[s];
            G.n[s] = nm[s];
            G.group_off[s] = g_off[s];
            G.ngroups[s] = (nm[s] + kGroup - 1) / kGroup;
            P.max_groups[s] = std::max(P.max_groups[s], G.ngroups[s]);
            p_off[s] += nm[s];
            g_off[s] += G.ngroups[s];
        }
        G.cell_off = cell_off;
        G.key_off = key_off;
        cell_off += ncells;
        key_off += (int64_t)1 << (3 * mb);
        max_cells = std::max<int64_t>(max_cells, (int64_t)ncells);
        max_hkeys = std::max<int64_t>(max_hkeys, (int64_t)1 << (3 * mb));
        if (nm[1] > 0 && (double)nm[1] * 27.0 >= 4.0e9) return fail("too many points in one mock for 32-bit histograms");
        // the pair kernel addresses a mock's stream rows (32 bytes per point) and
        // weights with 32-bit byte offsets
        for (int s = 0; s < 2; s++)
            if ((double)nm[s] * std::max(32.0, 8.0 * (double)std::max(P.nw, 1)) >= 4294967296.0)
                return fail("too many points (times weights per point) in one mock for 32-bit offsets");
    }
    P.n[0] = p_off[0];
    P.n[1] = p_off[1];
    if (P.n[0] >= (int64_t)0x7fffffff || P.n[1] >= (int64_t)0x7fffffff) return fail("too many points (>= 2^31)");
    P.ngroups[0] = g_off[0];
    P.ngroups[1] = g_off[1];
    P.total_cells = cell_off;
    P.total_keys = key_off;
    if (cell_off >= (int64_t)0xffffffff || key_off > (int64_t)0xffffffff) return fail("mesh too large for 32-bit keys");
    P.bits_row = bits_for((uint64_t)max_cells);
    P.bits_hilbert = bits_for((uint64_t)max_hkeys);
    // sort tiles per side: a mock's points start a new tile
    for (int s = 0; s < 2; s++) {
        P.tile[s] = sort_tile(P.n[s]);
        P.tfirst[s].assign((size_t)P.M + 1, 0u);
        int64_t t = 0;
        for (int m = 0; m < P.M; m++) {
            P.tfirst[s][m] = (uint32_t)t;
            t += (P.geom[m].n[s] + P.tile[s] - 1) / P.tile[s];
        }
        P.tfirst[s][P.M] = (uint32_t)t;
        if (t >= (int64_t)0x7fffffff) return fail("too many sort tiles");
        P.ntiles[s] = (int)std::max<int64_t>(t, 1);
    }

    // ---- workspace layout
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = align_up(off, 256);
        off = o + bytes;
        return o;
    };
    P.o_geom = take(sizeof(MockGeom) * (size_t)P.M);
    P.o_er = take(P.tr.e.size() * 8);
    P.o_ep = take(P.tp.e.size() * 8);
    P.o_tr = take(P.tr.t.size() * sizeof(TabEntry));
    P.o_tp = take(P.tp.t.size() * sizeof(TabEntry));
    P.o_tfirst[0] = take(((size_t)P.M + 1) * 4);
    P.o_tfirst[1] = take(((size_t)P.M + 1) * 4);
    const size_t consts_end = off;
    P.o_status = take(16);
    P.o_gm[0] = take((size_t)P.ngroups[0] * 4 + 4);
    if (P.rows2) P.o_gm[1] = take((size_t)(P.same ? P.ngroups[0] : P.ngroups[1]) * 4 + 4);
    // sorted sides: pass A lanes = (0, Hilbert), stream = (s2, RowMajor);
    // pass B lanes = (s2, Hilbert), stream = (0, RowMajor)
    const int s2 = P.same ? 0 : 1;
    P.sorted[0][kHilbert].needed = true;
    P.sorted[s2][kRowMajor].needed = true;
    if (P.rows2) {
        P.sorted[s2][kHilbert].needed = true;
        P.sorted[0][kRowMajor].needed = true;
    }
    int64_t nmax = 1;
    for (int s = 0; s < 2; s++) {
        for (int o = 0; o < 2; o++) {
            Side& S = P.sorted[s][o];
            if (!S.needed) continue;
            const size_t n = (size_t)std::max<int64_t>(P.n[s], 1);
            nmax = std::max<int64_t>(nmax, (int64_t)n);
            if (o == kRowMajor) {
                // one 32-byte row per point: x, y, z and the Hilbert order key
                // of the symmetric auto-count (else zero)
                S.x = take(n * 32);
                S.y = S.x + 8;
                S.z = S.x + 16;
                if (P.weighted()) S.w = take(n * 8 * (size_t)P.nw);
            }
            S.perm = take(n * 4);
            if (o == kRowMajor) S.cs = take((size_t)(P.total_cells + 1) * 4);
        }
    }
    // digit counts per tile: room for the smaller tiles whichever size runs
    const int64_t ntiles = (nmax + kSortTile - 1) / kSortTile + P.M + 1;
    P.o_keys[0] = take((size_t)nmax * 4);
    P.o_keys_row = take((size_t)nmax * 4);
    P.o_keys[1] = take((size_t)nmax * 4);
    P.o_vals[0] = take((size_t)nmax * 4);
    P.o_vals[1] = take((size_t)nmax * 4);
    P.o_tiles = take((size_t)ntiles * kRadixDigits * 4);
    P.o_digits = take((size_t)kRadixDigits * 4);
    const int64_t gA = std::max<int64_t>(P.ngroups[0], 1);
    // count accumulators: up to kCountReplicas copies per mock (a group adds
    // into copy group % crep), so the atomics of many groups spread over cache
    // lines; wide histograms (xy_z) spread over their own bins already
    P.crep = (int)std::max<int64_t>(
        1, std::min<int64_t>({(int64_t)kCountReplicas, P.max_groups[0], (int64_t)(2048 / std::max(P.nb_count, 1))}));
    if (P.counts) P.o_pcount = take((size_t)P.M * (size_t)P.crep * (size_t)P.nb_count * 8);
    if (P.sums) P.o_psum = take((size_t)gA * (size_t)P.nb_count * 8);
    if (P.tangent) P.o_ptan = take((size_t)gA * (size_t)P.Kr * 8);
    P.o_pevals = take((size_t)gA * 8);
    P.o_queue = take(64 * 4);  // work-queue counters of the pair launches
    const int64_t red = reduce_scratch_elems(std::max<int64_t>(P.max_groups[0], 1), P.M,
                                             P.nb_count);  // sums (Kr or Kr * Kp) and evals (1)
    P.o_red = take((size_t)red * 8);
    // auto-count with shared weights: rows2 = rows1 + seam corrections, no swapped pass
    P.elim = P.rows2 && P.same && P.marked && r->wfunc == 1;
    if (P.elim) {
        P.o_corr = take((size_t)std::max<int64_t>(P.n[0], 1) * (size_t)P.Kr * 8);
        if (!P.rows1) P.o_rows_tmp = take((size_t)std::max<int64_t>(P.n[0], 1) * (size_t)P.Kr * 8);
    }
    // unweighted auto-counts evaluate each unshifted unordered pair once; they
    // need every row-major point's position in the Hilbert order
    P.sym = P.same && !P.obj && (!P.marked || P.tangent);
    // packed rows (x, y, z, Hilbert position[, up to 4 weights]) of every side
    // the stream side is built from: one random row read per point in the gather
    for (int s = 0; s < 2; s++) {
        if (!P.sorted[s][kRowMajor].needed) continue;
        const bool pack_w = P.weighted() && P.nw <= 4;
        P.row_nw[s] = pack_w ? P.nw : 0;
        P.row_stride[s] = pack_w ? 8 : 4;
        P.o_rows[s] = take((size_t)std::max<int64_t>(P.n[s], 1) * (size_t)P.row_stride[s] * 8);
    }
    if (P.jack) {
        // (w, tag) rows of each side, and the per-mock subsample tables
        P.o_pk[0] = take((size_t)std::max<int64_t>(P.n[0], 1) * 16);
        if (!P.same) P.o_pk[1] = take((size_t)std::max<int64_t>(P.n[1], 1) * 16);
        const double tab = (double)P.M * (P.nsamp + 1.0) * (double)P.nb_count * 8.0;
        if (tab > 64.0 * 1024 * 1024 * 1024) return fail("jackknife tables exceed 64 GiB");
        P.o_jacc = take((size_t)tab);
    }
    P.total = align_up(off, 256);

    // ---- constant blob (geometry, edges, tables) that goes to the workspace
    P.tr.spec.edges_off = (int64_t)P.o_er;
    P.tr.spec.table_off = (int64_t)P.o_tr;
    P.tp.spec.edges_off = (int64_t)P.o_ep;
    P.tp.spec.table_off = (int64_t)P.o_tp;
    P.consts.assign(consts_end, 0);
    std::memcpy(P.consts.data() + P.o_geom, P.geom.data(), sizeof(MockGeom) * (size_t)P.M);
    std::memcpy(P.consts.data() + P.o_er, P.tr.e.data(), P.tr.e.size() * 8);
    std::memcpy(P.consts.data() + P.o_ep, P.tp.e.data(), P.tp.e.size() * 8);
    std::memcpy(P.consts.data() + P.o_tr, P.tr.t.data(), P.tr.t.size() * sizeof(TabEntry));
    std::memcpy(P.consts.data() + P.o_tp, P.tp.t.data(), P.tp.t.size() * sizeof(TabEntry));
    for (int s = 0; s < 2; s++) std::memcpy(P.consts.data() + P.o_tfirst[s], P.tfirst[s].data(), P.tfirst[s].size() * 4);
    return 0;
}

static SortedSide side_ptrs(char* ws, const Side& S) {
    SortedSide o;
    o.x = S.x ? reinterpret_cast<double*>(ws + S.x) : nullptr;
    o.y = S.y ? reinterpret_cast<double*>(ws + S.y) : nullptr;
    o.z = S.z ? reinterpret_cast<double*>(ws + S.z) : nullptr;
    o.w = S.w ? reinterpret_cast<double*>(ws + S.w) : nullptr;
    o.perm = reinterpret_cast<uint32_t*>(ws + S.perm);
    o.cell_start = S.cs ? reinterpret_cast<uint32_t*>(ws + S.cs) : nullptr;
    o.xs = 4;  // the row-major copies are 32-byte rows (x, y, z, order key)
    return o;
}

// Lane side: the input arrays (point i at x[i * xs] ...), read through the
// Hilbert permutation.
static SortedSide lane_ptrs(char* ws, const Side& S, const double* x, const double* y, const double* z,
                            const double* w, int64_t xs) {
    SortedSide o = side_ptrs(ws, S);
    o.xs = xs;
    o.x = const_cast<double*>(x);
    o.y = const_cast<double*>(y);
    o.z = const_cast<double*>(z);
    o.w = const_cast<double*>(w);
    return o;
}

// Named host-side ranges of a request's phases (planning, cell build, pair
// kernels, reductions) for timeline tools (rocprofv3 --marker-trace); a range
// spans the enqueueing of that phase's work on the stream.  Without a tool
// attached a push or pop costs a function call.
struct PhaseRange {
    bool open = false;
    void begin(const char* name) {
        end();
        roctxRangePush(name);
        open = true;
    }
    void end() {
        if (open) roctxRangePop();
        open = false;
    }
    ~PhaseRange() { end(); }
};

static int run(const hp_request* r, int kind) {
    PhaseRange phase;
    phase.begin("halopair: plan");
    Plan P;
    if (make_plan(r, kind, P)) return -1;
    if (!r->workspace) return fail("workspace must not be null");
    if (r->workspace_bytes < P.total) return fail("workspace too small");
    if (P.n[0] > 0 && (!r->x1 || !r->y1 || !r->z1)) return fail("sample1 coordinates must not be null");
    if (r->stride1 < 0 || r->stride2 < 0) return fail("coordinate strides must not be negative");
    if (!P.same && P.n[1] > 0 && (!r->x2 || !r->y2 || !r->z2))
        return fail("sample2 coordinates must not be null");
    if (P.marked && ((P.n[0] > 0 && !r->w1) || (!P.same && P.n[1] > 0 && !r->w2)))
        return fail("weights must not be null");
    if (P.jack && ((P.n[0] > 0 && !r->jtags1) || (!P.same && P.n[1] > 0 && !r->jtags2)))
        return fail("jackknife tags must not be null");
    if (P.jack && !r->sums) return fail("sums output must not be null");
    if ((P.counts || P.obj) && !r->counts) return fail("counts output must not be null");
    if (P.sums && !r->sums) return fail("sums output must not be null");
    if (P.tangent && !r->tangents) return fail("tangents output must not be null");
    if (P.rows1 && !r->rows1) return fail("rows1 output must not be null");
    if (P.rows2 && !r->rows2) return fail("rows2 output must not be null");
    hipStream_t s = static_cast<hipStream_t>(r->stream);
    char* ws = static_cast<char*>(r->workspace);
    const MockGeom* dgeom = reinterpret_cast<const MockGeom*>(ws + P.o_geom);

    if (launch_store_bytes(ws, P.consts.data(), P.consts.size(), s)) return fail("constant upload failed");
    int32_t* status = r->status ? r->status : reinterpret_cast<int32_t*>(ws + P.o_status);
    launch_fill_u32(reinterpret_cast<uint32_t*>(status), 0u, 1, s);

    const double* X[2] = {r->x1, P.same ? r->x1 : r->x2};
    // doubles between consecutive points of each sample's coordinate arrays
    const int64_t XS[2] = {r->stride1 > 1 ? r->stride1 : 1,
                           P.same ? (r->stride1 > 1 ? r->stride1 : 1) : (r->stride2 > 1 ? r->stride2 : 1)};
    const double* Y[2] = {r->y1, P.same ? r->y1 : r->y2};
    const double* Z[2] = {r->z1, P.same ? r->z1 : r->z2};
    const double* W[2] = {r->w1, P.same ? r->w1 : r->w2};
    if (P.jack) {
        double* pk0 = reinterpret_cast<double*>(ws + P.o_pk[0]);
        double* pk1 = P.same ? pk0 : reinterpret_cast<double*>(ws + P.o_pk[1]);
        if (launch_pack_jack(r->w1, r->jtags1, P.n[0], P.nsamp, pk0, status, s) ||
            (!P.same && launch_pack_jack(r->w2, r->jtags2, P.n[1], P.nsamp, pk1, status, s)))
            return fail("jackknife pack launch failed");
        W[0] = pk0;
        W[1] = pk1;
    }

    SortBufs sb;
    sb.keys[0] = reinterpret_cast<uint32_t*>(ws + P.o_keys[0]);
    sb.keys[1] = reinterpret_cast<uint32_t*>(ws + P.o_keys[1]);
    sb.vals[0] = reinterpret_cast<uint32_t*>(ws + P.o_vals[0]);
    sb.vals[1] = reinterpret_cast<uint32_t*>(ws + P.o_vals[1]);
    sb.tile_counts = reinterpret_cast<uint32_t*>(ws + P.o_tiles);
    sb.digit_totals = reinterpret_cast<uint32_t*>(ws + P.o_digits);
    phase.begin("halopair: cell build");
    if (r->build_events[0] && hipEventRecord(static_cast<hipEvent_t>(r->build_events[0]), s) != hipSuccess)
        return fail("event record failed");
    for (int sd = 0; sd < 2; sd++) {
        for (int o = 0; o < 2; o++) {
            const Side& S = P.sorted[sd][o];
            if (!S.needed) continue;
            // with SAME_SAMPLES only side 0 exists; the geometry rows of side 0 describe it.
            // Symmetric auto-counts: the Hilbert key pass (first) leaves each
            // point's Hilbert key in its row, the row-major gather turns it into
            // the order keys.
            const bool rank = P.sym && sd == 0 && P.n[0] > 0;
            // a side sorted in both orders computes its keys once (Hilbert build first)
            const bool both = P.sorted[sd][kHilbert].needed && P.sorted[sd][kRowMajor].needed;
            uint32_t* krow = both ? reinterpret_cast<uint32_t*>(ws + P.o_keys_row) : nullptr;
            PackedRows rows;
            if (P.sorted[sd][kRowMajor].needed) {
                rows.p = reinterpret_cast<double*>(ws + P.o_rows[sd]);
                rows.stride = P.row_stride[sd];
                rows.nw = P.row_nw[sd];
            }
            static_assert(kHilbert < kRowMajor, "the Hilbert order must be built before the row-major order");
            SortTiles st;
            st.first = reinterpret_cast<const uint32_t*>(ws + P.o_tfirst[sd]);
            st.tile = P.tile[sd];
            st.ntiles = P.ntiles[sd];
            if (build_sorted_side(dgeom, P.M, sd, (Order)o, X[sd], Y[sd], Z[sd], P.weighted() ? W[sd] : nullptr, P.nw,
                                  P.n[sd], P.total_cells, o == kRowMajor ? P.bits_row : P.bits_hilbert, sb, st,
                                  side_ptrs(ws, S), status, s, rows,
                                  o == kRowMajor && rank, o == kHilbert ? krow : nullptr,
                                  o == kRowMajor ? krow : nullptr, XS[sd]))
                return fail("cell build launch failed");
        }
    }
    // the order keys: the fourth word of each row of side 0's row-major copy
    const uint64_t* hkey =
        (P.sym && P.n[0] > 0) ? reinterpret_cast<const uint64_t*>(ws + P.sorted[0][kRowMajor].x) + 3 : nullptr;
    if (r->build_events[1] && hipEventRecord(static_cast<hipEvent_t>(r->build_events[1]), s) != hipSuccess)
        return fail("event record failed");

    const int s2 = P.same ? 0 : 1;
    int32_t* gm0 = reinterpret_cast<int32_t*>(ws + P.o_gm[0]);
    if (launch_group_mock(dgeom, P.M, 0, gm0, P.ngroups[0], s)) return fail("group table launch failed");

    PairLaunch L;
    std::memset(&L, 0, sizeof(L));
    L.geom = dgeom;
    L.group_mock = gm0;
    L.ngroups = P.ngroups[0];
    L.lane = lane_ptrs(ws, P.sorted[0][kHilbert], X[0], Y[0], Z[0], P.weighted() ? W[0] : nullptr, XS[0]);
    L.strm = side_ptrs(ws, P.sorted[s2][kRowMajor]);
    L.nw = P.nw;
    L.br = P.tr.spec;
    L.bp = P.tp.spec;
    L.ws = ws;
    L.wfunc = r->wfunc;
    L.waves = r->tile_waves;
    L.hkey = hkey;
    L.hrows = hkey ? reinterpret_cast<const double*>(ws + P.o_rows[0]) : nullptr;
    L.hrow_stride = P.row_stride[0];
    L.pcount = P.counts ? reinterpret_cast<uint64_t*>(ws + P.o_pcount) : nullptr;
    L.psum = P.sums ? reinterpret_cast<double*>(ws + P.o_psum) : nullptr;
    L.ptan = P.tangent ? reinterpret_cast<double*>(ws + P.o_ptan) : nullptr;
    L.pevals = reinterpret_cast<uint64_t*>(ws + P.o_pevals);
    L.queue = reinterpret_cast<uint32_t*>(ws + P.o_queue);
    switch (kind) {
        case HP_NPAIRS_3D: L.mode = kModeCount3D; break;
        case HP_NPAIRS_PROJECTED: L.mode = kModeProj; break;
        case HP_NPAIRS_XY_Z: L.mode = kModeXYZ; break;
        case HP_NPAIRS_PER_OBJECT_3D:
            L.mode = kModeObj;
            L.orows = r->counts;
            break;
        case HP_MARKED_NPAIRS_XY_Z: L.mode = kModeWXYZ; break;
        case HP_NPAIRS_JACKKNIFE_3D:
        case HP_NPAIRS_JACKKNIFE_XY_Z:
            L.mode = (kind == HP_NPAIRS_JACKKNIFE_XY_Z) ? kModeJackXY : kModeJack;
            L.jacc = reinterpret_cast<double*>(ws + P.o_jacc);
            L.nsamp = P.nsamp;
            L.jslots = P.jslots;
            launch_fill_u32(reinterpret_cast<uint32_t*>(L.jacc), 0u, (int64_t)P.M * (P.nsamp + 1) * P.nb_count * 2,
                            s);
            break;
        default:
            L.mode = P.tangent ? kModeTan
                               : ((r->wfunc == 1 && P.nw == 1 && (P.rows1 || P.rows2)) ? kModeJac : kModeWSum);
            L.also_count = P.counts ? 1 : 0;
            L.rows = P.rows1 ? r->rows1 : nullptr;
            break;
    }
    double* corr = nullptr;
    if (P.elim) {
        corr = reinterpret_cast<double*>(ws + P.o_corr);
        if (!P.rows1) L.rows = reinterpret_cast<double*>(ws + P.o_rows_tmp);
        L.corr = corr;
        launch_fill_u32(reinterpret_cast<uint32_t*>(corr), 0u, P.n[0] * (int64_t)P.Kr * 2, s);
    }
    L.crep = P.crep;
    if (P.counts)
        launch_fill_u32(reinterpret_cast<uint32_t*>(L.pcount), 0u, (int64_t)P.M * P.crep * P.nb_count * 2, s);
    int rc = 0;
    phase.begin("halopair: pair kernels");
    if (r->pair_events[0] && hipEventRecord(static_cast<hipEvent_t>(r->pair_events[0]), s) != hipSuccess)
        return fail("event record failed");
    if (!(P.marked && !P.sums && !P.counts && !P.rows1 && !P.elim)) {
        rc = launch_pairs(L, s);
        if (rc == -2)
            return fail("len(pi_bins) too large: one rp row of the (rp, pi) histogram must fit in the LDS of a "
                        "workgroup");
        if (rc) return fail("pair kernel launch failed");
    }
    if (P.elim) {
        if (launch_add_rows(L.rows, corr, r->rows2, P.n[0] * (int64_t)P.Kr, s)) return fail("row add launch failed");
    } else if (P.rows2) {
        // pass B: lanes are sample 2, the stream is sample 1 (same arithmetic,
        // (x1 - x2shift) - x2, with the shift applied to the sample-1 point)
        int32_t* gm1 = reinterpret_cast<int32_t*>(ws + P.o_gm[1]);
        if (launch_group_mock(dgeom, P.M, s2, gm1, P.ngroups[s2], s)) return fail("group table launch failed");
        PairLaunch B = L;
        B.mode = kModeJac;
        B.swap = 1;
        B.also_count = 0;
        B.group_mock = gm1;
        B.ngroups = P.ngroups[s2];
        B.lane = lane_ptrs(ws, P.sorted[s2][kHilbert], X[s2], Y[s2], Z[s2], W[s2], XS[s2]);
        B.strm = side_ptrs(ws, P.sorted[0][kRowMajor]);
        B.rows = r->rows2;
        B.pcount = nullptr;
        B.psum = nullptr;
        B.pevals = nullptr;
        B.hkey = nullptr;
        B.hrows = nullptr;
        // with SAME_SAMPLES the side-1 rows of the geometry equal the side-0 rows
        if (launch_pairs(B, s)) return fail("pair kernel (swapped) launch failed");
    }
    if (r->pair_events[1] && hipEventRecord(static_cast<hipEvent_t>(r->pair_events[1]), s) != hipSuccess)
        return fail("event record failed");
    phase.begin("halopair: reductions");
    uint64_t* red = reinterpret_cast<uint64_t*>(ws + P.o_red);
    if (P.counts) {
        if (launch_prefix_counts(L.pcount, P.Kr, kind == HP_NPAIRS_XY_Z ? P.Kp : 0, P.M, P.crep, r->counts, s))
            return fail("reduction launch failed");
    }
    if (P.jack) {
        if (launch_jack_finish(L.jacc, P.Kr, kind == HP_NPAIRS_JACKKNIFE_XY_Z ? P.Kp : 0, P.M, P.nsamp, r->sums, s))
            return fail("jackknife reduction launch failed");
    }
    if (P.sums) {
        if (launch_reduce_sums_s(L.psum, P.Kr, P.mxyz ? P.Kp : 0, dgeom, P.M, 0, P.max_groups[0],
                                 reinterpret_cast<double*>(red), r->sums, s))
            return fail("reduction launch failed");
    }
    if (P.tangent) {
        if (launch_reduce_sums_s(L.ptan, P.Kr, 0, dgeom, P.M, 0, P.max_groups[0], reinterpret_cast<double*>(red),
                                 r->tangents, s))
            return fail("reduction launch failed");
    }
    if (r->evals) {
        if (launch_reduce_evals_s(L.pevals, dgeom, P.M, 0, P.max_groups[0], red, r->evals, s))
            return fail("reduction launch failed");
    }
    if (hipGetLastError() != hipSuccess) return fail("HIP launch error");
    g_err.clear();
    return 0;
}

// ------------------------------------------------------------ no-allocation variant
// hp_npairs_3d_noalloc runs on a fixed scratch array that every GPU the
// library runs on holds from module load until the process ends, instead of
// a caller-allocated workspace.  Host code must never pass the array's name
// to a kernel (it is a host-side stand-in there); each device's own address
// of it comes from hipGetSymbolAddress, called once per device with that
// device current, and only that address is used.  Calls that overlap on one
// GPU share the array and race.
static_assert(HP_NOALLOC_SCRATCH_BYTES % 256 == 0, "scratch size must keep 256-byte alignment");
static __device__ __align__(256) unsigned char scratch[HP_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

static int scratch_of_current_device(void** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return fail("hp_npairs_3d_noalloc: no current HIP device");
    if (dev >= kMaxDevices) return fail("hp_npairs_3d_noalloc: device index beyond the library's table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(std::string("hp_npairs_3d_noalloc: hipGetSymbolAddress(scratch) failed: ") +
                        hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

extern "C" {

const char* hp_last_error(void) { return g_err.c_str(); }

int hp_abi_version(void) { return HP_ABI_VERSION; }

size_t hp_workspace_size(const hp_request* req) {
    Plan P;
    if (!req) return 0;
    if (make_plan(req, req->kind, P)) return 0;
    return P.total;
}

int hp_npairs_3d(const hp_request* req) { return run(req, HP_NPAIRS_3D); }

int hp_npairs_3d_noalloc(const hp_request* req) {
    if (!req) return fail("null request");
    if (req->workspace || req->workspace_bytes)
        return fail("hp_npairs_3d_noalloc takes no workspace (workspace must be NULL and workspace_bytes 0)");
    Plan P;
    if (make_plan(req, HP_NPAIRS_3D, P)) return -1;
    if (P.total > (size_t)HP_NOALLOC_SCRATCH_BYTES)
        return fail("hp_npairs_3d_noalloc: the request needs " + std::to_string(P.total) +
                    " bytes of scratch, more than the library's fixed " +
                    std::to_string((size_t)HP_NOALLOC_SCRATCH_BYTES) +
                    " bytes (use hp_npairs_3d with a workspace of hp_workspace_size() bytes)");
    void* ptr = nullptr;
    if (scratch_of_current_device(&ptr)) return -1;
    hp_request r = *req;
    r.workspace = ptr;
    r.workspace_bytes = (size_t)HP_NOALLOC_SCRATCH_BYTES;
    return run(&r, HP_NPAIRS_3D);
}
int hp_npairs_projected(const hp_request* req) { return run(req, HP_NPAIRS_PROJECTED); }
int hp_npairs_xy_z(const hp_request* req) { return run(req, HP_NPAIRS_XY_Z); }
int hp_marked_npairs_3d(const hp_request* req) { return run(req, HP_MARKED_NPAIRS_3D); }
int hp_npairs_per_object_3d(const hp_request* req) { return run(req, HP_NPAIRS_PER_OBJECT_3D); }
int hp_marked_npairs_xy_z(const hp_request* req) { return run(req, HP_MARKED_NPAIRS_XY_Z); }
int hp_npairs_jackknife_3d(const hp_request* req) { return run(req, HP_NPAIRS_JACKKNIFE_3D); }
int hp_npairs_jackknife_xy_z(const hp_request* req) { return run(req, HP_NPAIRS_JACKKNIFE_XY_Z); }

int hp_marked_backward(const double* rows, int32_t nr, const double* gsuffix, const int64_t* offsets,
                       int32_t nmocks, double* grad, void* stream) {
    if (!rows || !gsuffix || !offsets || !grad) return fail("null pointer");
    if (nr < 1 || nmocks < 1) return fail("bad sizes");
    if (launch_rows_dot(rows, nr, gsuffix, offsets, nmocks, grad, static_cast<hipStream_t>(stream)))
        return fail("backward launch failed");
    g_err.clear();
    return 0;
}

int64_t hp_segment_sum_scratch(const int64_t* offsets, int32_t nmocks) {
    if (!offsets || nmocks < 1) return fail("bad arguments");
    for (int m = 0; m < nmocks; m++)
        if (offsets[m + 1] < offsets[m]) return fail("offsets must be non-decreasing");
    g_err.clear();
    return segment_sum_scratch(offsets, nmocks);
}

int hp_segment_sum(const double* values, const int64_t* offsets, int32_t nmocks, double* out, double* scratch,
                   int64_t scratch_elems, void* stream) {
    if (!offsets || !out || nmocks < 1) return fail("bad arguments");
    if (!values && offsets[nmocks] > offsets[0]) return fail("values must not be null");
    for (int m = 0; m < nmocks; m++)
        if (offsets[m + 1] < offsets[m]) return fail("offsets must be non-decreasing");
    if (!scratch || scratch_elems < segment_sum_scratch(offsets, nmocks)) return fail("scratch too small");
    if (launch_segment_sum(values, offsets, nmocks, out, scratch, static_cast<hipStream_t>(stream)))
        return fail("segment sum launch failed");
    g_err.clear();
    return 0;
}

int hp_zheng07_tangent(const double* log_mass, const int64_t* mock_id, const double* logMmin, double sigma,
                       int64_t n, double* wt, void* stream) {
    if (n < 0) return fail("n must be >= 0");
    if (n > 0 && (!log_mass || !mock_id || !logMmin || !wt)) return fail("null argument");
    if (!(sigma > 0.0) || !(sigma < INFINITY)) return fail("sigma must be positive and finite");
    if (launch_zheng07_tangent(log_mass, mock_id, logMmin, sigma, n, wt, static_cast<hipStream_t>(stream)))
        return fail("weights launch failed");
    g_err.clear();
    return 0;
}

// Host-only diagnostics: the bucketed bin lookup the kernels use, evaluated on
// the CPU for a set of squared distances (tests pin it against searchsorted).
int hp_debug_bin_lookup(const double* r2, int32_t K, const double* d, int64_t n, int32_t* out, int32_t* info) {
    if (!r2 || !d || !out || K < 1) return fail("bad arguments");
    std::vector<double> e = suffix_min(r2, K);
    Table T = make_table(e, K);
    for (int64_t i = 0; i < n; i++) out[i] = lookup_host(T, d[i]);
    if (info) {
        info[0] = T.spec.shift;
        info[1] = T.spec.base;
        info[2] = T.spec.nb;
        info[3] = T.spec.multi;
    }
    return 0;
}

// Host-only diagnostic: Hilbert index of a lane-ordering cell (tests check the
// curve is a bijection whose consecutive cells share a face).
uint32_t hp_debug_hilbert(uint32_t x, uint32_t y, uint32_t z, int32_t bits) { return hilbert3(x, y, z, bits); }

}  // extern "C"
