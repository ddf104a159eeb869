// pairs.hip -- workgroup shape and LDS layout of the pair kernels, and the
// dispatch to the per-mode launchers.  The kernel itself is pairs_kernel.h.
#include <cstdlib>

#include "pairs_kernel.h"

namespace hp {

// instantiated in pairs_mode_*.hip
extern template int launch_mode<kModeCount3D>(const PairArgs&, int, int, int, int, hipStream_t);
extern template int launch_mode<kModeProj>(const PairArgs&, int, int, int, int, hipStream_t);
extern template int launch_mode<kModeXYZ>(const PairArgs&, int, int, int, int, hipStream_t);
extern template int launch_mode<kModeObj>(const PairArgs&, int, int, int, int, hipStream_t);
extern template int launch_mode<kModeWXYZ>(const PairArgs&, int, int, int, int, hipStream_t);
extern template int launch_mode<kModeJack>(const PairArgs&, int, int, int, int, hipStream_t);
extern template int launch_mode<kModeJackXY>(const PairArgs&, int, int, int, int, hipStream_t);
extern template int launch_mode<kModeWSum>(const PairArgs&, int, int, int, int, hipStream_t);
extern template int launch_mode<kModeJac>(const PairArgs&, int, int, int, int, hipStream_t);
extern template int launch_mode<kModeTan>(const PairArgs&, int, int, int, int, hipStream_t);

static int align16(int v) { return (v + 15) & ~15; }

// LDS layout of one wave's area for a launch that histograms kc radial / rp
// bins (plus the trash row); returns the bytes before the survivor ring.
static int wave_layout(const PairLaunch& L, int kc, int jslots, PairArgs& A) {
    const int Kp = L.bp.nbins;
    A.shared_count = 0;
    A.cnt_off = 0;
    A.jrow = 0;
    A.jr0_off = 0;
    switch (L.mode) {
        case kModeCount3D:
        case kModeProj:
            if (L.br.nbins > 64) { A.shared_count = 1; return (kc + 1) * 4; }
            return (kc + 1) * 32 * 4;  // rows shared by lanes l and l+32
        case kModeXYZ: return (kc + 1) * (Kp + 1) * 4;
        case kModeWXYZ: return kc * Kp * 8;  // only the cells pairs in range write
        case kModeJack:
        case kModeJackXY: {
            // row 0 plus one row per subsample tag the wave keeps in LDS
            A.jrow = (kc + 1) * (L.mode == kModeJackXY ? Kp + 1 : 1) * 8;
            int b = (jslots + 1) * A.jrow;
            if (L.mode == kModeJack) {
                // all-pairs rows in 16 lane columns for uniform-tag groups
                A.jr0_off = b;
                b += (kc + 1) * 16 * 8;
            }
            return b;
        }
        case kModeObj: return (kc + 1) * 64 * 4;
        case kModeJac:
            A.cnt_off = (kc + 1) * 64 * 8;
            return A.cnt_off + (L.also_count ? (kc + 1) * 64 * 4 : 0);
        case kModeWSum:  // rows shared by 4 lanes (16 columns), counts by 2 (32 columns)
            A.cnt_off = (kc + 1) * 16 * 8;
            return A.cnt_off + (L.also_count ? (kc + 1) * 32 * 4 : 0);
        case kModeTan:
            // per bin one block: W row, dW row (16 columns of 8 B each), then
            // the count row (32 columns of 4 B) when counts are wanted (the
            // kernel's kTanRow and kTanCnt)
            A.cnt_off = 256;
            return (kc + 1) * (L.also_count ? 384 : 256);
        default: return -1;
    }
}

// Jackknife: LDS rows for subsample tags, ~24 KB of table per wave.
static int jack_slots(const PairLaunch& L, int kc) {
    const int jrow = (kc + 1) * (L.mode == kModeJackXY ? L.bp.nbins + 1 : 1) * 8;
    return std::max(0, std::min(L.jslots, 24576 / jrow - 1));
}

int launch_pairs(const PairLaunch& L, hipStream_t s) {
    PairArgs A;
    A.L = L;
    const int Kr = L.br.nbins;
    if (L.br.nb > kTableMax || L.bp.nb > kTableMax) return -1;
    const bool xy = (L.mode == kModeXYZ || L.mode == kModeWXYZ || L.mode == kModeJackXY);
    A.tab_bytes = align16((int)sizeof(LdsEntry) * (L.br.nb + (xy ? L.bp.nb : 0)));
    // Workgroup shape: the number of waves per workgroup that keeps the most
    // waves resident per CU (160 KB of LDS; register-limited waves per CU as
    // the kernels' launch bounds hold them); on ties 4 waves, then 8 (4 was
    // never slower than 8 in measurements of every mode).
    constexpr int kLdsPerCU = 160 * 1024;
    static int lds_per_block = 0;  // the device's limit per workgroup (queried once)
    if (lds_per_block == 0) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
            lds_per_block = std::min(v, kLdsPerCU);
        else
            lds_per_block = 64 * 1024;
    }
    const int kLdsPerBlock = lds_per_block;
    static int num_cus = 0;  // compute units of the device (queried once)
    if (num_cus == 0) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
            num_cus = v;
        else
            num_cus = 256;
    }
    // product marks (weight_func_id 1, one weight per point) have a compiled-in variant
    const int prod = (L.wfunc == 1 && L.nw == 1) ? 1 : 0;
    // waves per SIMD the kernels' registers allow (their launch bounds; the
    // per-lane-row Jacobian kernel is bounded by LDS long before registers)
    const int reg_waves =
        4 * (L.mode == kModeJac ? 6
                                : (((L.mode == kModeWSum || L.mode == kModeWXYZ) && prod) ? 5
                                   : ((L.mode == kModeWSum || L.mode == kModeTan || L.mode == kModeWXYZ ||
                                       L.mode == kModeJack || L.mode == kModeJackXY) ? 4
                                                                                     : (L.mode == kModeXYZ ? 5 : (L.mode == kModeObj ? 6 : HP_COUNT_WAVES)))));
    // shape for a launch of kc bins: sets A's layout, returns the waves per
    // workgroup (0 if nothing fits) and the resident waves per CU
    auto shape = [&](int kc, int* resident) -> int {
        const int wb = wave_layout(L, kc, jack_slots(L, kc), A);
        if (wb < 0) return 0;
        A.ring_off = align16(wb);
        A.wave_bytes = align16(A.ring_off + kRing * 4);
        int waves = 0, best = 0;
        for (const int w : {4, 8, 2, 1}) {
            if (L.waves > 0 && w != L.waves) continue;
            const int bytes = A.tab_bytes + w * A.wave_bytes;
            if (bytes > kLdsPerBlock) continue;
            const int res = std::min(w * (kLdsPerCU / bytes), reg_waves);
            if (res > best) { best = res; waves = w; }
        }
        *resident = best;
        return waves;
    };
    // Bins per launch: all of them when their histograms fit in LDS;
    // otherwise the largest share that keeps 4 waves per CU resident (fewer
    // if even one bin cannot), each launch histogramming its own range of
    // start bins over every pair (the pairs are evaluated once per launch).
    int res = 0;
    int kc = Kr;
    if (shape(Kr, &res) == 0) {
        int lo = 1, hi = Kr - 1, fit = 0;
        while (lo <= hi) {
            const int mid = (lo + hi) / 2;
            int r = 0;
            if (shape(mid, &r) != 0 && r >= std::min(4, reg_waves)) { fit = mid; lo = mid + 1; }
            else hi = mid - 1;
        }
        if (fit == 0) {
            lo = 1, hi = Kr - 1;
            while (lo <= hi) {
                const int mid = (lo + hi) / 2;
                int r = 0;
                if (shape(mid, &r) != 0) { fit = mid; lo = mid + 1; }
                else hi = mid - 1;
            }
        }
        if (fit == 0) return -2;
        kc = fit;
    }
    // launches over a range of the bins take the scanning lookup, which maps
    // start bins outside the range to the trash row
    const int multi = (L.br.multi || L.bp.multi || kc < Kr) ? 1 : 0;
    const int also = L.also_count ? 1 : 0;
    for (int k0 = 0; k0 < Kr; k0 += kc) {
        const int kn = std::min(kc, Kr - k0);
        const int waves = shape(kn, &res);
        if (waves == 0) return -2;
        A.L.jslots = jack_slots(L, kn);
        A.k0 = k0;
        A.kc = kn;
        A.smem = A.tab_bytes + waves * A.wave_bytes;
        A.L.pevals = k0 == 0 ? L.pevals : nullptr;  // every launch evaluates the same pairs
        A.grid = num_cus;  // persistent launches: times the workgroups one CU holds (launch_t)
        // below 2^32 by the largest batch of increments between two checks;
        // HALOPAIR_SHARED_FLUSH_AT lowers it (tests exercise the flush path)
        A.flush_at = 0xF0000000ull;
        if (const char* e = std::getenv("HALOPAIR_SHARED_FLUSH_AT")) {
            const unsigned long long v = std::strtoull(e, nullptr, 10);
            if (v > 0 && v < A.flush_at) A.flush_at = v;
        }
        int rc = -1;
        switch (L.mode) {
            case kModeCount3D: rc = launch_mode<kModeCount3D>(A, waves, multi, also, 0, s); break;
            case kModeProj: rc = launch_mode<kModeProj>(A, waves, multi, also, 0, s); break;
            case kModeXYZ: rc = launch_mode<kModeXYZ>(A, waves, multi, also, 0, s); break;
            case kModeObj: rc = launch_mode<kModeObj>(A, waves, multi, also, 0, s); break;
            case kModeWXYZ: rc = launch_mode<kModeWXYZ>(A, waves, multi, also, prod, s); break;
            case kModeJack: rc = launch_mode<kModeJack>(A, waves, multi, also, 0, s); break;
            case kModeJackXY: rc = launch_mode<kModeJackXY>(A, waves, multi, also, 0, s); break;
            case kModeWSum: rc = launch_mode<kModeWSum>(A, waves, multi, also, prod, s); break;
            case kModeJac: rc = launch_mode<kModeJac>(A, waves, multi, also, L.swap, s); break;
            case kModeTan: rc = launch_mode<kModeTan>(A, waves, multi, also, 0, s); break;
        }
        if (rc) return rc;
    }
    return 0;
}

}  // namespace hp
