// pairs.hip -- workgroup shape and LDS layout of the pair kernels, and the
// dispatch to the per-mode launchers.  The kernel itself is pairs_kernel.h.
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

int launch_pairs(const PairLaunch& L, hipStream_t s) {
    PairArgs A;
    A.L = L;
    const int Kr = L.br.nbins, Kp = L.bp.nbins;
    A.shared_count = 0;
    A.cnt_off = 0;
    A.tan_off = 0;
    A.jrow = 0;
    A.jr0_off = 0;
    int wave_bytes = 0;
    switch (L.mode) {
        case kModeCount3D:
        case kModeProj:
            if (Kr > 64) { A.shared_count = 1; wave_bytes = (Kr + 1) * 4; }
            else wave_bytes = (Kr + 1) * 32 * 4;  // rows shared by lanes l and l+32
            break;
        case kModeXYZ: wave_bytes = (Kr + 1) * (Kp + 1) * 4; break;
        case kModeWXYZ: wave_bytes = (Kr + 1) * (Kp + 1) * 8; break;
        case kModeJack:
        case kModeJackXY:
            // row 0 plus one row per subsample tag the wave keeps in LDS
            A.jrow = (Kr + 1) * (L.mode == kModeJackXY ? Kp + 1 : 1) * 8;
            wave_bytes = (L.jslots + 1) * A.jrow;
            if (L.mode == kModeJack) {
                // all-pairs rows in 16 lane columns for uniform-tag groups
                A.jr0_off = wave_bytes;
                wave_bytes += (Kr + 1) * 16 * 8;
            }
            break;
        case kModeObj: wave_bytes = (Kr + 1) * 64 * 4; break;
        case kModeJac:
            A.cnt_off = (Kr + 1) * 64 * 8;
            wave_bytes = A.cnt_off + (L.also_count ? (Kr + 1) * 64 * 4 : 0);
            break;
        case kModeWSum:  // rows shared by 4 lanes (16 columns), counts by 2 (32 columns)
            A.cnt_off = (Kr + 1) * 16 * 8;
            wave_bytes = A.cnt_off + (L.also_count ? (Kr + 1) * 32 * 4 : 0);
            break;
        case kModeTan:
            A.tan_off = (Kr + 1) * 16 * 8;
            A.cnt_off = 2 * A.tan_off;
            wave_bytes = A.cnt_off + (L.also_count ? (Kr + 1) * 32 * 4 : 0);
            break;
        default: return -1;
    }
    if (L.br.nb > kTableMax || L.bp.nb > kTableMax) return -1;
    A.ring_off = align16(wave_bytes);
    A.wave_bytes = align16(A.ring_off + kRing * 4);
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
    const int reg_waves =
        4 * (L.mode == kModeJac ? 6
                                : ((L.mode == kModeWSum || L.mode == kModeTan || L.mode == kModeWXYZ ||
                                    L.mode == kModeJack || L.mode == kModeJackXY) ? 4
                                                                                                      : (L.mode == kModeXYZ ? 5 : HP_COUNT_WAVES)));
    int waves = 0, best = -1;
    if (L.waves > 0) {
        waves = L.waves;
    } else {
        for (const int w : {4, 8, 2, 1}) {
            const int bytes = A.tab_bytes + w * A.wave_bytes;
            if (bytes > kLdsPerBlock) continue;
            const int resident = std::min(w * (kLdsPerCU / bytes), reg_waves);
            if (resident > best) { best = resident; waves = w; }
        }
        if (waves == 0) return -2;
    }
    A.smem = A.tab_bytes + waves * A.wave_bytes;
    if (A.smem > kLdsPerBlock) return -2;
    const int multi = (L.br.multi || L.bp.multi) ? 1 : 0;
    const int also = L.also_count ? 1 : 0;
    switch (L.mode) {
        case kModeCount3D: return launch_mode<kModeCount3D>(A, waves, multi, also, 0, s);
        case kModeProj: return launch_mode<kModeProj>(A, waves, multi, also, 0, s);
        case kModeXYZ: return launch_mode<kModeXYZ>(A, waves, multi, also, 0, s);
        case kModeObj: return launch_mode<kModeObj>(A, waves, multi, also, 0, s);
        case kModeWXYZ: return launch_mode<kModeWXYZ>(A, waves, multi, also, 0, s);
        case kModeJack: return launch_mode<kModeJack>(A, waves, multi, also, 0, s);
        case kModeJackXY: return launch_mode<kModeJackXY>(A, waves, multi, also, 0, s);
        case kModeWSum: return launch_mode<kModeWSum>(A, waves, multi, also, 0, s);
        case kModeJac: return launch_mode<kModeJac>(A, waves, multi, also, L.swap, s);
        case kModeTan: return launch_mode<kModeTan>(A, waves, multi, also, 0, s);
    }
    return -1;
}

}  // namespace hp
