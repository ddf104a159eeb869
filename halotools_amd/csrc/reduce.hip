// reduce.hip -- fixed-order reductions of per-group partial histograms into
// the reference's cumulative outputs, and the backward contraction.
//
// The reference sums per-process count vectors with np.sum(..., axis=0)
// (PC/npairs_3d.py:142-146) and returns cumulative counts N(<= r_k) because its
// engines walk the bins from the top down (PC/cpairs/npairs_3d_engine.pyx:178-182).
// Here each group contributed the histogram of the *first* bin a pair falls in;
// an inclusive prefix sum over bins (2-D for xy_z) turns it into the same
// cumulative counts.  All sums run in a fixed order, so results are
// run-to-run bitwise reproducible, including the FP64 marked sums.
#include "hp_internal.h"

namespace hp {

constexpr int kChunk = 512;  // groups per stage-1 block

template <typename T>
__global__ void __launch_bounds__(256) k_reduce_s1(const T* __restrict__ partial, int nb, const MockGeom* __restrict__ geom,
                                                   int side, int nchunk_max, T* __restrict__ s1) {
    __shared__ T sh[256];
    const int m = blockIdx.y, c = blockIdx.x;
    const int64_t go = geom[m].group_off[side], ng = geom[m].ngroups[side];
    const int64_t r0 = (int64_t)c * kChunk;
    const int64_t r1 = r0 + kChunk < ng ? r0 + kChunk : ng;
    T* out = s1 + ((int64_t)m * nchunk_max + c) * nb;
    const int nbt = nb < 256 ? nb : 256;
    const int rstride = 256 / nbt;
    for (int b0 = 0; b0 < nb; b0 += nbt) {
        const int t = threadIdx.x;
        const int b = b0 + t % nbt;
        const int rr = t / nbt;
        T acc = 0;
        if (rr < rstride && b < nb) {
            for (int64_t r = r0 + rr; r < r1; r += rstride) acc += partial[(go + r) * nb + b];
        }
        sh[t] = acc;
        __syncthreads();
        if (t < nbt && b0 + t < nb) {
            T tot = 0;
            for (int q = 0; q < rstride; q++) tot += sh[q * nbt + t];
            out[b0 + t] = tot;
        }
        __syncthreads();
    }
}

// Stage 2: per mock, sum the chunk rows in order, then the inclusive prefix.
// dims: nr (and np for the 2-D xy_z histogram; np = 0 means 1-D).
template <typename T, typename TO>
__global__ void __launch_bounds__(256) k_reduce_s2(const T* __restrict__ s1, int nr, int np, const MockGeom* __restrict__ geom,
                                                   int side, int nchunk_max, TO* __restrict__ out) {
    const int m = blockIdx.x;
    const int nb = np > 0 ? nr * np : nr;
    const int64_t ng = geom[m].ngroups[side];
    const int nchunks = (int)((ng + kChunk - 1) / kChunk);
    TO* o = out + (int64_t)m * nb;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        T tot = 0;
        for (int c = 0; c < nchunks; c++) tot += s1[((int64_t)m * nchunk_max + c) * nb + b];
        o[b] = (TO)tot;
    }
    __threadfence_block();
    __syncthreads();
    if (np <= 0) {
        if (threadIdx.x == 0) {
            TO run = 0;
            for (int k = 0; k < nr; k++) { run += o[k]; o[k] = run; }
        }
    } else {
        for (int k = threadIdx.x; k < nr; k += blockDim.x) {
            TO run = 0;
            for (int q = 0; q < np; q++) { run += o[k * np + q]; o[k * np + q] = run; }
        }
        __threadfence_block();
        __syncthreads();
        for (int q = threadIdx.x; q < np; q += blockDim.x) {
            TO run = 0;
            for (int k = 0; k < nr; k++) { run += o[k * np + q]; o[k * np + q] = run; }
        }
    }
}

// Counts: the pair kernel accumulates each mock's histogram directly (exact
// integer atomics); this turns it into the reference's cumulative counts.
__global__ void __launch_bounds__(256) k_prefix_counts(const uint64_t* __restrict__ acc, int nr, int np, int crep,
                                                       int64_t* __restrict__ out) {
    __shared__ uint64_t sh[256];
    const int m = blockIdx.x;
    const int nb = np > 0 ? nr * np : nr;
    const uint64_t* a = acc + (int64_t)m * crep * nb;
    int64_t* o = out + (int64_t)m * nb;
    // replica sums: threads split over (bin, replica stripe), then the stripes
    // are added in LDS (integer sums: exact in any order)
    const int nbt = nb < 256 ? nb : 256;
    const int rstride = 256 / nbt;
    for (int b0 = 0; b0 < nb; b0 += nbt) {
        const int t = threadIdx.x;
        const int b = b0 + t % nbt;
        const int rr = t / nbt;
        uint64_t acc_t = 0;
        if (rr < rstride && b < nb)
            for (int r = rr; r < crep; r += rstride) acc_t += a[(int64_t)r * nb + b];
        sh[t] = acc_t;
        __syncthreads();
        if (t < nbt && b0 + t < nb) {
            uint64_t tot = 0;
            for (int q = 0; q < rstride; q++) tot += sh[q * nbt + t];
            o[b0 + t] = (int64_t)tot;
        }
        __syncthreads();
    }
    __threadfence_block();
    __syncthreads();
    if (np <= 0) {
        if (threadIdx.x == 0) {
            int64_t run = 0;
            for (int k = 0; k < nr; k++) { run += o[k]; o[k] = run; }
        }
    } else {
        for (int k = threadIdx.x; k < nr; k += blockDim.x) {
            int64_t run = 0;
            for (int q = 0; q < np; q++) { run += o[k * np + q]; o[k * np + q] = run; }
        }
        __threadfence_block();
        __syncthreads();
        for (int q = threadIdx.x; q < np; q += blockDim.x) {
            int64_t run = 0;
            for (int k = 0; k < nr; k++) { run += o[k * np + q]; o[k * np + q] = run; }
        }
    }
}

int launch_prefix_counts(const uint64_t* acc, int nr, int np, int nmocks, int crep, int64_t* counts, hipStream_t s) {
    hipLaunchKernelGGL(k_prefix_counts, dim3(nmocks), dim3(256), 0, s, acc, nr, np, crep, counts);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Plain per-mock totals (no prefix): used for the candidate-evaluation counter.
__global__ void __launch_bounds__(256) k_reduce_plain(const uint64_t* __restrict__ s1, const MockGeom* __restrict__ geom,
                                                      int side, int nchunk_max, int64_t* __restrict__ out) {
    const int m = blockIdx.x;
    const int64_t ng = geom[m].ngroups[side];
    const int nchunks = (int)((ng + kChunk - 1) / kChunk);
    if (threadIdx.x == 0) {
        uint64_t tot = 0;
        for (int c = 0; c < nchunks; c++) tot += s1[(int64_t)m * nchunk_max + c];
        out[m] = (int64_t)tot;
    }
}

static int64_t nchunk_of(int64_t max_groups) { return (max_groups + kChunk - 1) / kChunk; }

// Wide histograms (the (rp, pi) cells of marked xy_z, hundreds of cells per
// group): stage 1 takes chunks of kChunkW groups, one cell per thread, so a
// launch has (groups / 64) x (cells / 256) blocks instead of groups / 512;
// stage 2 sums a mock's chunk rows in four fixed stripes per cell, then the
// stripes in order; the 2-D prefix runs once per mock.
constexpr int kChunkW = 64;
constexpr int kWideCells = 64;  // histograms above this many cells take the wide path
static int64_t nchunk_w(int64_t max_groups) { return (max_groups + kChunkW - 1) / kChunkW; }

__global__ void __launch_bounds__(256) k_reduce_w1(const double* __restrict__ partial, int nb,
                                                   const MockGeom* __restrict__ geom, int side, int nchunk_max,
                                                   double* __restrict__ s1) {
    const int m = blockIdx.y, c = blockIdx.x;
    const int b = blockIdx.z * 256 + threadIdx.x;
    const int64_t go = geom[m].group_off[side], ng = geom[m].ngroups[side];
    const int64_t r0 = (int64_t)c * kChunkW;
    if (b >= nb || r0 >= ng) return;
    const int64_t r1 = r0 + kChunkW < ng ? r0 + kChunkW : ng;
    double acc = 0.0;
    for (int64_t r = r0; r < r1; r++) acc += partial[(go + r) * nb + b];
    s1[((int64_t)m * nchunk_max + c) * nb + b] = acc;
}

__global__ void __launch_bounds__(256) k_reduce_w2(const double* __restrict__ s1, int nb,
                                                   const MockGeom* __restrict__ geom, int side, int nchunk_max,
                                                   double* __restrict__ out) {
    __shared__ double sh[4][64];
    const int m = blockIdx.y;
    const int bl = threadIdx.x & 63, stripe = threadIdx.x >> 6;
    const int b = blockIdx.x * 64 + bl;
    const int64_t ng = geom[m].ngroups[side];
    const int nchunks = (int)((ng + kChunkW - 1) / kChunkW);
    double acc = 0.0;
    if (b < nb)
        for (int c = stripe; c < nchunks; c += 4) acc += s1[((int64_t)m * nchunk_max + c) * nb + b];
    sh[stripe][bl] = acc;
    __syncthreads();
    if (stripe == 0 && b < nb) out[(int64_t)m * nb + b] = ((sh[0][bl] + sh[1][bl]) + sh[2][bl]) + sh[3][bl];
}

// inclusive 2-D prefix of each mock's [nr][np] sums (over pi, then over rp)
__global__ void __launch_bounds__(256) k_prefix_2d(double* __restrict__ out, int nr, int np) {
    double* o = out + (int64_t)blockIdx.x * nr * np;
    for (int k = threadIdx.x; k < nr; k += blockDim.x) {
        double run = 0.0;
        for (int q = 0; q < np; q++) { run += o[k * np + q]; o[k * np + q] = run; }
    }
    __threadfence_block();
    __syncthreads();
    for (int q = threadIdx.x; q < np; q += blockDim.x) {
        double run = 0.0;
        for (int k = 0; k < nr; k++) { run += o[k * np + q]; o[k * np + q] = run; }
    }
}

int64_t reduce_scratch_elems(int64_t max_groups_per_mock, int nmocks, int nb) {
    const int64_t g = max_groups_per_mock > 0 ? max_groups_per_mock : 1;
    const int64_t nc = nb > kWideCells ? nchunk_w(g) : nchunk_of(g);
    return nc * nmocks * (int64_t)nb;
}

// FP64 partial histograms [groups][nr] (2-D: [groups][nr * np]) -> cumulative sums per mock
int launch_reduce_sums_s(const double* partial, int nr, int np, const MockGeom* geom, int nmocks, int side,
                         int64_t max_groups, double* scratch, double* sums, hipStream_t s) {
    const int nb = np > 0 ? nr * np : nr;
    if (np > 0 && nb > kWideCells) {
        const int nc = (int)(nchunk_w(max_groups) > 0 ? nchunk_w(max_groups) : 1);
        hipLaunchKernelGGL(k_reduce_w1, dim3(nc, nmocks, (nb + 255) / 256), dim3(256), 0, s, partial, nb, geom, side,
                           nc, scratch);
        hipLaunchKernelGGL(k_reduce_w2, dim3((nb + 63) / 64, nmocks), dim3(256), 0, s, scratch, nb, geom, side, nc,
                           sums);
        hipLaunchKernelGGL(k_prefix_2d, dim3(nmocks), dim3(256), 0, s, sums, nr, np);
        return hipGetLastError() == hipSuccess ? 0 : -1;
    }
    const int nc = (int)(nchunk_of(max_groups) > 0 ? nchunk_of(max_groups) : 1);
    hipLaunchKernelGGL(k_reduce_s1<double>, dim3(nc, nmocks), dim3(256), 0, s, partial, nb, geom, side, nc, scratch);
    hipLaunchKernelGGL((k_reduce_s2<double, double>), dim3(nmocks), dim3(256), 0, s, scratch, nr, np, geom, side, nc,
                       sums);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_reduce_evals_s(const uint64_t* pevals, const MockGeom* geom, int nmocks, int side, int64_t max_groups,
                          uint64_t* scratch, int64_t* evals, hipStream_t s) {
    const int nc = (int)(nchunk_of(max_groups) > 0 ? nchunk_of(max_groups) : 1);
    hipLaunchKernelGGL(k_reduce_s1<uint64_t>, dim3(nc, nmocks), dim3(256), 0, s, pevals, 1, geom, side, nc, scratch);
    hipLaunchKernelGGL(k_reduce_plain, dim3(nmocks), dim3(64), 0, s, scratch, geom, side, nc, evals);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ------------------------------------------------------------- backward
constexpr int kDotMocks = 240;
struct DotArgs {
    const double* rows;
    const double* g;
    double* grad;
    int nr;
    int nm;           // mocks in this launch
    int m0;           // first mock of this launch (row of g)
    int64_t off[kDotMocks + 1];
};

// grad[p] = sum_s rows[p][s] * g[mock(p)][s]: one thread per point, the
// contraction of the per-point Jacobian rows with the suffix-summed upstream
// gradient.  HBM-bound (reads nr doubles per point).
__global__ void __launch_bounds__(256) k_rows_dot(const DotArgs a) {
    const int64_t p = a.off[0] + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= a.off[a.nm]) return;
    int lo = 0, hi = a.nm - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.off[mid] <= p) lo = mid; else hi = mid - 1;
    }
    const double* r = a.rows + p * a.nr;
    const double* g = a.g + (int64_t)(a.m0 + lo) * a.nr;
    double acc = 0.0;
    for (int s = 0; s < a.nr; s++) acc += r[s] * g[s];
    a.grad[p] = acc;
}

int launch_rows_dot(const double* rows, int nr, const double* gsuffix, const int64_t* offsets, int nmocks,
                    double* grad, hipStream_t s) {
    for (int m0 = 0; m0 < nmocks; m0 += kDotMocks) {
        DotArgs a;
        a.rows = rows;
        a.g = gsuffix;
        a.grad = grad;
        a.nr = nr;
        a.nm = nmocks - m0 < kDotMocks ? nmocks - m0 : kDotMocks;
        a.m0 = m0;
        for (int i = 0; i <= a.nm; i++) a.off[i] = offsets[m0 + i];
        const int64_t n = a.off[a.nm] - a.off[0];
        if (n <= 0) continue;
        hipLaunchKernelGGL(k_rows_dot, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ---------------------------------------------------------------- segment sums
// out[m] = sum of values[off[m] .. off[m+1]) for every mock m, in a fixed
// order: stage 1 sums chunks of kSegChunk values (one block each: strided
// per-thread sums, then a tree), stage 2 adds each mock's chunk sums in
// chunk order with one wavefront.  No atomics, so results are bitwise
// reproducible.
constexpr int kSegMocks = 120;
constexpr int kSegChunk = 4096;
struct SegArgs {
    const double* v;
    double* part;     // chunk sums of this launch
    double* out;
    int nm;           // mocks in this launch
    int m0;
    int64_t off[kSegMocks + 1];   // point offsets
    int64_t cc[kSegMocks + 1];    // cumulative chunk counts (cc[0] = 0)
};

__global__ void __launch_bounds__(256) k_seg_chunks(const SegArgs a) {
    __shared__ double t[256];
    const int64_t c = blockIdx.x;
    int lo = 0, hi = a.nm - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.cc[mid] <= c) lo = mid; else hi = mid - 1;
    }
    const int64_t b = a.off[lo] + (c - a.cc[lo]) * kSegChunk;
    const int64_t e = min(b + (int64_t)kSegChunk, a.off[lo + 1]);
    double acc = 0.0;
    for (int64_t i = b + threadIdx.x; i < e; i += 256) acc += a.v[i];
    t[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) t[threadIdx.x] += t[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.part[c] = t[0];
}

__global__ void __launch_bounds__(64) k_seg_mocks(const SegArgs a) {
    const int m = blockIdx.x;
    const int lane = threadIdx.x;
    double acc = 0.0;
    for (int64_t c = a.cc[m] + lane; c < a.cc[m + 1]; c += 64) acc += a.part[c];
    for (int w = 32; w > 0; w >>= 1) acc += __shfl_xor(acc, w, 64);
    if (lane == 0) a.out[a.m0 + m] = acc;
}

int64_t segment_sum_scratch(const int64_t* offsets, int nmocks) {
    int64_t need = 1;
    for (int m0 = 0; m0 < nmocks; m0 += kSegMocks) {
        const int nm = nmocks - m0 < kSegMocks ? nmocks - m0 : kSegMocks;
        int64_t chunks = 0;
        for (int i = 0; i < nm; i++) chunks += (offsets[m0 + i + 1] - offsets[m0 + i] + kSegChunk - 1) / kSegChunk;
        need = chunks > need ? chunks : need;
    }
    return need;
}

int launch_segment_sum(const double* v, const int64_t* offsets, int nmocks, double* out, double* scratch,
                       hipStream_t s) {
    for (int m0 = 0; m0 < nmocks; m0 += kSegMocks) {
        SegArgs a;
        a.v = v;
        a.part = scratch;
        a.out = out;
        a.nm = nmocks - m0 < kSegMocks ? nmocks - m0 : kSegMocks;
        a.m0 = m0;
        a.cc[0] = 0;
        for (int i = 0; i <= a.nm; i++) a.off[i] = offsets[m0 + i];
        for (int i = 0; i < a.nm; i++) a.cc[i + 1] = a.cc[i] + (a.off[i + 1] - a.off[i] + kSegChunk - 1) / kSegChunk;
        if (a.cc[a.nm] > 0)
            hipLaunchKernelGGL(k_seg_chunks, dim3((unsigned)a.cc[a.nm]), dim3(256), 0, s, a);
        hipLaunchKernelGGL(k_seg_mocks, dim3((unsigned)a.nm), dim3(64), 0, s, a);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

__global__ void k_add_rows(const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ out,
                           int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = a[i] + b[i];
}

int launch_add_rows(const double* a, const double* b, double* out, int64_t n, hipStream_t s) {
    if (n <= 0) return 0;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_add_rows, dim3((unsigned)blocks), dim3(256), 0, s, a, b, out, n);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ------------------------------------------------------------ jackknife
// (w, tag) rows for the pair kernel: the weight (1 when none is given) and the
// subsample tag as a double (exact below 2^53).  Tags outside [1, nsamp] set
// status bit 2; the pair kernel never indexes a table with them.
__global__ void k_pack_jack(const double* __restrict__ w, const int32_t* __restrict__ tags, int64_t n, int nsamp,
                            double* __restrict__ out, int32_t* __restrict__ status) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t t = tags[i];
        if ((t < 1 || t > nsamp) && status) atomicOr(status, 2);
        out[2 * i] = w ? w[i] : 1.0;
        out[2 * i + 1] = (double)t;
    }
}

int launch_pack_jack(const double* w, const int32_t* tags, int64_t n, int nsamp, double* out, int32_t* status,
                     hipStream_t s) {
    if (n <= 0) return 0;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_pack_jack, dim3((unsigned)blocks), dim3(256), 0, s, w, tags, n, nsamp, out, status);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Inclusive prefix of every (mock, subsample) table row: over bins (1-D) or
// over pi then rp (2-D, the order of k_reduce_s2).
__global__ void __launch_bounds__(256) k_jack_prefix(double* __restrict__ acc, int nr, int np) {
    double* o = acc + (int64_t)blockIdx.x * (np > 0 ? nr * np : nr);
    if (np <= 0) {
        if (threadIdx.x == 0) {
            double run = 0.0;
            for (int k = 0; k < nr; k++) { run += o[k]; o[k] = run; }
        }
        return;
    }
    for (int k = threadIdx.x; k < nr; k += blockDim.x) {
        double run = 0.0;
        for (int q = 0; q < np; q++) { run += o[k * np + q]; o[k * np + q] = run; }
    }
    __threadfence_block();
    __syncthreads();
    for (int q = threadIdx.x; q < np; q += blockDim.x) {
        double run = 0.0;
        for (int k = 0; k < nr; k++) { run += o[k * np + q]; o[k * np + q] = run; }
    }
}

// out[m][0] = all pairs; out[m][s] = all pairs - weight lost without subsample s.
__global__ void k_jack_out(const double* __restrict__ acc, int64_t nb, int nsamp, int64_t total,
                           double* __restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = i % nb;
        const int64_t row = i / nb;
        const int64_t sub = row % (nsamp + 1);
        const double h = acc[(row - sub) * nb + b];
        out[i] = sub == 0 ? h : h - acc[i];
    }
}

int launch_jack_finish(double* acc, int nr, int np, int nmocks, int nsamp, double* out, hipStream_t s) {
    const int64_t rows = (int64_t)nmocks * (nsamp + 1);
    const int64_t nb = np > 0 ? (int64_t)nr * np : nr;
    hipLaunchKernelGGL(k_jack_prefix, dim3((unsigned)rows), dim3(np > 0 ? 256 : 64), 0, s, acc, nr, np);
    const int64_t total = rows * nb;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_jack_out, dim3((unsigned)blocks), dim3(256), 0, s, acc, nb, nsamp, total, out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Zheng07 central occupation of each point and its derivative with respect to
// logMmin (empirical_models/occupation_models/zheng07_components.py:149-153):
// u = (log_mass - logMmin[mock]) / sigma, w = 0.5 * (1 + erf(u)),
// dw = -exp(-u * u) / (sqrt(pi) * sigma), written as (w, dw) rows: the weights
// and tangents of the batched forward-mode step, in one pass over the points
// (the divisions by the host constants are multiplications by their
// reciprocals, as torch evaluates a tensor divided by a Python float).
__global__ void k_zheng07_tangent(const double* __restrict__ logm, const int64_t* __restrict__ mock_id,
                                  const double* __restrict__ logMmin, double inv_sigma, double inv_dnorm, int64_t n,
                                  double2* __restrict__ wt) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double u = (logm[i] - logMmin[mock_id[i]]) * inv_sigma;
    const double w = 0.5 * (1.0 + erf(u));
    const double dw = -exp(-(u * u)) * inv_dnorm;
    wt[i] = make_double2(w, dw);
}

int launch_zheng07_tangent(const double* logm, const int64_t* mock_id, const double* logMmin, double sigma,
                           int64_t n, double* wt, hipStream_t s) {
    if (n <= 0) return 0;
    const double inv_sigma = 1.0 / sigma, inv_dnorm = 1.0 / (sqrt(M_PI) * sigma);
    hipLaunchKernelGGL(k_zheng07_tangent, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, logm, mock_id, logMmin,
                       inv_sigma, inv_dnorm, n, reinterpret_cast<double2*>(wt));
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace hp
