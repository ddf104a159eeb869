/*
 * cpu_ref.c -- CPU restatement of the halotools pair-counting engines.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle that the GPU path is
 * checked against (tests/, __graft_entry__.smoke) and, built with -Ofast, the
 * CPU baseline that bench.py times beside the GPU.  Nothing in the product
 * (halotools_amd/) may link, load or call it.
 *
 * Every engine below is a loop-for-loop restatement of the reference Cython
 * engine it names: same cell-1 / window / i / j loop nest, the same periodic
 * shift applied to the cell-1 point (x1tmp = x1 - shift), the same FP64
 * expression order dsq = dx*dx + dy*dy + dz*dz and the same top-down
 * "while dsq <= r2[k]" cumulative binning.  The mesh itself (idx_sorted,
 * cell_id_indices, divisions, covering steps) is computed by oracle/mesh.py,
 * which restates PC/rectangular_mesh.py; this file receives the already
 * gathered, cell-sorted coordinates exactly as the Cython engines do after
 * their `x1in[double_mesh.mesh1.idx_sorted]` copies.
 *
 * PC/ abbreviates halotools/mock_observables/pair_counters/.
 *
 * Build exact mode with -O2 -ffp-contract=off -fno-fast-math (parity oracle)
 * and baseline mode with -Ofast and no -march, as PC/cpairs/setup_package.py:27-29
 * compiles the reference (x86-64 baseline ISA: SSE2, no FMA).
 *
 * A test-only mutation switch (ref_set_mutation) plants known mistakes so the
 * test-suite can show that its golden checks would catch them.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* mutation switch (test infrastructure: prove the golden checks have teeth) */
/* 0: none; 1: strict '<' in the bin test; 2: periodic shift dropped;        */
/* 3: window one cell too narrow on the high side; 4: rbins not squared.     */
static int g_mutation = 0;
void ref_set_mutation(int m) { g_mutation = m; }

/* compensated summation switch (test infrastructure): marked sums accumulate
 * with Neumaier's compensated sum instead of the reference's plain serial sum,
 * giving a reference value accurate to ~1 ulp for large inputs, where any two
 * summation orders of ~1e9 terms (the reference's among them) differ by more
 * than 1e-12 relative.  Only the exact build honours it. */
static int g_compensated = 0;
void ref_set_compensated(int c) { g_compensated = c; }
int ref_get_mutation(void) { return g_mutation; }

/* Mesh description handed over by oracle/mesh.py.  Field meaning follows the
 * Cython engines' local variables (PC/cpairs/npairs_3d_engine.pyx:47-93). */
typedef struct ref_mesh {
    int64_t first_cell1, last_cell1;  /* cell1_tuple                          */
    int32_t n1div[3];                 /* mesh1.num_{x,y,z}divs                */
    int32_t n2div[3];                 /* mesh2.num_{x,y,z}divs                */
    int32_t ncover[3];                /* num_{x,y,z}2_covering_steps          */
    int32_t pbc;                      /* double_mesh._PBCs                    */
    double period[3];                 /* double_mesh.{x,y,z}period            */
    const int64_t* cell1_indices;     /* mesh1.cell_id_indices (ncell1+1)     */
    const int64_t* cell2_indices;     /* mesh2.cell_id_indices (ncell2+1)     */
} ref_mesh;

/* ---------------------------------------------------------------------- */
/* marking functions: PC/marked_cpairs/marking_functions.pyx:14-217        */
/* (ids 1-17) and custom_marking_func.pyx:12-16 (id 0).                    */
static double wf_product(const double* a, const double* b) { return a[0] * b[0]; }          /* 1  :14-21  */
static double wf_sum(const double* a, const double* b) { return a[0] + b[0]; }              /* 2  :24-31  */
static double wf_eq(const double* a, const double* b) { return a[0] == b[0] ? a[1] * b[1] : 0.0; }   /* 3 :34-45 */
static double wf_ineq(const double* a, const double* b) { return a[0] != b[0] ? a[1] * b[1] : 0.0; } /* 4 :48-59 */
static double wf_gt(const double* a, const double* b) { return b[0] > a[0] ? a[1] * b[1] : 0.0; }   /* 5 :62-73 */
static double wf_lt(const double* a, const double* b) { return b[0] < a[0] ? a[1] * b[1] : 0.0; }   /* 6 :76-87 */
static double wf_tgt(const double* a, const double* b) { return b[0] > (a[0] + a[1]) ? b[1] : 0.0; } /* 7 :90-101 */
/* id 8: the reference tests w2[0] < (w1[0] + w1[1]) although its docstring says '-' (:104-115) */
static double wf_tlt(const double* a, const double* b) { return b[0] < (a[0] + a[1]) ? b[1] : 0.0; }
static double wf_tol(const double* a, const double* b) { return fabs(a[0] - b[0]) < a[1] ? b[1] : 0.0; }  /* 9 :118-129 */
static double wf_excl(const double* a, const double* b) { return fabs(a[0] - b[0]) > a[1] ? b[1] : 0.0; } /* 10 :132-143 */
static double wf_ratio(const double* a, const double* b) { return b[0] > a[0] * a[1] ? b[1] : 0.0; }     /* 11 :146-157 */
static double wf_dot(const double* a, const double* b) {                                                  /* 12 :160-168 */
    return a[0] * b[0] * (a[1] * b[1] + a[2] * b[2] + a[3] * b[3]);
}
static double wf_sqdot(const double* a, const double* b) {                                                /* 13 :171-179 */
    return a[0] * b[0] * (a[1] * b[1] + a[2] * b[2] + a[3] * b[3]) * (a[1] * b[1] + a[2] * b[2] + a[3] * b[3]);
}
static double wf_projdot(const double* a, const double* b) {                                              /* 14 :181-189 */
    return a[0] * b[0] * (a[1] * b[1] + a[2] * b[2]);
}
static double wf_projsqdot(const double* a, const double* b) {                                            /* 15 :192-200 */
    return a[0] * b[0] * (a[1] * b[1] + a[2] * b[2]) * (a[1] * b[1] + a[2] * b[2]);
}
static double wf_sqdot_eq(const double* a, const double* b) {                                             /* 16 :202-213 */
    if (a[4] == b[4])
        return a[0] * b[0] * (a[1] * b[1] + a[2] * b[2] + a[3] * b[3]) * (a[1] * b[1] + a[2] * b[2] + a[3] * b[3]);
    return 0.0;
}
static double wf_sqdot_ineq(const double* a, const double* b) {                                           /* 17 :215-226 */
    if (a[4] != b[4])
        return a[0] * b[0] * (a[1] * b[1] + a[2] * b[2] + a[3] * b[3]) * (a[1] * b[1] + a[2] * b[2] + a[3] * b[3]);
    return 0.0;
}

typedef double (*wfunc_t)(const double*, const double*);

/* PC/marked_cpairs/marked_npairs_3d_engine.pyx:221-263 */
static wfunc_t wfunc_from_id(int id) {
    switch (id) {
        case 0: return wf_product; /* custom_func is w1[0]*w2[0] */
        case 1: return wf_product;
        case 2: return wf_sum;
        case 3: return wf_eq;
        case 4: return wf_ineq;
        case 5: return wf_gt;
        case 6: return wf_lt;
        case 7: return wf_tgt;
        case 8: return wf_tlt;
        case 9: return wf_tol;
        case 10: return wf_excl;
        case 11: return wf_ratio;
        case 12: return wf_dot;
        case 13: return wf_sqdot;
        case 14: return wf_projdot;
        case 15: return wf_projsqdot;
        case 16: return wf_sqdot_eq;
        case 17: return wf_sqdot_ineq;
        default: return NULL;
    }
}

double ref_wfunc(int id, const double* w1, const double* w2) {
    wfunc_t f = wfunc_from_id(id);
    return f ? f(w1, w2) : NAN;
}

/* ---------------------------------------------------------------------- */
/* The shared window walk: PC/cpairs/npairs_3d_engine.pyx:103-165.          */
/* The visitor is invoked for every (cell1, cell2) pair with the cell-2     */
/* shift values, exactly where the Cython engines enter their i/j loops.    */

enum { KIND_3D = 0, KIND_PROJ = 1, KIND_XYZ = 2, KIND_MARKED = 3, KIND_OBJ = 4, KIND_MARKED_XYZ = 5,
       KIND_JACK = 6, KIND_JACK_XYZ = 7 };

typedef struct ref_job {
    int kind;
    const ref_mesh* m;
    const double *x1, *y1, *z1, *x2, *y2, *z2; /* cell-sorted copies */
    const double *w1, *w2;                     /* cell-sorted rows (n, nw) */
    int nw;
    int wfunc_id;
    const int64_t *jt1, *jt2;                  /* jackknife: cell-sorted subvolume tags */
    int64_t nsamp;                             /* jackknife: N_samples */
    const double* r2;   /* squared radial / rp bins (K)   */
    int K;
    const double* pi2;  /* squared pi bins (G) for xy_z   */
    int G;
    double pimax2;      /* npairs_projected               */
    int64_t* counts;    /* int64 output (K) or (K*G)      */
    double* wcounts;    /* float64 output (K)             */
    double* wcomp;      /* compensation terms of wcounts (g_compensated) */
} ref_job;

/* s += x, optionally compensated (Neumaier 1974, "Rundungsfehleranalyse einiger
 * Verfahren zur Summation endlicher Summen"): the rounding error of each
 * addition is carried in c and added back at the end. */
static inline void acc_add(double* s, double* c, double x) {
    if (!c) {
        *s += x;
        return;
    }
    const double t = *s + x;
    if (fabs(*s) >= fabs(x)) *c += (*s - t) + x;
    else *c += (x - t) + *s;
    *s = t;
}

/* jackknife weight of a pair for subsample s (0 = whole sample):
 * PC/cpairs/npairs_jackknife_3d_engine.pyx:239-287 */
static double jweight(int64_t s, int64_t j1, int64_t j2, double w1, double w2) {
    double result = 0.0;
    if (s == 0) result = w1 * w2;
    else if ((j1 == j2) && (j1 == s)) result = 0.0;                  /* both outside */
    else if ((j1 != s) && (j2 != s)) result = (w1 * w2);             /* both inside */
    else if ((j1 != j2) && ((j1 == s) || (j2 == s))) result = 0.5 * (w1 * w2);
    return result;
}

static inline int bin_pass(double dsq, double r2k) {
    return g_mutation == 1 ? (dsq < r2k) : (dsq <= r2k);
}

static void run_cells(const ref_job* J, int64_t first, int64_t last) {
    const ref_mesh* m = J->m;
    const int nx1 = m->n1div[0], ny1 = m->n1div[1], nz1 = m->n1div[2];
    const int nx2 = m->n2div[0], ny2 = m->n2div[1], nz2 = m->n2div[2];
    const int px = nx2 / nx1, py = ny2 / ny1, pz = nz2 / nz1;   /* num_x2_per_x1 ... :91-93 */
    const int cx = m->ncover[0], cy = m->ncover[1];
    const int cz = m->ncover[2];
    const int hi_trim = (g_mutation == 3) ? 1 : 0;
    const double PBCs = (g_mutation == 2) ? 0.0 : (double)m->pbc;
    const int K = J->K, G = J->G;

    for (int64_t icell1 = first; icell1 < last; icell1++) {                   /* :103 */
        const int64_t ifirst1 = m->cell1_indices[icell1];
        const int64_t ilast1 = m->cell1_indices[icell1 + 1];
        const int64_t Ni = ilast1 - ifirst1;
        if (Ni <= 0) continue;                                                  /* :111 */
        const int ix1 = (int)(icell1 / ((int64_t)ny1 * nz1));                   /* :113-115 */
        const int iy1 = (int)((icell1 - (int64_t)ix1 * ny1 * nz1) / nz1);
        const int iz1 = (int)(icell1 - (int64_t)ix1 * ny1 * nz1 - (int64_t)iy1 * nz1);
        const int lx = ix1 * px - cx, ly = iy1 * py - cy, lz = iz1 * pz - cz;   /* :117-119 */
        const int rx = (ix1 + 1) * px + cx - hi_trim;                           /* :121-123 */
        const int ry = (iy1 + 1) * py + cy - hi_trim;
        const int rz = (iz1 + 1) * pz + cz - hi_trim;
        for (int ux = lx; ux < rx; ux++) {                                      /* :125 */
            double sx = ux < 0 ? -m->period[0] * PBCs : (ux >= nx2 ? +m->period[0] * PBCs : 0.0);
            int ix2 = ((ux % nx2) + nx2) % nx2;                                 /* Python-style modulo :133 */
            for (int uy = ly; uy < ry; uy++) {
                double sy = uy < 0 ? -m->period[1] * PBCs : (uy >= ny2 ? +m->period[1] * PBCs : 0.0);
                int iy2 = ((uy % ny2) + ny2) % ny2;
                for (int uz = lz; uz < rz; uz++) {
                    double sz = uz < 0 ? -m->period[2] * PBCs : (uz >= nz2 ? +m->period[2] * PBCs : 0.0);
                    int iz2 = ((uz % nz2) + nz2) % nz2;
                    const int64_t icell2 = (int64_t)ix2 * ny2 * nz2 + (int64_t)iy2 * nz2 + iz2;   /* :155 */
                    const int64_t ifirst2 = m->cell2_indices[icell2];
                    const int64_t ilast2 = m->cell2_indices[icell2 + 1];
                    const int64_t Nj = ilast2 - ifirst2;
                    if (Nj <= 0) continue;                                      /* :165 */
                    for (int64_t i = ifirst1; i < ilast1; i++) {                /* :166 */
                        const double x1tmp = J->x1[i] - sx;                     /* :167-169 */
                        const double y1tmp = J->y1[i] - sy;
                        const double z1tmp = J->z1[i] - sz;
                        for (int64_t j = ifirst2; j < ilast2; j++) {            /* :171 */
                            const double dx = x1tmp - J->x2[j];                 /* :173-175 */
                            const double dy = y1tmp - J->y2[j];
                            const double dz = z1tmp - J->z2[j];
                            if (J->kind == KIND_3D) {
                                const double dsq = dx * dx + dy * dy + dz * dz; /* :176 */
                                int k = K - 1;                                  /* :178-182 */
                                while (bin_pass(dsq, J->r2[k])) {
                                    J->counts[k] += 1;
                                    k = k - 1;
                                    if (k < 0) break;
                                }
                            } else if (J->kind == KIND_OBJ) {
                                /* PC/cpairs/npairs_per_object_3d_engine.pyx:183-199: the
                                 * counts of sorted sample-1 point i, row ifirst1 + i */
                                const double dsq = dx * dx + dy * dy + dz * dz;
                                int k = K - 1;
                                while (bin_pass(dsq, J->r2[k])) {
                                    J->counts[i * K + k] += 1;
                                    k = k - 1;
                                    if (k < 0) break;
                                }
                            } else if (J->kind == KIND_MARKED) {
                                /* PC/marked_cpairs/marked_npairs_3d_engine.pyx:207-216 */
                                const double dsq = dx * dx + dy * dy + dz * dz;
                                const double weight = wfunc_from_id(J->wfunc_id)(J->w1 + i * J->nw, J->w2 + j * J->nw);
                                int k = K - 1;
                                while (bin_pass(dsq, J->r2[k])) {
                                    acc_add(&J->wcounts[k], J->wcomp ? &J->wcomp[k] : NULL, weight);
                                    k = k - 1;
                                    if (k < 0) break;
                                }
                            } else if (J->kind == KIND_MARKED_XYZ) {
                                /* PC/marked_cpairs/marked_npairs_xy_z_engine.pyx:214-226 */
                                const double dxy_sq = dx * dx + dy * dy;
                                const double dz_sq = dz * dz;
                                const double weight = wfunc_from_id(J->wfunc_id)(J->w1 + i * J->nw, J->w2 + j * J->nw);
                                int k = K - 1;
                                while (bin_pass(dxy_sq, J->r2[k])) {
                                    int g = G - 1;
                                    while (bin_pass(dz_sq, J->pi2[g])) {
                                        acc_add(&J->wcounts[(int64_t)k * G + g],
                                                J->wcomp ? &J->wcomp[(int64_t)k * G + g] : NULL, weight);
                                        g = g - 1;
                                        if (g < 0) break;
                                    }
                                    k = k - 1;
                                    if (k < 0) break;
                                }
                            } else if (J->kind == KIND_JACK) {
                                /* PC/cpairs/npairs_jackknife_3d_engine.pyx:218-234 */
                                const double dsq = dx * dx + dy * dy + dz * dz;
                                const double w1 = J->w1[i], w2 = J->w2[j];
                                const int64_t j1 = J->jt1[i], j2 = J->jt2[j];
                                for (int64_t s = 0; s < J->nsamp + 1; s++) {
                                    int k = K - 1;
                                    while (bin_pass(dsq, J->r2[k])) {
                                        J->wcounts[s * K + k] += jweight(s, j1, j2, w1, w2);
                                        k = k - 1;
                                        if (k < 0) break;
                                    }
                                }
                            } else if (J->kind == KIND_JACK_XYZ) {
                                /* PC/cpairs/npairs_jackknife_xy_z_engine.pyx:222-243 */
                                const double dxy_sq = dx * dx + dy * dy;
                                const double dz_sq = dz * dz;
                                const double w1 = J->w1[i], w2 = J->w2[j];
                                const int64_t j1 = J->jt1[i], j2 = J->jt2[j];
                                for (int64_t s = 0; s < J->nsamp + 1; s++) {
                                    int k = K - 1;
                                    while (bin_pass(dxy_sq, J->r2[k])) {
                                        int g = G - 1;
                                        while (bin_pass(dz_sq, J->pi2[g])) {
                                            J->wcounts[(s * K + k) * G + g] += jweight(s, j1, j2, w1, w2);
                                            g = g - 1;
                                            if (g < 0) break;
                                        }
                                        k = k - 1;
                                        if (k < 0) break;
                                    }
                                }
                            } else if (J->kind == KIND_PROJ) {
                                /* PC/cpairs/npairs_projected_engine.pyx:182-190 */
                                const double dxy_sq = dx * dx + dy * dy;
                                const double dz_sq = dz * dz;
                                int k = K - 1;
                                while (bin_pass(dxy_sq, J->r2[k])) {
                                    if (bin_pass(dz_sq, J->pimax2)) J->counts[k] += 1;
                                    k = k - 1;
                                    if (k < 0) break;
                                }
                            } else { /* KIND_XYZ: PC/cpairs/npairs_xy_z_engine.pyx:183-194 */
                                const double dxy_sq = dx * dx + dy * dy;
                                const double dz_sq = dz * dz;
                                int k = K - 1;
                                while (bin_pass(dxy_sq, J->r2[k])) {
                                    int g = G - 1;
                                    while (bin_pass(dz_sq, J->pi2[g])) {
                                        J->counts[(int64_t)k * G + g] += 1;
                                        g = g - 1;
                                        if (g < 0) break;
                                    }
                                    k = k - 1;
                                    if (k < 0) break;
                                }
                            }
                        }
                    }
                }
            }
        }
    }
}

/* ---------------------------------------------------------------------- */
/* Threaded driver: contiguous chunks of the cell-1 range with np.array_split */
/* semantics (PC/mesh_helpers.py:183-221), one private output per worker,   */
/* summed in worker order like np.sum(np.array(result), axis=0)             */
/* (PC/npairs_3d.py:142-146).                                               */

typedef struct worker_arg {
    ref_job job;
    int64_t first, last;
} worker_arg;

static void* worker_main(void* p) {
    worker_arg* a = (worker_arg*)p;
    run_cells(&a->job, a->first, a->last);
    return NULL;
}

static int run_threads(const ref_job* J, int nthreads) {
    const ref_mesh* m = J->m;
    const int64_t first = m->first_cell1, last = m->last_cell1;
    const int64_t ncells = last - first;
    int64_t outlen = (J->kind == KIND_XYZ || J->kind == KIND_MARKED_XYZ) ? (int64_t)J->K * J->G : J->K;
    if (J->kind == KIND_JACK) outlen = (J->nsamp + 1) * J->K;
    if (J->kind == KIND_JACK_XYZ) outlen = (J->nsamp + 1) * J->K * J->G;
    const int fp_out = (J->kind == KIND_MARKED || J->kind == KIND_MARKED_XYZ || J->kind == KIND_JACK ||
                        J->kind == KIND_JACK_XYZ);
    /* per-object rows of different cell-1 chunks are disjoint: workers share
     * the output (the reference sums per-worker arrays that are zero elsewhere) */
    const int shared_out = (J->kind == KIND_OBJ);
    const int comp = g_compensated && (J->kind == KIND_MARKED || J->kind == KIND_MARKED_XYZ);
    if (nthreads <= 1 || ncells <= 1) {
        ref_job one = *J;
        one.wcomp = comp ? (double*)calloc((size_t)outlen, sizeof(double)) : NULL;
        run_cells(&one, first, last);
        if (comp) {
            for (int64_t k = 0; k < outlen; k++) J->wcounts[k] += one.wcomp[k];
            free(one.wcomp);
        }
        return 0;
    }
    if (nthreads > ncells) nthreads = (int)ncells;
    worker_arg* args = (worker_arg*)calloc((size_t)nthreads, sizeof(worker_arg));
    pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
    if (!args || !th) { free(args); free(th); return -1; }
    /* np.array_split: the first (ncells % n) chunks get one extra element */
    const int64_t base = ncells / nthreads, extra = ncells % nthreads;
    int64_t start = first;
    for (int t = 0; t < nthreads; t++) {
        const int64_t len = base + (t < extra ? 1 : 0);
        args[t].job = *J;
        if (fp_out) {
            args[t].job.wcounts = (double*)calloc((size_t)outlen, sizeof(double));
            args[t].job.wcomp = comp ? (double*)calloc((size_t)outlen, sizeof(double)) : NULL;
        } else if (!shared_out) {
            args[t].job.counts = (int64_t*)calloc((size_t)outlen, sizeof(int64_t));
        }
        args[t].first = start;
        args[t].last = start + len;
        start += len;
        pthread_create(&th[t], NULL, worker_main, &args[t]);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    double* ctot = comp ? (double*)calloc((size_t)outlen, sizeof(double)) : NULL;
    for (int t = 0; t < nthreads; t++) {
        if (shared_out) continue;
        for (int64_t k = 0; k < outlen; k++) {
            if (comp) {
                acc_add(&J->wcounts[k], &ctot[k], args[t].job.wcounts[k]);
                acc_add(&J->wcounts[k], &ctot[k], args[t].job.wcomp[k]);
            } else if (fp_out) {
                J->wcounts[k] += args[t].job.wcounts[k];
            } else {
                J->counts[k] += args[t].job.counts[k];
            }
        }
        free(args[t].job.counts);
        free(args[t].job.wcounts);
        free(args[t].job.wcomp);
    }
    if (comp) {
        for (int64_t k = 0; k < outlen; k++) J->wcounts[k] += ctot[k];
        free(ctot);
    }
    free(args);
    free(th);
    return 0;
}

/* rbins_squared = rbins*rbins (PC/cpairs/npairs_3d_engine.pyx:46) */
static double* square_bins(const double* b, int n) {
    double* r2 = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (int k = 0; k < n; k++) r2[k] = (g_mutation == 4) ? b[k] : b[k] * b[k];
    return r2;
}

/* ---------------------------------------------------------------------- */
/* public entry points (called from oracle/mesh.py through ctypes)          */

int ref_npairs_3d(const ref_mesh* m, const double* x1, const double* y1, const double* z1,
                  const double* x2, const double* y2, const double* z2,
                  const double* rbins, int K, int nthreads, int64_t* counts) {
    ref_job J;
    memset(&J, 0, sizeof(J));
    J.kind = KIND_3D; J.m = m;
    J.x1 = x1; J.y1 = y1; J.z1 = z1; J.x2 = x2; J.y2 = y2; J.z2 = z2;
    double* r2 = square_bins(rbins, K);
    J.r2 = r2; J.K = K; J.counts = counts;
    memset(counts, 0, sizeof(int64_t) * (size_t)K);
    int rc = run_threads(&J, nthreads);
    free(r2);
    return rc;
}

int ref_npairs_projected(const ref_mesh* m, const double* x1, const double* y1, const double* z1,
                         const double* x2, const double* y2, const double* z2,
                         const double* rp_bins, int K, double pi_max, int nthreads, int64_t* counts) {
    ref_job J;
    memset(&J, 0, sizeof(J));
    J.kind = KIND_PROJ; J.m = m;
    J.x1 = x1; J.y1 = y1; J.z1 = z1; J.x2 = x2; J.y2 = y2; J.z2 = z2;
    double* r2 = square_bins(rp_bins, K);
    J.r2 = r2; J.K = K; J.counts = counts;
    J.pimax2 = (g_mutation == 4) ? pi_max : pi_max * pi_max;   /* npairs_projected_engine.pyx:52 */
    memset(counts, 0, sizeof(int64_t) * (size_t)K);
    int rc = run_threads(&J, nthreads);
    free(r2);
    return rc;
}

int ref_npairs_xy_z(const ref_mesh* m, const double* x1, const double* y1, const double* z1,
                    const double* x2, const double* y2, const double* z2,
                    const double* rp_bins, int K, const double* pi_bins, int G, int nthreads,
                    int64_t* counts) {
    ref_job J;
    memset(&J, 0, sizeof(J));
    J.kind = KIND_XYZ; J.m = m;
    J.x1 = x1; J.y1 = y1; J.z1 = z1; J.x2 = x2; J.y2 = y2; J.z2 = z2;
    double* r2 = square_bins(rp_bins, K);
    double* p2 = square_bins(pi_bins, G);
    J.r2 = r2; J.K = K; J.pi2 = p2; J.G = G; J.counts = counts;
    memset(counts, 0, sizeof(int64_t) * (size_t)K * (size_t)G);
    int rc = run_threads(&J, nthreads);
    free(r2);
    free(p2);
    return rc;
}

/* counts: (n1, K) rows in the cell-sorted order of sample 1 */
int ref_npairs_per_object_3d(const ref_mesh* m, const double* x1, const double* y1, const double* z1,
                             const double* x2, const double* y2, const double* z2,
                             const double* rbins, int K, int64_t n1, int nthreads, int64_t* counts) {
    ref_job J;
    memset(&J, 0, sizeof(J));
    J.kind = KIND_OBJ; J.m = m;
    J.x1 = x1; J.y1 = y1; J.z1 = z1; J.x2 = x2; J.y2 = y2; J.z2 = z2;
    double* r2 = square_bins(rbins, K);
    J.r2 = r2; J.K = K; J.counts = counts;
    memset(counts, 0, sizeof(int64_t) * (size_t)K * (size_t)n1);
    int rc = run_threads(&J, nthreads);
    free(r2);
    return rc;
}

int ref_marked_npairs_3d(const ref_mesh* m, const double* x1, const double* y1, const double* z1,
                         const double* x2, const double* y2, const double* z2,
                         const double* w1, const double* w2, int nw, int wfunc_id,
                         const double* rbins, int K, int nthreads, double* wcounts) {
    if (!wfunc_from_id(wfunc_id)) return -2;
    ref_job J;
    memset(&J, 0, sizeof(J));
    J.kind = KIND_MARKED; J.m = m;
    J.x1 = x1; J.y1 = y1; J.z1 = z1; J.x2 = x2; J.y2 = y2; J.z2 = z2;
    J.w1 = w1; J.w2 = w2; J.nw = nw; J.wfunc_id = wfunc_id;
    double* r2 = square_bins(rbins, K);
    J.r2 = r2; J.K = K; J.wcounts = wcounts;
    memset(wcounts, 0, sizeof(double) * (size_t)K);
    int rc = run_threads(&J, nthreads);
    free(r2);
    return rc;
}

int ref_marked_npairs_xy_z(const ref_mesh* m, const double* x1, const double* y1, const double* z1,
                           const double* x2, const double* y2, const double* z2,
                           const double* w1, const double* w2, int nw, int wfunc_id,
                           const double* rp_bins, int K, const double* pi_bins, int G, int nthreads,
                           double* wcounts) {
    if (!wfunc_from_id(wfunc_id)) return -2;
    ref_job J;
    memset(&J, 0, sizeof(J));
    J.kind = KIND_MARKED_XYZ; J.m = m;
    J.x1 = x1; J.y1 = y1; J.z1 = z1; J.x2 = x2; J.y2 = y2; J.z2 = z2;
    J.w1 = w1; J.w2 = w2; J.nw = nw; J.wfunc_id = wfunc_id;
    double* r2 = square_bins(rp_bins, K);
    double* p2 = square_bins(pi_bins, G);
    J.r2 = r2; J.K = K; J.pi2 = p2; J.G = G; J.wcounts = wcounts;
    memset(wcounts, 0, sizeof(double) * (size_t)K * (size_t)G);
    int rc = run_threads(&J, nthreads);
    free(r2);
    free(p2);
    return rc;
}

/* npairs_jackknife_3d / _xy_z: wcounts is (N_samples+1, K) or (N_samples+1, K, G);
 * weights and tags are cell-sorted (PC/cpairs/npairs_jackknife_3d_engine.pyx:76-84) */
int ref_npairs_jackknife(const ref_mesh* m, const double* x1, const double* y1, const double* z1,
                         const double* x2, const double* y2, const double* z2,
                         const double* w1, const double* w2, const int64_t* jt1, const int64_t* jt2,
                         int64_t nsamp, const double* rp_bins, int K, const double* pi_bins, int G,
                         int nthreads, double* wcounts) {
    ref_job J;
    memset(&J, 0, sizeof(J));
    J.kind = pi_bins ? KIND_JACK_XYZ : KIND_JACK; J.m = m;
    J.x1 = x1; J.y1 = y1; J.z1 = z1; J.x2 = x2; J.y2 = y2; J.z2 = z2;
    J.w1 = w1; J.w2 = w2; J.nw = 1; J.jt1 = jt1; J.jt2 = jt2; J.nsamp = nsamp;
    double* r2 = square_bins(rp_bins, K);
    double* p2 = pi_bins ? square_bins(pi_bins, G) : NULL;
    J.r2 = r2; J.K = K; J.pi2 = p2; J.G = pi_bins ? G : 1; J.wcounts = wcounts;
    memset(wcounts, 0, sizeof(double) * (size_t)(nsamp + 1) * (size_t)K * (size_t)J.G);
    int rc = run_threads(&J, nthreads);
    free(r2);
    free(p2);
    return rc;
}

/* Candidate-pair count of the reference mesh: sum over cell1 of
 * n1(cell1) * sum over its window of n2(cell2).  Used to express the GPU's
 * throughput in the reference engine's own units of work. */
int64_t ref_candidate_pairs(const ref_mesh* m) {
    const int nx1 = m->n1div[0], ny1 = m->n1div[1], nz1 = m->n1div[2];
    const int nx2 = m->n2div[0], ny2 = m->n2div[1], nz2 = m->n2div[2];
    const int px = nx2 / nx1, py = ny2 / ny1, pz = nz2 / nz1;
    const int cx = m->ncover[0], cy = m->ncover[1], cz = m->ncover[2];
    int64_t total = 0;
    for (int64_t icell1 = m->first_cell1; icell1 < m->last_cell1; icell1++) {
        const int64_t Ni = m->cell1_indices[icell1 + 1] - m->cell1_indices[icell1];
        if (Ni <= 0) continue;
        const int ix1 = (int)(icell1 / ((int64_t)ny1 * nz1));
        const int iy1 = (int)((icell1 - (int64_t)ix1 * ny1 * nz1) / nz1);
        const int iz1 = (int)(icell1 - (int64_t)ix1 * ny1 * nz1 - (int64_t)iy1 * nz1);
        int64_t nj = 0;
        for (int ux = ix1 * px - cx; ux < (ix1 + 1) * px + cx; ux++) {
            const int ix2 = ((ux % nx2) + nx2) % nx2;
            for (int uy = iy1 * py - cy; uy < (iy1 + 1) * py + cy; uy++) {
                const int iy2 = ((uy % ny2) + ny2) % ny2;
                for (int uz = iz1 * pz - cz; uz < (iz1 + 1) * pz + cz; uz++) {
                    const int iz2 = ((uz % nz2) + nz2) % nz2;
                    const int64_t c2 = (int64_t)ix2 * ny2 * nz2 + (int64_t)iy2 * nz2 + iz2;
                    nj += m->cell2_indices[c2 + 1] - m->cell2_indices[c2];
                }
            }
        }
        total += Ni * nj;
    }
    return total;
}
