/*
 * scrimp_oracle.c -- CPU restatement of the reference SCRIMP self-join.
 * TEST INFRASTRUCTURE ONLY (see scrimp_oracle.h).  Compiled with
 * -ffp-contract=off: every fused multiply-add below is an explicit fma().
 */
#define _GNU_SOURCE
#include "scrimp_oracle.h"

#include <immintrin.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static int g_fault = 0;

void orc_set_fault(int code) { g_fault = code; }

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ScrimpSequOpt.cpp:33-79 */
void orc_window_stats(const double* A, long n, int m, double* sInv, double* mus) {
    const long P = n - m + 1;
    if (P <= 0) return;
    double* cs = (double*)malloc(sizeof(double) * n);
    double* sq = (double*)malloc(sizeof(double) * n);
    cs[0] = A[0];
    for (long i = 1; i < n; ++i) cs[i] = A[i] + cs[i - 1];
    sq[0] = A[0] * A[0];
    for (long i = 1; i < n; ++i) sq[i] = fma(A[i], A[i], sq[i - 1]);
    const double sqrt_m = sqrt((double)m);
    for (long k = 0; k < P; ++k) {
        const double sum = (k == 0) ? cs[m - 1] : cs[m - 1 + k] - cs[k - 1];
        const double ssq = (k == 0) ? sq[m - 1] : sq[m - 1 + k] - sq[k - 1];
        const double mean = sum / (double)m;
        double var = fma(-mean, mean, ssq / (double)m);
        if (g_fault == 5 && m > 1) var = var * (double)m / (double)(m - 1);
        const double si = 1.0 / sqrt(var);
        sInv[k] = si;
        mus[k] = (mean * si) * sqrt_m;
    }
    free(cs);
    free(sq);
}

/* ScrimpSequOpt.cpp:81-96 */
void orc_seeds(const double* A, long n, int m, long first, long last, double* qt) {
    (void)n;
    for (long d = first; d <= last; ++d) {
        double z = 0.0;
        for (int k = 0; k < m; ++k) z = fma(A[k + d], A[k], z);
        qt[d] = z;
    }
}

/* kernels.cpp:310-366 */
void orc_eval_block_scalar(double prof_colmin[], long idx_colmin[], double prof_rowmin[],
                           long idx_rowmin[], double tmpQ[], long blocklen, long trianglen,
                           const double A_hor[], const double A_vert[], long w,
                           const double s_hor[], const double mu_hor[], const double s_vert[],
                           const double mu_vert[], long baserow, long basecol) {
    for (long j = 0; j < trianglen; ++j) {
        double profile_j = prof_colmin[j];
        long index_j = idx_colmin[j];
        const long lim = j + (blocklen < trianglen - j ? blocklen : trianglen - j);
        for (long i = j; i < lim; ++i) {
            const long diag = i - j;
            const double mprod = (g_fault == 2) ? -(mu_vert[i] * mu_hor[j]) : mu_vert[i] * mu_hor[j];
            const double score = fma(tmpQ[diag], s_vert[i] * s_hor[j], -mprod);
            tmpQ[diag] = fma(A_vert[i + w], A_hor[j + w], tmpQ[diag]);
            if (g_fault != 1) tmpQ[diag] = fma(-A_vert[i], A_hor[j], tmpQ[diag]);
            if (score > prof_rowmin[i]) {
                prof_rowmin[i] = score;
                idx_rowmin[i] = j + basecol;
            }
            if (score > profile_j) {
                profile_j = score;
                index_j = (g_fault == 3) ? j + basecol : i + baserow;
            }
        }
        if (profile_j > prof_colmin[j]) {
            prof_colmin[j] = profile_j;
            idx_colmin[j] = index_j;
        }
    }
}

/*
 * kernels.cpp:169-306.  Lanes cover 4 adjacent diagonals at one column; the
 * row maximum is a masked store per lane, the column maximum a per-lane
 * running max reduced at the end of the column (lower lanes win ties), then a
 * scalar remainder loop over the last diagonals.
 */
__attribute__((target("avx2,fma")))
void orc_eval_block_abba(double prof_colmin[], long idx_colmin[], double prof_rowmin[],
                         long idx_rowmin[], double tmpQ[], long blocklen, long trianglen,
                         const double A_hor[], const double A_vert[], long w,
                         const double s_hor[], const double mu_hor[], const double s_vert[],
                         const double mu_vert[], long baserow, long basecol) {
    if (g_fault != 0) {
        orc_eval_block_scalar(prof_colmin, idx_colmin, prof_rowmin, idx_rowmin, tmpQ, blocklen,
                              trianglen, A_hor, A_vert, w, s_hor, mu_hor, s_vert, mu_vert, baserow,
                              basecol);
        return;
    }
    for (long j = 0; j < trianglen; ++j) {
        const double muh = mu_hor[j], sh = s_hor[j], ah = A_hor[j], ahf = A_hor[j + w];
        const __m256d v_muh = _mm256_set1_pd(muh), v_sh = _mm256_set1_pd(sh);
        const __m256d v_ah = _mm256_set1_pd(ah), v_ahf = _mm256_set1_pd(ahf);
        __m256d colv = _mm256_set1_pd(prof_colmin[j]);
        __m256i coli = _mm256_set1_epi64x(-1);
        const long lim = j + (blocklen < trianglen - j ? blocklen : trianglen - j);
        const __m256i col_in_mat = _mm256_set1_epi64x(j + basecol);
        long i = j;
        for (; i + 4 <= lim; i += 4) {
            const long vd = i - j;
            const __m256d sv = _mm256_loadu_pd(s_vert + i);
            const __m256d mv = _mm256_loadu_pd(mu_vert + i);
            __m256d q = _mm256_loadu_pd(tmpQ + vd);
            const __m256d corr = _mm256_mul_pd(sv, v_sh);
            const __m256d mprod = _mm256_mul_pd(v_muh, mv);
            const __m256d score = _mm256_fmsub_pd(q, corr, mprod);
            q = _mm256_fmadd_pd(_mm256_loadu_pd(A_vert + w + i), v_ahf, q);
            q = _mm256_fnmadd_pd(_mm256_loadu_pd(A_vert + i), v_ah, q);
            _mm256_storeu_pd(tmpQ + vd, q);
            const __m256d rowv = _mm256_loadu_pd(prof_rowmin + i);
            const __m256d rmask = _mm256_cmp_pd(score, rowv, _CMP_GT_OS);
            _mm256_maskstore_pd(prof_rowmin + i, _mm256_castpd_si256(rmask), score);
            _mm256_maskstore_epi64((long long*)(idx_rowmin + i), _mm256_castpd_si256(rmask), col_in_mat);
            const __m256d cmask = _mm256_cmp_pd(score, colv, _CMP_GT_OS);
            colv = _mm256_blendv_pd(colv, score, cmask);
            const __m256i lanei = _mm256_add_epi64(_mm256_set1_epi64x(i + baserow),
                                                   _mm256_set_epi64x(3, 2, 1, 0));
            coli = _mm256_castpd_si256(_mm256_blendv_pd(_mm256_castsi256_pd(coli),
                                                        _mm256_castsi256_pd(lanei), cmask));
        }
        /* lane reduction: lower lanes win ties (kernels.cpp:255-271) */
        double pv[4];
        long long pi[4];
        _mm256_storeu_pd(pv, colv);
        _mm256_storeu_si256((__m256i*)pi, coli);
        if (pv[1] > pv[0]) { pv[0] = pv[1]; pi[0] = pi[1]; }
        if (pv[3] > pv[2]) { pv[2] = pv[3]; pi[2] = pi[3]; }
        if (pv[2] > pv[0]) { pv[0] = pv[2]; pi[0] = pi[2]; }
        double profile_j = prof_colmin[j];
        long index_j = idx_colmin[j];
        if (pv[0] > profile_j) {
            profile_j = pv[0];
            index_j = (long)pi[0];
        }
        for (; i < lim; ++i) {
            const long diag = i - j;
            const double score = fma(tmpQ[diag], s_vert[i] * sh, -(mu_vert[i] * muh));
            tmpQ[diag] = fma(A_vert[i + w], ahf, tmpQ[diag]);
            tmpQ[diag] = fma(-A_vert[i], ah, tmpQ[diag]);
            if (score > prof_rowmin[i]) {
                prof_rowmin[i] = score;
                idx_rowmin[i] = j + basecol;
            }
            if (score > profile_j) {
                profile_j = score;
                index_j = i + baserow;
            }
        }
        if (profile_j > prof_colmin[j]) {
            prof_colmin[j] = profile_j;
            idx_colmin[j] = index_j;
        }
    }
}

/* partitioning_1d.cpp:8-48 */
void orc_get_partition(int part_id, int num_partitions, int ids_to_process, int* first, int* last) {
    const int nfull = ids_to_process - (ids_to_process / num_partitions) * num_partitions;
    const int sfull = ids_to_process / num_partitions + 1;
    const int overhead = part_id - nfull;
    int f = (overhead > 0) ? part_id * sfull - overhead : part_id * sfull;
    int l = (part_id < nfull) ? f + sfull - 1 : f + sfull - 2;
    *first = f;
    *last = l;
}

typedef void (*orc_kernel_fn)(double*, long*, double*, long*, double*, long, long, const double*,
                              const double*, long, const double*, const double*, const double*,
                              const double*, long, long);

typedef struct {
    const double* A;
    long n, P, excl;
    int m, blocklen, rank, nranks, block_stride;
    orc_kernel_fn kern;
    double* profile;
    long* index;
    double seconds, cells;
} rank_job;

static void run_rank(rank_job* job) {
    const double t0 = now_s();
    const long P = job->P, n = job->n;
    const int m = job->m;
    for (long i = 0; i < P; ++i) {
        job->profile[i] = -1.0;
        job->index[i] = -1;
    }
    double* sInv = (double*)malloc(sizeof(double) * P);
    double* mus = (double*)malloc(sizeof(double) * P);
    double* qt = (double*)malloc(sizeof(double) * (n + 1));
    orc_window_stats(job->A, n, m, sInv, mus);
    const long excl = job->excl;
    const int W = (int)(P - excl - 1);
    const int nparts = 2 * job->nranks;
    int parts[2][2];
    for (int h = 0; h < 2; ++h) {
        const int pid = (h == 0) ? job->rank : nparts - job->rank - 1;
        int f, l;
        orc_get_partition(pid, nparts, W > 0 ? W : 0, &f, &l);
        parts[h][0] = (int)(excl + 1 + f);
        parts[h][1] = (int)((P - 1) < excl + 1 + l ? (P - 1) : excl + 1 + l);
    }
    double cells = 0.0;
    long blockno = 0;
    for (int h = 0; h < 2; ++h) {
        const int first = parts[h][0], last = parts[h][1];
        for (int prev = first - 1; prev < last;) {
            const int bf = prev + 1;
            const int bl = (bf + job->blocklen - 1 < last) ? bf + job->blocklen - 1 : last;
            prev = bl;
            if (job->block_stride > 1 && (blockno++ % job->block_stride) != 0) continue;
            orc_seeds(job->A, n, m, bf, bl, qt);
            const long blen = bl + 1 - bf;
            job->kern(job->profile, job->index, job->profile + bf, job->index + bf, qt + bf, blen,
                      P - bf, job->A, job->A + bf, m, sInv, mus, sInv + bf, mus + bf, bf, 0);
            for (long d = bf; d <= bl; ++d) cells += (double)(P - d);
        }
    }
    const double twice_m = 2.0 * (double)m;
    for (long i = 0; i < P; ++i) job->profile[i] = twice_m - 2.0 * job->profile[i];
    free(sInv);
    free(mus);
    free(qt);
    job->cells = cells;
    job->seconds = now_s() - t0;
}

int orc_triv_par_rank(const double* A, long n, int m, long excl, int blocklen, int rank, int nranks,
                      int kernel, int block_stride, double* profile, long* index, double* cells,
                      double* seconds) {
    const long P = n - m + 1;
    if (P <= 0 || m < 1 || rank < 0 || rank >= nranks) return -1;
    if (excl < 0) excl = m / 4;
    if (g_fault == 4) excl = excl - 1;
    rank_job job;
    memset(&job, 0, sizeof job);
    job.A = A;
    job.n = n;
    job.P = P;
    job.excl = excl;
    job.m = m;
    job.blocklen = blocklen;
    job.rank = rank;
    job.nranks = nranks;
    job.block_stride = block_stride;
    job.kern = kernel ? orc_eval_block_abba : orc_eval_block_scalar;
    job.profile = profile;
    job.index = index;
    run_rank(&job);
    if (cells) *cells = job.cells;
    if (seconds) *seconds = job.seconds;
    return 0;
}

typedef struct {
    rank_job* jobs;
    int first_rank, stride, nranks;
} thread_arg;

static void* thread_main(void* p) {
    thread_arg* ta = (thread_arg*)p;
    for (int r = ta->first_rank; r < ta->nranks; r += ta->stride) run_rank(&ta->jobs[r]);
    return NULL;
}

int orc_triv_par(const double* A, long n, int m, long excl, int blocklen, int nranks,
                 int nthreads, int kernel, int block_stride, double* profile, long* index,
                 double* cells, double* seconds) {
    const long P = n - m + 1;
    if (P <= 0 || m < 1) return -1;
    if (excl < 0) excl = m / 4;
    if (nranks < 1) nranks = 1;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > nranks) nthreads = nranks;
    if (g_fault == 4) excl = excl - 1;
    rank_job* jobs = (rank_job*)calloc((size_t)nranks, sizeof(rank_job));
    for (int r = 0; r < nranks; ++r) {
        jobs[r].A = A;
        jobs[r].n = n;
        jobs[r].P = P;
        jobs[r].excl = excl;
        jobs[r].m = m;
        jobs[r].blocklen = blocklen;
        jobs[r].rank = r;
        jobs[r].nranks = nranks;
        jobs[r].block_stride = block_stride;
        jobs[r].kern = kernel ? orc_eval_block_abba : orc_eval_block_scalar;
        if (r == 0) {
            jobs[r].profile = profile;
            jobs[r].index = index;
        } else {
            jobs[r].profile = (double*)malloc(sizeof(double) * P);
            jobs[r].index = (long*)malloc(sizeof(long) * P);
        }
    }
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
    thread_arg* ta = (thread_arg*)malloc(sizeof(thread_arg) * nthreads);
    for (int t = 0; t < nthreads; ++t) {
        ta[t].jobs = jobs;
        ta[t].first_rank = t;
        ta[t].stride = nthreads;
        ta[t].nranks = nranks;
        pthread_create(&th[t], NULL, thread_main, &ta[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    /* ScrimpTrivialPar.cpp:68-108: inout = rank 0, take in when strictly smaller */
    const double tr0 = now_s();
    for (int r = 1; r < nranks; ++r) {
        for (long i = 0; i < P; ++i) {
            if (jobs[r].profile[i] < profile[i]) {
                profile[i] = jobs[r].profile[i];
                index[i] = jobs[r].index[i];
            }
        }
    }
    const double tred = now_s() - tr0;
    double tmax = 0.0, csum = 0.0;
    for (int r = 0; r < nranks; ++r) {
        if (jobs[r].seconds > tmax) tmax = jobs[r].seconds;
        csum += jobs[r].cells;
        if (r > 0) {
            free(jobs[r].profile);
            free(jobs[r].index);
        }
    }
    if (cells) *cells = csum;
    if (seconds) *seconds = tmax + tred;
    free(th);
    free(ta);
    free(jobs);
    return 0;
}

/* ScrimpVertBlocked.cpp:89-186 without the shuffle */
int orc_vert_blocked(const double* A, long n, int m, long excl, int blocklen, int kernel,
                     double* profile, long* index) {
    const long P = n - m + 1;
    if (P <= 0 || m < 1) return -1;
    if (excl < 0) excl = m / 4;
    for (long i = 0; i < P; ++i) {
        profile[i] = -1.0;
        index[i] = 0;
    }
    double* sInv = (double*)malloc(sizeof(double) * P);
    double* mus = (double*)malloc(sizeof(double) * P);
    double* qt = (double*)malloc(sizeof(double) * (n + 1));
    orc_window_stats(A, n, m, sInv, mus);
    orc_kernel_fn kern = kernel ? orc_eval_block_abba : orc_eval_block_scalar;
    for (long b = excl + 1; b < P; b += blocklen) {
        const long last = (b + blocklen - 1 < P - 1) ? b + blocklen - 1 : P - 1;
        orc_seeds(A, n, m, b, last, qt);
        kern(profile, index, profile + b, index + b, qt + b, blocklen, P - b, A, A + b, m, sInv,
             mus, sInv + b, mus + b, b, 0);
    }
    const double twice_m = 2.0 * (double)m;
    for (long i = 0; i < P; ++i) profile[i] = twice_m - 2.0 * profile[i];
    free(sInv);
    free(mus);
    free(qt);
    return 0;
}

void orc_window_moments_ld(const double* A, long n, int m, double* mean, double* sigma) {
    const long P = n - m + 1;
    for (long i = 0; i < P; ++i) {
        long double s = 0.0L;
        for (int t = 0; t < m; ++t) s += (long double)A[i + t];
        const long double mu = s / (long double)m;
        long double v = 0.0L;
        for (int t = 0; t < m; ++t) {
            const long double d = (long double)A[i + t] - mu;
            v += d * d;
        }
        mean[i] = (double)mu;
        sigma[i] = (double)sqrtl(v / (long double)m);
    }
}

double orc_pair_dist2(const double* A, long n, int m, long i, long j) {
    (void)n;
    long double si = 0.0L, sj = 0.0L;
    for (int t = 0; t < m; ++t) {
        si += (long double)A[i + t];
        sj += (long double)A[j + t];
    }
    const long double mi = si / m, mj = sj / m;
    long double vi = 0.0L, vj = 0.0L, c = 0.0L;
    for (int t = 0; t < m; ++t) {
        const long double a = (long double)A[i + t] - mi, b = (long double)A[j + t] - mj;
        vi += a * a;
        vj += b * b;
        c += a * b;
    }
    const long double rho = c / sqrtl(vi * vj);
    return (double)(2.0L * (long double)m * (1.0L - rho));
}

void orc_brute_force(const double* A, long n, int m, long excl, double* profile, long* index) {
    const long P = n - m + 1;
    if (P <= 0) return;
    if (excl < 0) excl = m / 4;
    long double* cen = (long double*)malloc(sizeof(long double) * (size_t)P * (size_t)m);
    long double* nrm = (long double*)malloc(sizeof(long double) * (size_t)P);
    for (long i = 0; i < P; ++i) {
        long double s = 0.0L;
        for (int t = 0; t < m; ++t) s += (long double)A[i + t];
        const long double mu = s / m;
        long double v = 0.0L;
        for (int t = 0; t < m; ++t) {
            const long double d = (long double)A[i + t] - mu;
            cen[(size_t)i * m + t] = d;
            v += d * d;
        }
        nrm[i] = sqrtl(v);
    }
    for (long i = 0; i < P; ++i) {
        profile[i] = INFINITY;
        index[i] = -1;
    }
    for (long i = 0; i < P; ++i) {
        for (long j = i + excl + 1; j < P; ++j) {
            if (nrm[i] == 0.0L || nrm[j] == 0.0L) continue;
            long double c = 0.0L;
            const long double* a = cen + (size_t)i * m;
            const long double* b = cen + (size_t)j * m;
            for (int t = 0; t < m; ++t) c += a[t] * b[t];
            const double d2 = (double)(2.0L * m * (1.0L - c / (nrm[i] * nrm[j])));
            if (d2 < profile[i]) { profile[i] = d2; index[i] = j; }
            if (d2 < profile[j]) { profile[j] = d2; index[j] = i; }
        }
    }
    free(cen);
    free(nrm);
}

/* ScrimpppOrig.cpp:79-283, PreSCRIMP phase (arithmetic in scrimp_oracle.h) */
void orc_prescrimp(const double* A, long n, int m, double stride, const long* order, long norder,
                   double* profile, long* index) {
    const long P = n - m + 1;
    if (P <= 0) return;
    const long excl = m / 4;
    long step = (long)floor(stride * (double)m);
    if (step < 1) step = 1;
    double* cs = (double*)malloc(sizeof(double) * n);
    double* sq = (double*)malloc(sizeof(double) * n);
    double* mean = (double*)malloc(sizeof(double) * P);
    double* sig = (double*)malloc(sizeof(double) * P);
    /* ScrimpppOrig.cpp:84-113 */
    cs[0] = A[0];
    sq[0] = A[0] * A[0];
    for (long i = 1; i < n; ++i) {
        cs[i] = A[i] + cs[i - 1];
        sq[i] = fma(A[i], A[i], sq[i - 1]);
    }
    for (long k = 0; k < P; ++k) {
        const double sum = (k == 0) ? cs[m - 1] : cs[m - 1 + k] - cs[k - 1];
        const double sumsq = (k == 0) ? sq[m - 1] : sq[m - 1 + k] - sq[k - 1];
        mean[k] = sum / m;
        sig[k] = sqrt(fma(-mean[k], mean[k], sumsq / m));
    }
    for (long i = 0; i < P; ++i) {
        profile[i] = INFINITY;
        index[i] = 0;
    }
    const double dm = (double)m;
    long nrows = (P + step - 1) / step;
    for (long r = 0; r < (order ? norder : nrows); ++r) {
        const long i = order ? order[r] : r * step;
        double best = INFINITY;
        long bestj = 0;
        for (long j = 0; j < P; ++j) {
            double d;
            if (j > i - excl && j < i + excl) {
                d = INFINITY;
            } else {
                double qt = 0.0;
                for (int t = 0; t < m; ++t) qt = fma(A[j + t], A[i + t], qt);
                const double mp = (g_fault == 2) ? (dm * mean[j]) : -(dm * mean[j]);
                d = 2.0 * (dm - fma(mp, mean[i], qt) / (sig[j] * sig[i]));
            }
            if (d < best) {
                best = d;
                bestj = j;
            }
            if (d < profile[j]) {
                profile[j] = d;
                index[j] = i;
            }
        }
        profile[i] = best;
        index[i] = bestj;
        if (!(best < INFINITY)) continue;
        const long j = bestj;
        const double z0 = fma(dm - profile[i] / 2.0, sig[j] * sig[i], (dm * mean[j]) * mean[i]);
        double z = z0;
        for (long k = 1; k < step && i + k < P && j + k < P; ++k) {
            z = (g_fault == 1) ? fma(A[i + k + m - 1], A[j + k + m - 1], z)
                               : fma(A[i + k + m - 1], A[j + k + m - 1], fma(-A[i + k - 1], A[j + k - 1], z));
            const double d = 2.0 * (dm - fma(-(dm * mean[j + k]), mean[i + k], z) / (sig[j + k] * sig[i + k]));
            if (d < profile[i + k]) {
                profile[i + k] = d;
                index[i + k] = j + k;
            }
            if (d < profile[j + k]) {
                profile[j + k] = d;
                index[j + k] = i + k;
            }
        }
        z = z0;
        for (long k = 1; k < step && i - k >= 0 && j - k >= 0; ++k) {
            z = fma(A[i - k], A[j - k], fma(-A[i - k + m], A[j - k + m], z));
            const double d = 2.0 * (dm - fma(-(dm * mean[j - k]), mean[i - k], z) / (sig[j - k] * sig[i - k]));
            if (d < profile[i - k]) {
                profile[i - k] = d;
                index[i - k] = j - k;
            }
            if (d < profile[j - k]) {
                profile[j - k] = d;
                index[j - k] = i - k;
            }
        }
    }
    free(cs);
    free(sq);
    free(mean);
    free(sig);
}

/* ---- distrib_par: the k x k checkerboard of triangular tiles ---------------- */

/* checkerboard_partitioning.cpp:8-41: row r holds ranks r^2 .. r^2+2r; column
 * (rank - r^2)/2; upper when the rank's parity differs from the diagonal's. */
void orc_partition_coords(long rank, long* row, long* col, int* is_upper) {
    long r = 0, first = 0, next = 1;
    while (next <= rank) {
        r += 1;
        first = r * r;
        next = first + 2 * r + 1;
    }
    *row = r;
    *col = (rank - first) / 2;
    *is_upper = (r % 2 == 0) ? (rank % 2 == 1) : (rank % 2 == 0);
}

/* ScrimpDistribPar.cpp:580-596: the series is padded so that the triangle's
 * base W = n + pad - m - ez splits into k tiles of L diagonals. */
int orc_distrib_geometry(long n, int m, int k, long* pad, long* L) {
    const long ez = m / 4;
    if (k < 1 || m < 1) return -1;
    const long work = n - m - ez;
    const long floored = work / k;
    const long delta = work - k * floored;
    const long p = (delta > 0) ? k - delta : 0;
    const long PL = n + p - m + 1;
    const long W = PL - ez - 1;
    if (PL <= 0 || W < k) return -1; /* ScrimpDistribPar.cpp:605-606 */
    *pad = p;
    *L = W / k;
    return 0;
}

/* slice[x] = series[off + x] inside the series, 0 past its end
 * (ScrimpDistribPar.cpp:724-767 with DISTRIB_BCAST_INPUT, settings.h:34) */
static void distrib_slice(const double* series, long n, long off, long len, double* out) {
    for (long x = 0; x < len; ++x) out[x] = (off + x < n) ? series[off + x] : 0.0;
}

int orc_distrib_tile_inputs(const double* series, long n, int m, int k, long rank, double* A, double* B,
                            double* sA, double* muA, double* sB, double* muB, double* dots, long* row_offset,
                            long* col_offset, int* is_upper) {
    long pad, L, row, col;
    if (orc_distrib_geometry(n, m, k, &pad, &L) != 0 || rank < 0 || rank >= (long)k * k) return -1;
    const long ez = m / 4;
    orc_partition_coords(rank, &row, &col, is_upper);
    *row_offset = row * L + ez + 1; /* ScrimpDistribPar.cpp:666-668 */
    *col_offset = col * L;
    const long slen = L + m - 1; /* ts_slice_len */
    distrib_slice(series, n, *row_offset, slen, A);
    distrib_slice(series, n, *col_offset, slen, B);
    /* statistics of each slice on its own (ScrimpDistribPar.cpp:780-781) */
    orc_window_stats(A, slen, m, sA, muA);
    orc_window_stats(B, slen, m, sB, muB);
    /* padding invalidated after the statistics (ScrimpDistribPar.cpp:102-108, 783-802) */
    for (long x = slen - pad; x < slen; ++x) {
        if (col == k - 1) {
            B[x] = 0.0;
            if (x < L) sB[x] = muB[x] = 0.0;
        }
        if (row == k - 1) {
            A[x] = 0.0;
            if (x < L) sA[x] = muA[x] = 0.0;
        }
    }
    A[slen] = 0.0; /* the element pushed back for the kernel's last read (:805-806) */
    B[slen] = 0.0;
    /* init_dotproducts (ScrimpDistribPar.cpp:86-100, 810-815):
     * dots[d] = sum_t X[t + d] * Y[t], an FMA chain under -ffast-math */
    const int swap = *is_upper && g_fault != 7;
    const double* X = swap ? B : A;
    const double* Y = swap ? A : B;
    for (long d = 0; d < L; ++d) {
        double z = 0.0;
        for (int t = 0; t < m; ++t) z = fma(X[t + d], Y[t], z);
        dots[d] = z;
    }
    return 0;
}

/* ScrimpDistribPar::eval_tile_blocked (ScrimpDistribPar.cpp:205-247) */
void orc_eval_tile_blocked(int kernel, double prof_colmin[], long idx_colmin[], double prof_rowmin[],
                           long idx_rowmin[], double tmpQ[], long blocklen, long trianglen, long tilelen,
                           const double A_hor[], const double A_vert[], long w, const double s_hor[],
                           const double mu_hor[], const double s_vert[], const double mu_vert[], long baserow,
                           long basecol) {
    orc_kernel_fn kern = kernel ? orc_eval_block_abba : orc_eval_block_scalar;
    for (long bo = 0; bo < tilelen; bo += blocklen) {
        const long nd = (blocklen < tilelen - bo) ? blocklen : tilelen - bo;
        kern(prof_colmin, idx_colmin, prof_rowmin + bo, idx_rowmin + bo, tmpQ + bo, nd, trianglen - bo, A_hor,
             A_vert + bo, w, s_hor, mu_hor, s_vert + bo, mu_vert + bo, baserow + bo, basecol);
    }
}

int orc_distrib_par(const double* series, long n, int m, int k, int blocklen, int kernel, double* profile,
                    long* index) {
    long pad, L;
    if (orc_distrib_geometry(n, m, k, &pad, &L) != 0) return -1;
    const long ez = m / 4;
    const long PL = n + pad - m + 1, P = n - m + 1;
    const long slen = L + m - 1;
    double* score = (double*)malloc(sizeof(double) * PL);
    long* sidx = (long*)malloc(sizeof(long) * PL);
    double* buf = (double*)malloc(sizeof(double) * (2 * (slen + 1) + 5 * L + 2 * (L + ez + 1)));
    long* ibuf = (long*)malloc(sizeof(long) * (L + (L + ez + 1)));
    double *A = buf, *B = A + slen + 1, *sA = B + slen + 1, *muA = sA + L, *sB = muA + L, *muB = sB + L;
    double *dots = muB + L, *hor = dots + L, *vert = hor + L + ez + 1;
    long *hidx = ibuf, *vidx = ibuf + L + ez + 1;
    for (long g = 0; g < PL; ++g) {
        score[g] = -1.0;
        sidx[g] = -1;
    }
    for (long rank = 0; rank < (long)k * k; ++rank) {
        long ro, co;
        int up;
        orc_distrib_tile_inputs(series, n, m, k, rank, A, B, sA, muA, sB, muB, dots, &ro, &co, &up);
        for (long x = 0; x < L + ez + 1; ++x) {
            hor[x] = -1.0;
            hidx[x] = -1;
        }
        for (long x = 0; x < L; ++x) {
            vert[x] = -1.0;
            vidx[x] = -1;
        }
        if (up && g_fault == 6) continue;
        if (up) /* the upper triangle, transposed (ScrimpDistribPar.cpp:819-839) */
            orc_eval_tile_blocked(kernel, vert, vidx, hor + 1, hidx + 1, dots + 1, blocklen, L - 1, L - 1, A, B + 1,
                                  m, sA, muA, sB + 1, muB + 1, co + 1, ro);
        else /* ScrimpDistribPar.cpp:840-862 */
            orc_eval_tile_blocked(kernel, hor, hidx, vert, vidx, dots, blocklen, L, L, B, A, m, sB, muB, sA, muA, ro,
                                  co);
        /* score-max merge (merge_score_profiles, ScrimpDistribPar.cpp:175-200): the
         * horizontal result covers profile entries co + x, the vertical one ro + x */
        for (long x = 0; x < L; ++x) {
            if (hor[x] > score[co + x]) {
                score[co + x] = hor[x];
                sidx[co + x] = hidx[x];
            }
        }
        for (long x = 0; x < L; ++x) {
            if (vert[x] > score[ro + x]) {
                score[ro + x] = vert[x];
                sidx[ro + x] = vidx[x];
            }
        }
    }
    /* d2 = 2m - 2*score (ScrimpDistribPar.cpp:1010-1015); padding entries are not written */
    const double twice_m = 2.0 * (double)m;
    for (long g = 0; g < P; ++g) {
        profile[g] = twice_m - 2.0 * score[g];
        index[g] = sidx[g];
    }
    free(score);
    free(sidx);
    free(buf);
    free(ibuf);
    return 0;
}
