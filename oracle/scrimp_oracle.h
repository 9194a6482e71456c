/*
 * scrimp_oracle.h -- CPU restatement of the reference SCRIMP self-join.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or as the timed CPU baseline.  The product library never links it.
 *
 * Arithmetic definition (what "the reference" computes, written out because
 * the reference's bits depend on `-march=native -ffast-math`,
 * scrimppp/CMakeLists.txt:96-103; GCC contracts a*b+c into one FMA):
 *   stats  (ScrimpSequOpt.cpp:33-79)   cs[i] = A[i] + cs[i-1];  sq[i] = fma(A[i],A[i],sq[i-1])
 *                                      sum[k] = cs[m-1+k] - cs[k-1];  mean = sum/m
 *                                      var = fma(-mean, mean, sumsq/m); sInv = 1/sqrt(var)
 *                                      mus = (mean*sInv)*sqrt(m)
 *   seeds  (ScrimpSequOpt.cpp:81-96)   z = fma(A[k+d], A[k], z), k = 0..m-1 in order
 *   cell   (kernels.cpp:205-222)       score = fma(QT, sv*sh, -(mv*mh))
 *                                      QT = fma(A[i+m], A[j+m], QT); QT = fma(-A[i], A[j], QT)
 *   final  (ScrimpTrivialPar.cpp:336-340)  d2 = 2m - 2*score
 * Seeds are summed sequentially (the reference build may vectorise that sum);
 * everything else has one fixed evaluation order.
 */
#ifndef SCRIMP_ORACLE_H
#define SCRIMP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* window statistics: sInv[P], mus[P]  (ScrimpSequOpt.cpp:33-79) */
void orc_window_stats(const double* A, long n, int m, double* sInv, double* mus);

/* seeds qt[d] for d in [first, last] (ScrimpSequOpt.cpp:81-96); qt indexed by d */
void orc_seeds(const double* A, long n, int m, long first, long last, double* qt);

/* scalar kernel, kernels.cpp:310-366 (the specification of the AVX2 kernel) */
void orc_eval_block_scalar(double prof_colmin[], long idx_colmin[], double prof_rowmin[],
                           long idx_rowmin[], double tmpQ[], long blocklen, long trianglen,
                           const double A_hor[], const double A_vert[], long w,
                           const double s_hor[], const double mu_hor[], const double s_vert[],
                           const double mu_vert[], long baserow, long basecol);

/* AVX2+FMA "ABBA" kernel, kernels.cpp:169-306 (4 diagonals per vector, same math) */
void orc_eval_block_abba(double prof_colmin[], long idx_colmin[], double prof_rowmin[],
                         long idx_rowmin[], double tmpQ[], long blocklen, long trianglen,
                         const double A_hor[], const double A_vert[], long w,
                         const double s_hor[], const double mu_hor[], const double s_vert[],
                         const double mu_vert[], long baserow, long basecol);

/* 1-D partitioning, partitioning_1d.cpp:8-48 */
void orc_get_partition(int part_id, int num_partitions, int ids_to_process, int* first, int* last);

/*
 * scrimp_triv_par driver (ScrimpTrivialPar.cpp:142-368) with `nranks` ranks
 * executed by `nthreads` POSIX threads (each rank: own result SoA, 2
 * partitions r and 2p-1-r, blocks of `blocklen`, seeds per block, kernel per
 * block, d2 transform), followed by the element-wise min reduction of
 * ScrimpTrivialPar.cpp:68-108 in rank order.  `kernel`: 0 scalar, 1 AVX2.
 * `excl` < 0 selects m/4.  `block_stride` > 1 evaluates only every
 * block_stride-th block of each rank (a bounded timing sample); *cells gets
 * the number of cells evaluated and *seconds the maximum over ranks of the
 * stats+seeds+evaluation time plus the reduction.  Returns 0, or -1 when
 * the profile length is <= 0 (the reference throws there).
 */
int orc_triv_par(const double* A, long n, int m, long excl, int blocklen, int nranks,
                 int nthreads, int kernel, int block_stride, double* profile, long* index,
                 double* cells, double* seconds);

/*
 * One rank's share of scrimp_triv_par (ScrimpTrivialPar.cpp:201-340): its own
 * statistics, partitions `rank` and 2*nranks-1-rank, blocks with seeds and the
 * kernel, the d2 transform; profile/index (P entries) are this rank's partial
 * result, to be merged with the element-wise min reduction of
 * ScrimpTrivialPar.cpp:68-108.  Used by the MPI launcher (one process per rank).
 */
int orc_triv_par_rank(const double* A, long n, int m, long excl, int blocklen, int rank, int nranks,
                      int kernel, int block_stride, double* profile, long* index, double* cells,
                      double* seconds);

/*
 * scrimp_vert_blocked driver (ScrimpVertBlocked.cpp:89-186) without the
 * unseeded shuffle of block starts (order only changes ties): index starts
 * at 0 and profile at -1.0.
 */
int orc_vert_blocked(const double* A, long n, int m, long excl, int blocklen, int kernel,
                     double* profile, long* index);

/*
 * distrib_par (ScrimpDistribPar.cpp:444-1116) on one process: the k x k
 * checkerboard of tiles of k^2 ranks, restated without MPI.
 *   geometry (:580-596): W = n + pad - m - ez with pad making W a multiple of
 *     k, tile length L = W/k; rank -> (row, col, upper) by
 *     get_partition_coords (checkerboard_partitioning.cpp:8-41); row offset
 *     row*L + ez + 1, column offset col*L (:666-668).
 *   tile inputs: slices of L + m - 1 points (0 past the series) at both
 *     offsets, window statistics of each slice on its own (:780-781, the
 *     ScrimpSequOpt arithmetic), padding of the last row/column zeroed after
 *     them (:783-802), one 0.0 after each slice (:805-806), seeds
 *     dots[d] = sum_t X[t+d]*Y[t] with (X, Y) = (A, B) for a lower tile and
 *     (B, A) for an upper one (:86-100, 810-815).
 *   evaluation: eval_tile_blocked (:205-247) on the lower triangle, or on the
 *     upper triangle transposed (:819-862).
 *   merge: every tile's horizontal result (entries col*L + x) and vertical
 *     result (entries row*L + ez + 1 + x) into the profile with the score-max
 *     rule of merge_score_profiles (:175-200), in rank order; then
 *     d2 = 2m - 2*score (:1010-1015) for the n - m + 1 entries of the input.
 * The reference reduces over row/column communicators (:899-1004); every step
 * is the same strict maximum, so only exact ties can depend on the order.
 */
void orc_partition_coords(long rank, long* row, long* col, int* is_upper);
int orc_distrib_geometry(long n, int m, int k, long* pad, long* L);
/* A, B: L + m entries; sA, muA, sB, muB, dots: L entries */
int orc_distrib_tile_inputs(const double* series, long n, int m, int k, long rank, double* A, double* B,
                            double* sA, double* muA, double* sB, double* muB, double* dots, long* row_offset,
                            long* col_offset, int* is_upper);
void orc_eval_tile_blocked(int kernel, double prof_colmin[], long idx_colmin[], double prof_rowmin[],
                           long idx_rowmin[], double tmpQ[], long blocklen, long trianglen, long tilelen,
                           const double A_hor[], const double A_vert[], long w, const double s_hor[],
                           const double mu_hor[], const double s_vert[], const double mu_vert[], long baserow,
                           long basecol);
int orc_distrib_par(const double* series, long n, int m, int k, int blocklen, int kernel, double* profile,
                    long* index);

/*
 * Brute force in long double: for every pair |i-j| > excl,
 * d2 = 2m(1 - rho_ij) with rho from two-pass centred sums; profile[i] =
 * min over j, index[i] = the first j attaining it.  O(P^2 m).
 */
void orc_brute_force(const double* A, long n, int m, long excl, double* profile, long* index);

/* d2 of one pair in long double (two-pass, population sigma) */
double orc_pair_dist2(const double* A, long n, int m, long i, long j);

/* Window mean and population sigma of every window in long double (two-pass). */
void orc_window_moments_ld(const double* A, long n, int m, double* mean, double* sigma);

/*
 * PreSCRIMP, the anytime phase of scrimppp_orig (ScrimpppOrig.cpp:79-203 and
 * 130-283).  Statistics in that file's own form (ScrimpppOrig.cpp:84-113):
 *   cs[i] = A[i] + cs[i-1];  sq[i] = fma(A[i], A[i], sq[i-1])
 *   sum = cs[m-1+k] - cs[k-1];  mean = sum/m;  var = fma(-mean, mean, sumsq/m);  sigma = sqrt(var)
 * step = floor(stride*m), excl = m/4.  For each sampled row i (0, step, 2*step, ...
 * in the order given, ascending when `order` is NULL):
 *   QT(i,j) = sum_t A[j+t]*A[i+t] for every window j -- by FFT in the reference
 *             (ScrimpppOrig.cpp:137-186), a direct sequential fma sum here
 *   d(i,j)  = 2*(m - fma(-(m*mean_j), mean_i, QT) / (sigma_j*sigma_i)),
 *             +inf for i-excl < j < i+excl (ScrimpppOrig.cpp:191-200)
 *   profile[j] = d(i,j) where smaller (strict <, ascending j), then profile[i] =
 *   the row minimum and index[i] its first j (ScrimpppOrig.cpp:202-216); then
 *   the diagonal through (i, j = index[i]) is walked step-1 cells forward and
 *   back with lastz = fma(m - profile[i]/2, sigma_j*sigma_i, (m*mean_j)*mean_i)
 *   and lastz = fma(A[i+k+m-1], A[j+k+m-1], fma(-A[i+k-1], A[j+k-1], lastz))
 *   (backward: fma(A[i-k], A[j-k], fma(-A[i-k+m], A[j-k+m], lastz))), updating
 *   both profile[i+k] and profile[j+k] where smaller (ScrimpppOrig.cpp:218-281).
 * A row whose every window is excluded keeps +inf and is not walked (the
 * reference reads an unset index there).  profile[] holds d2 (+inf where never
 * reached), index[] starts at 0.  Only ties depend on the order of rows: every
 * update is a minimum, and the row minimum overwriting profile[i] can never
 * raise it (d(i,j) = d(j,i)).
 */
void orc_prescrimp(const double* A, long n, int m, double stride, const long* order, long norder,
                   double* profile, long* index);

/*
 * Planted-mistake switch for the oracle's own pinning tests: 0 = none,
 * 1 = drop the A[i]*A[j] term of the recurrence, 2 = flip the sign of the
 * mean product in the score, 3 = report the row index for column results,
 * 4 = exclusion zone one diagonal too narrow, 5 = sample instead of
 * population variance, 6 = distrib_par skips its upper tiles, 7 = distrib_par
 * seeds an upper tile without transposing it.
 */
void orc_set_fault(int code);

#ifdef __cplusplus
}
#endif

#endif /* SCRIMP_ORACLE_H */
