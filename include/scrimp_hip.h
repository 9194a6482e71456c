/*
 * scrimp_hip.h -- C ABI of the MI355X SCRIMP matrix-profile self-join.
 *
 * Plain C types only (pointers, sizes, status codes); no C++ or torch types
 * cross this boundary and no exception escapes it.  Every function returns
 * MP_OK (0) or a negative MP_ERR_* status.
 *
 * Result convention (same as the reference's MatProfSOA,
 * scrimppp/inc/ScrimpTrivialPar.hpp:16-20): two arrays of length
 * P = n - m + 1, profile[i] = squared z-normalised Euclidean distance
 * d^2 = 2m - 2*score of the best match of window i, index[i] = that match's
 * start or -1 when no admissible window scored above the initial -1.0
 * (scrimppp/src/ScrimpTrivialPar.cpp:167-168, 336-340).  The exclusion zone
 * defaults to m/4: pairs with |i-j| <= m/4 are never compared
 * (scrimppp/src/ScrimpTrivialPar.cpp:149-150, 202, 221-224).
 */
#ifndef SCRIMP_HIP_H
#define SCRIMP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
enum {
    MP_OK = 0,
    MP_ERR_BAD_WINDOW = -1,    /* m < 1 or exclusion < 0                                    */
    MP_ERR_SHORT_SERIES = -2,  /* n < m: profile length <= 0
                                  (reference throws, ScrimpTrivialPar.cpp:182-184)          */
    MP_ERR_NULL = -3,          /* a required pointer is NULL                                */
    MP_ERR_OOM = -4,           /* device or host allocation failed                          */
    MP_ERR_DEVICE = -5,        /* a HIP runtime call failed                                 */
    MP_ERR_TOO_LONG = -6,      /* profile length beyond the supported 2^30
                                  (reference: MAX_PROFILE_LENGTH, ScrimpTrivialPar.cpp:179)  */
    MP_ERR_BAD_ARG = -7,       /* an option or size is out of range                         */
    MP_ERR_NO_DEVICE = -8,     /* no HIP device / the library was built without one         */
    MP_ERR_ADMISSION = -9,     /* request does not fit the device memory budget             */
    MP_ERR_NOT_FOUND = -10     /* unknown request handle                                    */
};

const char* mp_status_string(int status);

/* ---- options and timing -------------------------------------------------- */
typedef struct mp_opts {
    int32_t exclusion;   /* exclusion zone; -1 selects the reference default m/4            */
    int32_t band_k;      /* diagonals per lane (band width = 64*band_k); 0 = default       */
    int32_t tile_len;    /* FP32 path: columns per tile = re-seed interval B; 0 = default  */
    int32_t refine;      /* FP32 path: 1 = recompute each winner's distance in FP64        */
    int32_t index_init;  /* index reported when nothing beat the initial score: -1 (the
                            parallel drivers) or 0 (scrimp_vert_blocked,
                            ScrimpVertBlocked.cpp:104)                                     */
    int32_t accum64;     /* FP32 path: 1 (default) = each diagonal's covariance kept in
                            FP64, advanced by FP32 partial sums of 16 steps folded in
                            every 16 columns; within 1e-4 of the FP64 profile. 0 = FP32
                            accumulation, about 10% faster, drifts to ~1e-3 on long
                            near-periodic series                                          */
    int32_t walk_form;   /* exact path: how the band walker runs.  0 = chosen from the
                            work size; at band_k = 8: 1 = one wave per SIMD, rows from
                            an LDS ring; 2 = two waves per SIMD, rows shifted by DPP;
                            3 = two waves per SIMD, rows in place and the diagonals'
                            running dot products rotated by DPP; at band_k = 1, 2 or 4
                            (default 4): 4 = every band cut into 2048-column segments,
                            one wave each, started from values a pre-pass advances
                            along the diagonals (series with W <= 2^18 diagonals; the
                            automatic choice on short series).  All give the same
                            profile values                                              */
    int32_t reserved;
} mp_opts;

void mp_opts_default(mp_opts* o);

typedef struct mp_timing {
    double total_ms;     /* device time of the whole self-join (HIP events)                */
    double stats_ms;     /* window statistics                                              */
    double seeds_ms;     /* FP64 diagonal seeds (exact path)                               */
    double kernel_ms;    /* diagonal band / tile kernel                                    */
    double finalize_ms;  /* refine + unpack to profile/index                               */
    double cells;        /* cells evaluated: W(W+1)/2, W = P - excl - 1                     */
    int64_t work_items;  /* bands or column segments (exact path), tiles (FP32 path)      */
    int32_t band_k, tile_len;
    int32_t walk_form;   /* exact path: the walker form that ran (mp_opts.walk_form
                            numbering: 1 LDS ring, 2 DPP rows, 3 rotating, 4 column
                            segments); 0 for whole bands at band width 1, 2 or 4       */
    int32_t reserved;
} mp_timing;

/* ---- device context (one per thread or stream; calls on one are serialised) */
typedef struct mp_ctx mp_ctx;

int mp_device_count(int* count);
int mp_ctx_create(int device, mp_ctx** out);
int mp_ctx_destroy(mp_ctx* ctx);
int mp_ctx_sync(mp_ctx* ctx);
/* plumbing for callers that keep inputs resident in HBM */
int mp_dev_alloc(mp_ctx* ctx, uint64_t bytes, void** out);
int mp_dev_free(mp_ctx* ctx, void* p);
int mp_memcpy_h2d(mp_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int mp_memcpy_d2h(mp_ctx* ctx, void* dst, const void* src, uint64_t bytes);

/* ---- self-join, plain FP64 (bit-exact restatement of the reference) -------
 * Replaces steps 5-8 of scrimp_triv_par::compute_matrix_profile
 * (scrimppp/src/ScrimpTrivialPar.cpp:263-340): window statistics
 * (ScrimpSequOpt.cpp:33-79), seeds (ScrimpSequOpt.cpp:81-96), the diagonal
 * kernel (kernels.cpp:169-306) over all diagonals ez+1..P-1, the rank merge
 * (ScrimpTrivialPar.cpp:68-108) and the d^2 transform (336-340).
 * Host pointers; ts has n elements, profile/index have P elements.          */
int mp_selfjoin_f64(const double* ts, int64_t n, int32_t m, double* profile,
                    int64_t* index, const mp_opts* opts, mp_timing* timing);
/* Same on device pointers (d_ts: n doubles, d_profile/d_index: P each). */
int mp_selfjoin_f64_dev(mp_ctx* ctx, const double* d_ts, int64_t n, int32_t m,
                        double* d_profile, int64_t* d_index, const mp_opts* opts,
                        mp_timing* timing);

/* ---- the exact self-join without device allocation ------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory while they run (tight serving loops, code that must not call
 * hipMalloc): the same computation and results as mp_selfjoin_f64_dev, bit for
 * bit, in a file-scope device array of the library instead of the context's
 * growable arena, so the call allocates and frees nothing.  The array holds
 * series of up to 524,288 points (2^19, about 50 MB on every GPU the library is
 * loaded on, kept until the process ends); a longer series returns
 * MP_ERR_TOO_LONG.  Column segments (walk_form 4) run where their start values
 * fit the array as well; otherwise the automatic choice walks whole bands (the
 * same results) and a forced walk_form 4 returns MP_ERR_TOO_LONG.  All calls on
 * one GPU share the array: they are serialised
 * inside the library (a lock per GPU), and a call on another stream or process
 * that overlaps one on the same GPU without this library's lock would race.
 * The array keeps the last call's contents; every buffer in it is written
 * before it is read.  Like mp_selfjoin_f64_dev it returns after the results
 * are in d_profile/d_index.                                                 */
int mp_selfjoin_f64_dev_noalloc(mp_ctx* ctx, const double* d_ts, int64_t n, int32_t m,
                                double* d_profile, int64_t* d_index, const mp_opts* opts,
                                mp_timing* timing);

/* ---- PreSCRIMP, the anytime phase of scrimppp_orig --------------------------
 * Replaces the PreSCRIMP loop of ScrimpppOrig::compute_matrix_profile
 * (scrimppp/src/ScrimpppOrig.cpp:130-283): every sampled row i = 0, step,
 * 2*step, ... (step = floor(stride*m), at least 1) gets its full distance
 * profile, which lowers profile[j] for every j and sets profile[i] to the
 * row's minimum; the diagonal through (i, that minimum's partner) is then
 * walked step-1 cells forward and back.  Partners with |i-j| < m/4 are
 * excluded (ScrimpppOrig.cpp:191-200).  profile[] gets d^2 (+inf where no
 * update reached), index[] the partner (0 where none, as the reference
 * initialises it).  The reference's shuffled order of rows only decides
 * ties, and its FFT-based dot products agree with the direct FP64 sums here
 * to rounding.  Host pointers; ts has n elements, profile/index P each.     */
int mp_prescrimp_f64(const double* ts, int64_t n, int32_t m, double stride, double* profile,
                     int64_t* index, mp_timing* timing);
/* Same on device pointers. */
int mp_prescrimp_f64_dev(mp_ctx* ctx, const double* d_ts, int64_t n, int32_t m, double stride,
                         double* d_profile, int64_t* d_index, mp_timing* timing);

/* ---- self-join, FP32 recurrence with FP64 statistics and re-seeding -------
 * Same result convention.  The covariance recurrence runs in FP32; window
 * statistics use compensated FP64 sums; every tile of tile_len columns is
 * seeded with an exact FP64 dot product; with opts->refine the reported
 * distance of each winner is recomputed in FP64.                          */
int mp_selfjoin_f32r(const double* ts, int64_t n, int32_t m, double* profile,
                     int64_t* index, const mp_opts* opts, mp_timing* timing);
int mp_selfjoin_f32r_dev(mp_ctx* ctx, const double* d_ts, int64_t n, int32_t m,
                         double* d_profile, int64_t* d_index, const mp_opts* opts,
                         mp_timing* timing);

/* ---- kernel-level drop-in ---------------------------------------------------
 * Exact 16-argument semantics of matrix_profile::eval_diag_block_triangle
 * (scrimppp/inc/kernels.h:7-24, scrimppp/src/kernels.cpp:368-410): for every
 * column j < trianglen and diagonal offset 0 <= i-j < min(blocklen,
 * trianglen-j), score = tmpQ[i-j]*(s_vert[i]*s_hor[j]) - mu_vert[i]*mu_hor[j];
 * tmpQ advances by A_vert[i+w]*A_hor[j+w] - A_vert[i]*A_hor[j]; the row
 * maximum goes to prof_rowmin[i]/idx_rowmin[i] (index j+basecol), the column
 * maximum to prof_colmin[j]/idx_colmin[j] (index i+baserow), replacing only on
 * a strictly greater score.  Row and column views may alias one array as in
 * the reference's self-join adapters (ScrimpTrivialPar.cpp:110-140).  Host
 * pointers; runs on the device of a library-owned context.  The final tmpQ
 * update at i = trianglen-1 reads A_vert[trianglen-1+windowSize] like the
 * reference (kernels.cpp:218-222); that element must be readable.        */
int mp_eval_diag_block_triangle_f64(double prof_colmin[], int64_t idx_colmin[],
                                    double prof_rowmin[], int64_t idx_rowmin[],
                                    double tmpQ[], int64_t blocklen, int64_t trianglen,
                                    const double A_hor[], const double A_vert[],
                                    int64_t windowSize, const double s_hor[],
                                    const double mu_hor[], const double s_vert[],
                                    const double mu_vert[], int64_t baserow,
                                    int64_t basecol);

/* ---- batched FP32-reseeded self-joins --------------------------------------
 * Thousands of independent series of uneven length and window in one
 * persistent launch.  Series s occupies lengths[s] consecutive doubles of the
 * concatenated input and P_s = lengths[s]-windows[s]+1 consecutive entries of
 * the concatenated profile/index outputs (device pointers).  The exclusion of
 * each series is windows[s]/4 unless opts->exclusion >= 0.  With use_graph
 * the whole batch (statistics, tiles, unpack) is captured once in a hipGraph
 * and replayed while the pointers stay the same.  mp_batch_tune times band
 * widths 4/8 and tile lengths 256..max_tile_len (<= 0: the longest automatic
 * one, 32768 with FP64 accumulation and 1024 with FP32) on the real
 * batch, keeps the fastest and records it in an on-disk cache that later
 * mp_batch_create calls consult ($SCRIMP_HIP_TUNE_CACHE, default
 * ~/.cache/scrimp_hip/).
 * No counterpart in the reference, which joins one series per run.        */
typedef struct mp_batch mp_batch;
int mp_batch_create(mp_ctx* ctx, int32_t nseries, const int64_t* lengths,
                    const int32_t* windows, const mp_opts* opts, mp_batch** out);
int mp_batch_run(mp_batch* batch, const double* d_ts, double* d_profile, int64_t* d_index,
                 int32_t use_graph, mp_timing* timing);
int mp_batch_tune(mp_batch* batch, const double* d_ts, double* d_profile, int64_t* d_index,
                  int32_t max_tile_len, mp_timing* best);
int mp_batch_info(const mp_batch* batch, int64_t* total_n, int64_t* total_p, int64_t* ntiles,
                  double* cells, int32_t* band_k, int32_t* tile_len);
int mp_batch_destroy(mp_batch* batch);

/* ---- serving: continuous batching of self-join requests --------------------
 * mp_server_create starts an engine thread on the context's device with an
 * HBM budget for resident request state (64 bytes per point: series,
 * statistics, recurrence-term record, partial-profile keys, outputs).
 * Requests are admitted in arrival order while they fit; each engine tick
 * launches the FP32-reseeded tile kernel once over a share of the remaining
 * tiles of every admitted request (about tick_cells cells per tick; <= 0 picks
 * 1.5e9), so
 * requests that arrive while others run join the next launch.  A request's
 * profile/index (P = n-m+1 entries each, host memory the caller keeps alive
 * until completion) are written when it completes.  opts: exclusion, band_k,
 * tile_len (re-seed interval), refine, index_init, accum64 as for
 * mp_selfjoin_f32r.
 * submit returns MP_ERR_ADMISSION when a request can never fit the budget.
 * poll/wait return the request's own status and, once done, its latency
 * (submit -> results written) and forget it.  No counterpart in the
 * reference (a batch program that joins one series per run).              */
typedef struct mp_server mp_server;
int mp_server_create(mp_ctx* ctx, uint64_t hbm_budget_bytes, int64_t tick_cells,
                     const mp_opts* opts, mp_server** out);
int mp_server_submit(mp_server* server, const double* ts, int64_t n, int32_t m,
                     double* profile_out, int64_t* index_out, uint64_t* request_id);
int mp_server_poll(mp_server* server, uint64_t request_id, int32_t* done, double* latency_ms,
                   double* queue_ms);
int mp_server_wait(mp_server* server, uint64_t request_id, double* latency_ms);
int mp_server_completed(mp_server* server, int32_t max, uint64_t* ids, int32_t* statuses,
                        double* latency_ms, double* queue_ms, int32_t* count);
int mp_server_stats(mp_server* server, int64_t* ticks, int64_t* completed, double* cells_done,
                    int64_t* queued, int64_t* active, int64_t* hbm_used_bytes);
int mp_server_destroy(mp_server* server);

/* ---- synthetic inputs (CPU only, bit-identical on every host) -------------
 * Counter-based Philox4x32-10 (Salmon et al., SC'11; key = 64-bit seed,
 * words {seed lo, seed hi}) + Box-Muller.  None of these has a counterpart in
 * the reference, which reads its series from files (ScrimpSequ.hpp:48-112). */
void mp_gen_philox4x32_10(const uint32_t ctr[4], uint64_t key, uint32_t out[4]);
void mp_gen_normals(uint64_t seed, uint32_t stream, int64_t first, int64_t count, double* out);
void mp_gen_uniforms(uint64_t seed, uint32_t stream, int64_t first, int64_t count, double* out);
void mp_gen_random_walk(uint64_t seed, uint32_t stream, int64_t n, double* out);
int mp_gen_precision_series(uint64_t seed, int64_t n, int32_t m, double* out, double* meta);

#ifdef __cplusplus
}
#endif

#endif /* SCRIMP_HIP_H */
