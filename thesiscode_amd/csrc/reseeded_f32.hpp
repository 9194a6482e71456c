// Descriptors shared by the host planner and the FP32-reseeded kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"

namespace mp {

// One self-join of a batch.  All per-series arrays (series A, mu, sigma, terms,
// keys, profile, index) live at element offset `off` of their concatenated
// buffers.
struct SeriesDesc {
    int64_t off;       // element offset of this series in every per-series buffer
    int64_t out_off;   // offset of its P entries in the profile/index outputs
    int64_t tile_base; // first tile of this series covered by this descriptor (a slice of its tiles)
    int n, m, P;       // series length, window, profile length
    int base;          // first diagonal: exclusion + 1
    int tl;            // diagonals = columns of the triangle: P - base
    int nbands;        // bands of 64*K diagonals
    int band_off;      // offset of this series' band prefix (nbands+1 entries)
    float init_c1;     // 1 + rho of the initial score -1.0: 1 - 1/m
    int pad;
};

struct F32rBatch {
    const SeriesDesc* descs;    // device, nseries
    const int64_t* entry_prefix;  // device, nseries+1: prefix of P
    const int64_t* stat_prefix;   // device, nseries+1: prefix of ceil(P / kStatChunk)
    const int64_t* tile_prefix;   // device, nseries+1: prefix of tiles
    const int64_t* band_prefix;   // device, concatenated per-series tile prefixes over bands
    int nseries;
    int64_t nentries, ntiles, nstat;  // entries, tiles, statistics chunks
    int band_k, tile_len;
    const double* A;
    double *mu, *sigma;
    void* terms;   // one TermRec per entry
    int accum64;   // covariance in FP64 (folded per-window FP32 partials) or FP32
    unsigned long long* keys;
    unsigned int* counter;
};

// The recurrence terms of window i, read together: df = (A[i+m] - A[i]) / 2,
// dg = (A[i+m] - mu[i+1]) + (A[i] - mu[i]) (both rounded to FP32; zero for the
// last window) and inv = 1 / (sigma sqrt(m)) (0 for a flat window).  The tile
// kernel loads one record per column and per row fed into its band.
struct alignas(16) TermRec {
    float df, dg, inv, pad;
};
inline size_t term_rec_bytes() { return sizeof(TermRec); }

// Tiles of a band of L columns (the band's longest diagonal).  Columns
// [0, band_main_cols) are cut into main tiles of at most T columns, made of
// whole flush windows in which every cell lies in the triangle; the remaining
// columns, where the triangle ends inside the band, form one edge tile whose
// windows test each cell.
__host__ __device__ inline int band_main_cols(int L, int band_k) {
    const int lm = L - (kWave * band_k + kFlush - 2);
    return lm > 0 ? (lm / kFlush) * kFlush : 0;
}
__host__ __device__ inline int64_t band_tiles(int64_t L, int band_k, int T) {
    if (L <= 0) return 0;
    const int lm = band_main_cols(static_cast<int>(L), band_k);
    return (lm + T - 1) / T + 1;
}

// Longest re-seed interval (tile length) that the library, the batch autotuner
// and the server pick by themselves.  With FP64 accumulation every interval up
// to 65536 keeps the profile within 1e-4 of FP64 (n = 2^22, m = 1024 near-periodic
// series: 1.9e-5 at every B from 256 to 65536; n = 2^20, m = 256 random walk:
// 2.0e-5 up to 16384, 3.2e-5 at 32768).  FP32 accumulation drifts with the
// interval: on random walks it stays within 1e-4 up to B = 1024 (6.5e-5) and
// leaves it from B = 4096 on (1.5e-4), so automatic choices stop at 1024 there;
// a caller that asks for a longer interval explicitly still gets it.
inline int max_auto_tile_len(int accum64) { return accum64 ? 32768 : 1024; }
// Tile lengths above kShortTileMax are taken only where they still leave 16
// tiles for each resident wave (16 per CU): longer tiles save seeds (2.7% of the
// kernel at n = 2^22, m = 1024 from 8192 to 32768), but too few of them leave
// waves idle at the end (n = 2^20, m = 256: 78 ms at 8192 or 16384, 81 ms at 32768).
constexpr int kShortTileMax = 8192;
inline int64_t long_tile_target(int cus) { return 256LL * (cus > 0 ? cus : 1); }

// Window statistics are computed in chunks of kStatChunk windows of one series;
// each chunk restarts its sums at its first window, so the statistics of a
// series do not depend on what else the launch carries.
constexpr int kStatChunk = 64;
inline int64_t stat_chunks(int P) { return (static_cast<int64_t>(P) + kStatChunk - 1) / kStatChunk; }

// Staging of served requests between pinned host memory and the device pools:
// one launch per tick moves every request's data (chunks of kStageChunk
// elements, located by binary search over chunk_base).
constexpr int kStageChunk = 4096;
struct StageDesc {
    int64_t pin_off;       // element offset in the pinned pool
    int64_t dev_off;       // element offset in the device pools
    int64_t band_dev_off;  // offset of the band prefix in the device band pool (stage in)
    int64_t chunk_base;    // first chunk of this request in the launch
    int n;                 // stage in: series length; stage out: profile length P
    int nb;                // stage in: band prefix entries (stored after the series)
};
int launch_stage_in(const StageDesc* d, int count, int64_t nchunks, const double* pinned,
                    double* A, int64_t* bandpool, hipStream_t st);
int launch_stage_out(const StageDesc* d, int count, int64_t nchunks, const double* prof,
                     const int64_t* idx, double* pinned, hipStream_t st);

int launch_f32r_prep(const F32rBatch& b, hipStream_t st);
int launch_f32r_tiles(const F32rBatch& b, int groups, hipStream_t st);
int launch_f32r_unpack(const F32rBatch& b, int refine, int64_t index_init, double* profile,
                       int64_t* index, hipStream_t st);

}  // namespace mp
