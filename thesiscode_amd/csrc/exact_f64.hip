// Exact FP64 SCRIMP path: the reference's arithmetic, bit for bit, on gfx950.
//
//   k_ref_cumsum   sequential prefix sums of A and A^2 (ScrimpSequOpt.cpp:47-54)
//   k_ref_stats    1/sigma and mu*sqrt(m)/sigma per window (ScrimpSequOpt.cpp:55-73)
//   k_ref_seeds    first-cell dot product of every diagonal (ScrimpSequOpt.cpp:81-96)
//   k_walk_f64     band walker: the diagonal kernel of kernels.cpp:169-306 over
//                  all diagonals, with the row/column maxima merged into HBM
//   k_unpack_f64   score -> d^2 = 2m - 2*score and index decode (ScrimpTrivialPar.cpp:336-340)
//
// Band walker layout.  A wave owns a band of 64*K adjacent diagonals; lane l
// owns diagonals dbase+K*l .. dbase+K*l+K-1 and the wave walks the band's
// columns j = 0, 1, ... from the first cell of every diagonal, so each
// diagonal's QT recurrence runs in exactly the reference's order (never
// re-seeded, kernels.cpp:221-222).  At column j lane l touches rows
// j+dbase+K*l+k.  A row's record (A[i], A[i+m], sInv[i], mus[i] and its
// running maximum) stays in register slot (j+k)%K while the row passes the
// lane's K diagonals, then moves to lane l-1 with one DPP wave_shl:1 per
// dword; lane 63 receives the next row from a scalar load.  The row leaving
// lane 0 has seen every diagonal of the band.  All lanes share column j: each
// lane writes its K-cell column maximum to LDS, and every 16 columns the wave
// reduces the 64 partials per column.  Row and column results are merged into
// HBM with two 64-bit atomicMax per result: one on the exact ordered score,
// one on a key of (truncated score, partner index).
//
// Lazy argmax.  A step keeps maxima only: the column maximum of a lane's K
// cells and, per row, the maximum over the lane's cells of that row (the
// exact scores, so the profile values stay bit-identical).  Which lane holds
// the argmax is tracked exactly (a row carries the lane that last raised it;
// a column's flush reduction keeps the lane of its maximum); which of that
// lane's K diagonals it was is left open: a result's key holds the lane's
// first candidate partner and the direction of the others, and the unpack
// recomputes the K candidates' scores of each profile entry's winning key from
// direct FP64 dot products and takes the one closest to the exact maximum.
#include <type_traits>
#include <utility>

#include "common.hpp"
#include "exact_f64.hpp"

namespace mp {

// The prefix sums of A and of A^2 are two independent sequential chains (the
// reference's order, an add chain and an FMA chain).  Each runs in its own
// workgroup (blockIdx.x 0: A, 1: A^2), where one lane scans the whole series:
// the dependent FP64 chain is the critical path.  The wave stages the series
// through LDS in chunks with coalesced vector loads issued one chunk ahead
// (vector loads retire in order, so the scan never waits on the chunk in
// flight).  The scanning lane takes its inputs 16 at a time (eight 16-byte
// LDS reads, one group ahead) and stores its results two per 16-byte store,
// one group behind: per element it issues one dependent add or FMA, half an
// LDS read and half a store (45 -> 34 ms at n = 2^22, where one LDS read and
// one store per element had the chain wait on them).
constexpr int kScanChunk = 2048;
constexpr int kScanGroup = 16;

__global__ __launch_bounds__(64) void k_ref_cumsum(const double* __restrict__ A, int n,
                                                  double* __restrict__ cs,
                                                  double* __restrict__ sq) {
    __shared__ __align__(16) double buf[2][kScanChunk];
    constexpr int R = kScanChunk / 64;
    const int lane = threadIdx.x;
    const bool squares = blockIdx.x == 1;
    double* __restrict__ out = squares ? sq : cs;
    if (n <= 0) return;
    const int nchunks = (n + kScanChunk - 1) / kScanChunk;
    double reg[R];
    auto load = [&](int ch) {
        const int base = ch * kScanChunk;
#pragma unroll
        for (int t = 0; t < R; ++t) {
            const int i = base + t * 64 + lane;
            reg[t] = A[min(i, n - 1)];
        }
    };
    auto stage = [&](int slot) {
#pragma unroll
        for (int t = 0; t < R; ++t) buf[slot][t * 64 + lane] = reg[t];
    };
    load(0);
    stage(0);
    __syncthreads();
    double c = 0.0;
    // the scanning lane's pass over one chunk; SQ: the chain of squares
    auto scan = [&](auto sq_tag, const double* b, double* dst, int cnt, bool first) __attribute__((always_inline)) {
        constexpr bool SQ = decltype(sq_tag)::value;
        auto next = [](double acc, double a) __attribute__((always_inline)) { return SQ ? fma(a, a, acc) : a + acc; };
        int t = 0;
        if (first) {
            // the first element starts the chain: c = A[0] (or A[0]^2), not 0 + A[0]
            c = SQ ? b[0] * b[0] : b[0];
            dst[0] = c;
            for (t = 1; t < kScanGroup && t < cnt; ++t) {
                c = next(c, b[t]);
                dst[t] = c;
            }
        }
        // A group's inputs are read while the previous group's chain runs, and
        // its results are stored after the next group's chain: two register
        // sets alternate for each, so the chain never writes a register that a
        // store in flight still reads (the read past the chunk's last group is
        // clamped into the chunk and not used).
        auto read_group = [&](int at, double2 (&v)[kScanGroup / 2]) __attribute__((always_inline)) {
            const double2* in2 = reinterpret_cast<const double2*>(b + min(at, kScanChunk - kScanGroup));
#pragma unroll
            for (int q = 0; q < kScanGroup / 2; ++q) v[q] = in2[q];
        };
        auto run_group = [&](const double2 (&v)[kScanGroup / 2], double2 (&r)[kScanGroup / 2])
            __attribute__((always_inline)) {
#pragma unroll
            for (int q = 0; q < kScanGroup / 2; ++q) {
                const double r0 = next(c, v[q].x);
                c = next(r0, v[q].y);
                r[q] = make_double2(r0, c);
            }
        };
        auto store_group = [&](int at, const double2 (&r)[kScanGroup / 2]) __attribute__((always_inline)) {
            double2* out2 = reinterpret_cast<double2*>(dst + at);
#pragma unroll
            for (int q = 0; q < kScanGroup / 2; ++q) out2[q] = r[q];
        };
        if (t + kScanGroup <= cnt) {
            double2 va[kScanGroup / 2], vb[kScanGroup / 2], ra[kScanGroup / 2], rb[kScanGroup / 2];
            read_group(t, va);
            read_group(t + kScanGroup, vb);
            run_group(va, ra);
            t += kScanGroup;  // ra holds group t - G, not yet stored
            for (; t + 2 * kScanGroup <= cnt; t += 2 * kScanGroup) {
                read_group(t + kScanGroup, va);
                run_group(vb, rb);
                store_group(t - kScanGroup, ra);
                read_group(t + 2 * kScanGroup, vb);
                run_group(va, ra);
                store_group(t, rb);
            }
            // ra: group t - G; vb: group t if it is whole
            if (t + kScanGroup <= cnt) {
                run_group(vb, rb);
                store_group(t - kScanGroup, ra);
                store_group(t, rb);
                t += kScanGroup;
            } else {
                store_group(t - kScanGroup, ra);
            }
        }
        for (; t < cnt; ++t) {
            c = next(c, b[t]);
            dst[t] = c;
        }
    };
    for (int ch = 0; ch < nchunks; ++ch) {
        if (ch + 1 < nchunks) load(ch + 1);
        const int base = ch * kScanChunk;
        const int cnt = min(kScanChunk, n - base);
        if (lane == 0) {
            if (squares)
                scan(std::true_type{}, buf[ch & 1], out + base, cnt, ch == 0);
            else
                scan(std::false_type{}, buf[ch & 1], out + base, cnt, ch == 0);
        }
        if (ch + 1 < nchunks) stage((ch + 1) & 1);
        __syncthreads();
    }
}

__global__ void k_ref_stats(const double* __restrict__ cs, const double* __restrict__ sq, int P,
                            int m, double* __restrict__ sInv, double* __restrict__ mus) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P) return;
    const double sum = (k == 0) ? cs[m - 1] : cs[m - 1 + k] - cs[k - 1];
    const double ssq = (k == 0) ? sq[m - 1] : sq[m - 1 + k] - sq[k - 1];
    const double dm = static_cast<double>(m);
    const double mean = sum / dm;
    const double var = fma(-mean, mean, ssq / dm);
    const double si = 1.0 / sqrt(var);
    sInv[k] = si;
    mus[k] = (mean * si) * sqrt(dm);
}

// qt[d] = sum_k A[base+d+k] * A[k], sequential FMA chain, for d in [0, count).
__global__ void k_ref_seeds(const double* __restrict__ A, int base, int count, int m,
                            double* __restrict__ qt) {
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= count) return;
    const double* a = A + base + d;
    double z = 0.0;
    for (int k = 0; k < m; ++k) z = fma(a[k], A[k], z);
    qt[d] = z;
}

__global__ void k_fill_u64(unsigned long long* __restrict__ p, int64_t count,
                           unsigned long long v) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}

__global__ void k_fill_u32(unsigned int* __restrict__ p, int count, unsigned int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}

// score -> d^2 = 2m - 2*score and the key's partner.  Keys of a walker of
// band width K > 1 name the K candidate partners of the winning lane's
// diagonals (code = p0 << 1 | dir: p0 + k, or p0 - k when dir = 1); the one
// whose recomputed score is closest to the exact maximum is the index (equal
// distances: the smaller partner).  The candidates' windows overlap, so each
// block of K window positions loads K values of window i and K of the
// candidates' (their other K-1 values are the previous block's last ones) and
// does K*K FMAs.  Candidates outside the profile or inside the
// exclusion zone are not cells and are skipped.
template <int K>
__device__ __forceinline__ int64_t pick_candidate(const double* __restrict__ A, const double* __restrict__ sInv,
                                                  const double* __restrict__ mus, int i, int64_t code, int P, int m,
                                                  int base, double score) {
    const int64_t p0 = code >> 1;
    const int64_t lo = (code & 1) ? p0 - (K - 1) : p0;  // the candidates are lo .. lo+K-1
    const int64_t ws = lo < 0 ? 0 : lo;                 // first window read (>= 0)
    const double* __restrict__ x = A + i;
    const double* __restrict__ y = A + ws;
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0.0;
    // ys slides along y: a block loads its last K values, the first K-1 are
    // the previous block's last ones
    double ys[2 * K - 1];
#pragma unroll
    for (int v = 0; v < K - 1; ++v) ys[v] = y[v];
    int t = 0;
    for (; t + K <= m; t += K) {
        double xs[K];
#pragma unroll
        for (int u = 0; u < K; ++u) xs[u] = x[t + u];
#pragma unroll
        for (int v = K - 1; v < 2 * K - 1; ++v) ys[v] = y[t + v];
#pragma unroll
        for (int k = 0; k < K; ++k) {
#pragma unroll
            for (int u = 0; u < K; ++u) acc[k] = fma(xs[u], ys[u + k], acc[k]);
        }
#pragma unroll
        for (int v = 0; v < K - 1; ++v) ys[v] = ys[v + K];
    }
    for (; t < m; ++t) {
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] = fma(x[t], y[t + k], acc[k]);
    }
    int64_t best = -1;
    double bestd = 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int64_t p = ws + k;
        const int64_t gap = p > i ? p - i : i - p;
        if (p < lo || p >= lo + K || p >= P || gap < base) continue;
        double dd = fabs(fma(acc[k], sInv[i] * sInv[p], -(mus[i] * mus[p])) - score);
        if (!(dd == dd)) dd = __builtin_inf();
        if (best < 0 || dd < bestd) {  // ascending p: an equal distance keeps the smaller partner
            best = p;
            bestd = dd;
        }
    }
    return best >= 0 ? best : p0;
}

template <int K>
__global__ void k_unpack_f64(const unsigned long long* __restrict__ val,
                             const unsigned long long* __restrict__ key, int P, int m, int ib,
                             int64_t index_init, const double* __restrict__ A,
                             const double* __restrict__ sInv, const double* __restrict__ mus, int base,
                             double* __restrict__ profile, int64_t* __restrict__ index) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const double score = unord_f64(val[i]);
    profile[i] = 2.0 * static_cast<double>(m) - 2.0 * score;
    const int64_t code = key_partner(key[i], ib);
    if (code < 0) {
        index[i] = index_init;
    } else if constexpr (K == 1) {
        index[i] = code;
    } else {
        index[i] = pick_candidate<K>(A, sInv, mus, i, code, P, m, base, score);
    }
}

// Block drop-in unpack: decode a device range back into score-domain values;
// a sentinel partner keeps the caller's original index.
__global__ void k_unpack_block(const unsigned long long* __restrict__ val,
                               const unsigned long long* __restrict__ key, int count, int ib,
                               int64_t pmin, const int64_t* __restrict__ idx_in,
                               double* __restrict__ prof_out, int64_t* __restrict__ idx_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    prof_out[i] = unord_f64(val[i]);
    const int64_t p = key_partner(key[i], ib);
    idx_out[i] = (p < 0) ? idx_in[i] : p + pmin;
}

__global__ void k_pack_block(const double* __restrict__ prof, int count, int ib,
                             unsigned long long* __restrict__ val,
                             unsigned long long* __restrict__ key) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint64_t o = ord_f64(prof[i]);
    val[i] = o;
    key[i] = o | ((1ull << ib) - 1ull);
}

__global__ void k_pack_rec(const double* __restrict__ A, const double* __restrict__ s,
                           const double* __restrict__ mu, int m, int valid, int total,
                           WalkRec* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    WalkRec r;
    if (i < valid) {
        r.a = A[i];
        r.am = A[i + m];
        r.s = s[i];
        r.mu = mu[i];
    } else {
        r.a = 0.0;
        r.am = 0.0;
        r.s = __builtin_nan("");
        r.mu = 0.0;
    }
    out[i] = r;
}

// A column result: its exact maximum score and the lane that holds it (the
// smallest such lane: the smaller diagonal, so the smaller partner, as the
// result keys rank exact ties); id -1: every partial was a masked cell (NaN).
struct ScoreId {
    double v;
    int id;
};
// Maximum of two scores with the plain instruction: it returns the other
// operand when one is a quiet NaN (a masked cell), and the scores here carry
// no signalling NaNs, so the per-operand quieting a generic fmax adds is not
// needed.
__device__ __forceinline__ double dmax(double a, double b) {
    double r;
    asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// dmax() and min() over lanes l, l^16, l^32 and l^48, in every lane.  Each
// gfx950 row swap leaves a lane with its own dword and the dword of lane l^32
// (l^16) in two registers, in some order; both operations are symmetric, so
// the order is free.
__device__ __forceinline__ double max_quarters(double x) {
    auto step = [](double v, auto swap) __attribute__((always_inline)) {
        const auto lo = swap(static_cast<unsigned>(__double2loint(v)));
        const auto hi = swap(static_cast<unsigned>(__double2hiint(v)));
        return dmax(__hiloint2double(static_cast<int>(hi[0]), static_cast<int>(lo[0])),
                    __hiloint2double(static_cast<int>(hi[1]), static_cast<int>(lo[1])));
    };
    x = step(x, [](unsigned v) { return __builtin_amdgcn_permlane32_swap(v, v, false, false); });
    return step(x, [](unsigned v) { return __builtin_amdgcn_permlane16_swap(v, v, false, false); });
}
__device__ __forceinline__ int min_quarters(int x) {
    auto step = [](int v, auto swap) __attribute__((always_inline)) {
        const auto s = swap(static_cast<unsigned>(v));
        return min(static_cast<int>(s[0]), static_cast<int>(s[1]));
    };
    x = step(x, [](unsigned v) { return __builtin_amdgcn_permlane32_swap(v, v, false, false); });
    return step(x, [](unsigned v) { return __builtin_amdgcn_permlane16_swap(v, v, false, false); });
}

// dmax() and min() over lanes l and l^32, in every lane (one gfx950 row swap).
__device__ __forceinline__ double max_halves(double x) {
    const auto lo = __builtin_amdgcn_permlane32_swap(static_cast<unsigned>(__double2loint(x)),
                                                     static_cast<unsigned>(__double2loint(x)), false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap(static_cast<unsigned>(__double2hiint(x)),
                                                     static_cast<unsigned>(__double2hiint(x)), false, false);
    return dmax(__hiloint2double(static_cast<int>(hi[0]), static_cast<int>(lo[0])),
                __hiloint2double(static_cast<int>(hi[1]), static_cast<int>(lo[1])));
}
__device__ __forceinline__ int min_halves(int x) {
    const auto s = __builtin_amdgcn_permlane32_swap(static_cast<unsigned>(x), static_cast<unsigned>(x), false, false);
    return min(static_cast<int>(s[0]), static_cast<int>(s[1]));
}

// Maximum of N integers, three at a time.  The instruction is written out (as
// in dmax) so that the scheduler keeps the tree where the cells are, instead of
// stretching their registers' lives.
__device__ __forceinline__ int imax3(int a, int b, int c) {
    int r;
    asm("v_max3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}
template <int N, int K0 = 0>
__device__ __forceinline__ int hi_max_tree(const int* h) {
    if constexpr (N == 1) {
        return h[K0];
    } else if constexpr (N == 2) {
        return imax3(h[K0], h[K0 + 1], h[K0 + 1]);
    } else if constexpr (N == 3) {
        return imax3(h[K0], h[K0 + 1], h[K0 + 2]);
    } else {
        constexpr int A = (N + 2) / 3, B = (N - A + 1) / 2;
        return imax3(hi_max_tree<A, K0>(h), hi_max_tree<B, K0 + A>(h), hi_max_tree<N - A - B, K0 + A + B>(h));
    }
}

// Maximum of N consecutive cells starting at cell K0 (pairwise tree).
template <int N, int K0 = 0>
__device__ __forceinline__ double max_tree(const double* sc) {
    if constexpr (N == 1) {
        return sc[K0];
    } else {
        return dmax(max_tree<N / 2, K0>(sc), max_tree<N - N / 2, K0 + N / 2>(sc));
    }
}

// WRITEBACK: store each diagonal's final QT (the block drop-in's tmpQ); the
// self-join does not need it and skips the per-cell end-of-diagonal test.
// RING: the row record a lane takes in each step comes from a per-wave ring of
// records in LDS instead of from its neighbour lane by DPP (two LDS reads in
// place of eight DPP moves and eight fills per step).  The ring holds rows of
// the band indexed by row - dbase modulo RR = BW + 2F: the rows the current
// window reads and the next window's new ones; its first F slots are mirrored
// after the end so that a lane's reads within a window never wrap.  The F new
// rows of window w+1 reach the ring at the start of window w from a staging
// array, which a direct-to-LDS load filled one window earlier: no register
// holds them and no wait falls inside a window.  Lanes read slots 8 apart, so
// the ring is padded by 16 bytes after every 8 slots (slot s at byte
// 32 s + 16 (s / 8)): the 16 lanes of an LDS cycle then cover all 64 banks.
// FEEDLDS (rows by DPP only): the row entering lane 63 comes from LDS staging
// (below) instead of from a scalar load and eight moves into vector registers.
//
// SEGS: the items are column segments of the bands (self-join only, seg_r =
// SEG / BW, start values in ckpt); a segment that ends before its band does
// flushes the rows still in the band, whose other cells the next segment walks.
template <int K, bool WRITEBACK, bool RING, bool FEEDLDS, bool SEGS = false>
__device__ __forceinline__ void walk_body(
    const WalkRec* __restrict__ colrec, const WalkRec* __restrict__ rowrec,
    double* qt, unsigned long long* rowval, unsigned long long* rowkey,
    unsigned long long* colval, unsigned long long* colkey, unsigned int* counter,
    int64_t row_off, int64_t col_off, int blocklen, int tl, int m, int ib, int nbands,
    int seg_r = 0, const double* __restrict__ ckpt = nullptr) {
    // the ring form (one wave per SIMD, latency-bound) flushes every 32 steps,
    // the others every 16
    constexpr int F = RING ? 2 * kFlush : kFlush;
    constexpr int NQ = kWave / F;   // parts of a column's partials, one per flush lane
    constexpr int PQ = kWave / NQ;  // partials per flush lane
    static_assert(NQ == 2 || NQ == 4, "a column's partials reduce in halves or quarters");
    constexpr int BW = kWave * K;
    static_assert(F % K == 0, "flush period must be a multiple of K");
    static_assert(!(RING && FEEDLDS), "the ring form reads every row from LDS");
    static_assert(!SEGS || (!RING && !WRITEBACK), "segments: DPP rows, self-join only");
    // Column partials, one row of 64 per step, padded by one entry: a flush
    // lane reads PQ consecutive entries of one row (immediate offsets), and
    // the lanes of an LDS phase then fall on distinct banks.  Without the
    // ring, 42 KB of LDS per group, so that 3 groups still fit in a CU; with
    // it (32-step windows), 158 KB.
    __shared__ double s_cv[kWavesPerGroup][F][kWave + 1];
    // [0, F): rows leaving lane 0; lane l > 0 writes scratch slot F + l + step
    __shared__ double s_rv[kWavesPerGroup][2 * F + kWave];
    __shared__ int s_ri[kWavesPerGroup][2 * F + kWave];
    constexpr int RR = BW + 2 * F;  // ring slots (the mirror adds F)
    static_assert(RR % F == 0, "a window's new rows take F consecutive slots");
    static_assert(RR % 8 == 0 && (K * 8) % 8 == 0, "ring bases stay multiples of 8 slots");
    constexpr int kRingBytes = 32 * (RR + F) + 16 * ((RR + F) / 8);
    char* ringb = nullptr;
    WalkRec* stage = nullptr;
    unsigned long long* filt = nullptr;  // [0, F): column results, [F, 2F): row results
    // FEEDLDS: the rows entering lane 63 (one per step) come from a staging
    // array in LDS, filled by a direct-to-LDS load one window ahead (two
    // buffers, alternating by window): a step reads its row with two broadcast
    // LDS reads into the registers the DPP shift then writes, so lane 63's
    // fill costs no vector instruction.  The high word of -inf (lane 63's fill
    // of the running row maximum, in every form) is read from LDS the same way.
    const WalkRec* feed = nullptr;
    __shared__ int s_neghi[kWavesPerGroup];
    const int* const neghi = &s_neghi[threadIdx.x / kWave];
    if (threadIdx.x % kWave == 0) s_neghi[threadIdx.x / kWave] = static_cast<int>(0xFFF00000u);
    __builtin_amdgcn_wave_barrier();
    if constexpr (FEEDLDS) {
        __shared__ __align__(16) WalkRec s_feed[kWavesPerGroup][2 * F];
        feed = &s_feed[threadIdx.x / kWave][0];
    }
    if constexpr (RING) {
        __shared__ __align__(16) char s_ring[kWavesPerGroup][kRingBytes];
        __shared__ WalkRec s_stage[kWavesPerGroup][F];
        __shared__ unsigned long long s_filt[kWavesPerGroup][2 * F];
        ringb = &s_ring[threadIdx.x / kWave][0];
        stage = &s_stage[threadIdx.x / kWave][0];
        filt = &s_filt[threadIdx.x / kWave][0];
    }
    // byte offset of ring slot s, and the two 16-byte halves of a record there
    auto ring_off = [](int s) __attribute__((always_inline)) { return 32 * s + 16 * (s >> 3); };
    auto ring_put = [&](int s, const WalkRec& r) __attribute__((always_inline)) {
        double2* p = reinterpret_cast<double2*>(ringb + ring_off(s));
        p[0] = make_double2(r.a, r.am);
        p[1] = make_double2(r.s, r.mu);
    };

    const int wv = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const double NEG = -__builtin_inf();
    const double QNAN = __builtin_nan("");
    const int lid0 = lane * K;
    double* const cv_out = &s_cv[wv][0][lane];
    double* const rv_out = &s_rv[wv][(lane == 0) ? 0 : F + lane];
    int* const ri_out = &s_ri[wv][(lane == 0) ? 0 : F + lane];

    // Flush the results of columns [jstart, jstart+cnt) and of the rows that
    // left lane 0 at those columns.  Row lanes are stored as lane+1 (0 = none).
    // pc/pr: the current results of this window's columns and rows, read one
    // window ahead.  Results only grow, so a plain (possibly stale) read filters
    // out the partials that cannot win; a stale read only lets an unneeded
    // atomic through.  The test compares the truncated scores the keys hold, so
    // which keys compete does not depend on timing.  `reload` issues the next
    // window's reads after the tests and before the atomics: the loads then
    // reuse pc/pr's registers, and no load has to wait for an atomic.
    auto flush = [&](int jstart, int cnt, int dbase, uint64_t& pc, uint64_t& pr, auto reload)
        __attribute__((always_inline)) {
        __builtin_amdgcn_wave_barrier();
        // Lane (cs, qq) holds part qq of column cs's 64 partials (those of
        // lanes PQ qq .. PQ qq + PQ - 1).  A column result passes the filter iff
        // its score reaches thr, the score of the truncated current value (the
        // order map is monotone), so a column can pass only if one of its
        // partials does.  The reduction runs only when some partial of the
        // window passes.  For cs >= cnt the quarter is stale and unused.
        const int cs = lane & (F - 1), qq = lane / F;
        const double* const pv = &s_cv[wv][cs][qq * PQ];
        double v[PQ];
#pragma unroll
        for (int u = 0; u < PQ; ++u) v[u] = pv[u];
        if constexpr (RING) {  // the filter values, read behind the partials: one LDS wait for both
            __builtin_amdgcn_sched_barrier(0);
            pc = filt[lane & (F - 1)];
            pr = filt[F + (lane & (F - 1))];
        }
        const double thr = unord_f64((pc >> ib) << ib);
        // the quarter's exact maximum (a NaN partial is a masked cell, and
        // v_max_f64 returns the other operand; all NaN: NaN, which fails)
        double mx = max_tree<PQ>(v);
        const bool cand = (mx >= thr) && (cs < cnt);
        const int jj = jstart + lane;
        uint64_t oc = 0;
        ScoreId b{NEG, -1};
        if (__builtin_amdgcn_ballot_w64(cand) != 0) {
            // the column's maximum, then the smallest lane holding it: this
            // part's first, then the smallest over the parts
            mx = (NQ == 4) ? max_quarters(mx) : max_halves(mx);
            int id = kWave;
#pragma unroll
            for (int u = PQ - 1; u >= 0; --u) id = (v[u] == mx) ? qq * PQ + u : id;
            id = (NQ == 4) ? min_quarters(id) : min_halves(id);
            b = ScoreId{mx, id < kWave ? id : -1};
            oc = ord_f64(mx);
        }
        // (non-short-circuit forms: both tests are done before the reload)
        // a NaN result: every partial of the column was a masked cell
        const bool do_col = (lane < cnt) & (b.id >= 0) & (b.v == b.v) & ((oc >> ib) >= (pc >> ib));
        const double rvl = s_rv[wv][lane & (F - 1)];
        const int rlane = s_ri[wv][lane & (F - 1)] - 1;
        const uint64_t orw = ord_f64(rvl);
        const bool do_row = (lane < cnt) & (rlane >= 0) & ((orw >> ib) >= (pr >> ib));
        __builtin_amdgcn_sched_barrier(0);  // tests first: the reload then reuses pc/pr
        reload(pc, pr);
        // Key codes.  K = 1: the partner itself (the lane is the diagonal).
        // K > 1: the lane's diagonals are candidates, and the code holds the
        // partner of its first diagonal (p0) and the direction in which the
        // candidates run: p0 + k for a column result (rows jj + d), p0 - k for
        // a row result (columns jj - d); the unpack picks among them.
        int64_t ccode, rcode;
        if constexpr (K == 1) {
            ccode = static_cast<int64_t>(jj) + dbase + b.id + col_off;
            rcode = static_cast<int64_t>(jj) - rlane + row_off;
        } else {
            ccode = (static_cast<int64_t>(jj) + dbase + b.id * K + col_off) << 1;
            rcode = ((static_cast<int64_t>(jj) - rlane * K + row_off) << 1) | 1;
        }
        if (do_col) {
            atomicMax(&colval[jj], static_cast<unsigned long long>(oc));
            atomicMax(&colkey[jj], static_cast<unsigned long long>(make_key(oc, ib, ccode)));
        }
        if (do_row) {
            const int i = jj + dbase;
            atomicMax(&rowval[i], static_cast<unsigned long long>(orw));
            atomicMax(&rowkey[i], static_cast<unsigned long long>(make_key(orw, ib, rcode)));
        }
        __builtin_amdgcn_wave_barrier();
    };

    const int nitems = SEGS ? static_cast<int>(seg_items(nbands, seg_r)) : nbands;
    for (;;) {
        // One band per item, longest first (band 0 has the most columns):
        // waves that finish early take the remaining, shorter bands, so all
        // waves end within about one short band of each other.  SEGS: one
        // segment per item, band by band; columns [cb, L) of the band.
        const int item = next_item(counter, lane);
        if (item >= nitems) break;
        int band = item, cb = 0;
        if constexpr (SEGS) {
            band = seg_band_of(item, nbands, seg_r);
            cb = static_cast<int>(item - seg_first_item(band, nbands, seg_r)) * seg_r * BW;
        }
        const int dbase = band * BW;
        const int Lb = tl - dbase;  // the band's columns
        const int L = SEGS ? min(cb + seg_r * BW, Lb) : Lb;
        // Cells outside the triangle are masked by NaN: a row past the last
        // one is a padding record with sInv = NaN, a diagonal past the block
        // gets QT = NaN, and a NaN score never wins a comparison (nor does
        // it in the reference).
        // per row slot: rin = the row's maximum over the lanes above this
        // one (exact), tag = the lane that holds it + 1 (0: none), lm = the
        // maximum over this lane's cells of the row so far
        double q[K], wa[K], wam[K], ws[K], wm[K], rin[K], lm[K];
        int tag[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int d = dbase + lid0 + k;
            if constexpr (SEGS) {
                q[k] = (d < blocklen) ? ((cb == 0) ? qt[d] : ckpt[static_cast<int64_t>(item) * BW + lid0 + k])
                                      : QNAN;
            } else {
                q[k] = (d < blocklen) ? qt[d] : QNAN;
            }
            const WalkRec r = rowrec[d + cb];
            wa[k] = r.a;
            wam[k] = r.am;
            ws[k] = r.s;
            wm[k] = r.mu;
            rin[k] = NEG;
            lm[k] = NEG;
            tag[k] = 0;
            if constexpr (RING) {  // rows 0 .. BW-1 of the band, slot = row - dbase
                ring_put(lid0 + k, r);
                if (lid0 + k < F) ring_put(RR + lid0 + k, r);
            }
        }
        if constexpr (RING) {
            // rows BW .. BW+F-1 of the band to the ring, rows BW+F .. BW+2F-1
            // (the first window's new rows) to the staging array
            const WalkRec r = rowrec[dbase + BW + (lane & (F - 1))];
            const WalkRec r2 = rowrec[dbase + BW + F + (lane & (F - 1))];
            if (lane < F) {
                ring_put(BW + lane, r);
                stage[lane] = r2;
            }
        }

        // One column j = c0 + W for all K diagonals of every lane; W is the
        // step's place in the flush window, so the row slot (W mod K) and
        // the LDS offsets are constants.  TAIL: the window may end before
        // F steps.  Column records are a linear scalar stream.
        WalkRec cur_c, cur_f;  // the next step's column and feed records
        int cur_nh = 0;        // the next step's fill word
        int rbase = 0;         // ring byte offset of this window's first refill row
        const WalkRec* fw = feed;  // this window's staged feed rows (FEEDLDS)
        // A step's column partial and the row leaving lane 0 are written to
        // LDS by the next step, after its wait for the scalar records: the
        // wait (lgkmcnt, which also counts LDS writes) then never waits for
        // writes issued just before it.
        double pcv = NEG, prv = NEG;
        int pri = 0;
        auto put_partials = [&](int w) __attribute__((always_inline)) {
            cv_out[w * (kWave + 1)] = pcv;
            rv_out[w] = prv;
            ri_out[w] = pri;
        };
        auto step = [&](int c0, auto w_tag, auto tail_tag) __attribute__((always_inline)) -> bool {
            constexpr int W = decltype(w_tag)::value;
            constexpr int S = W % K;
            constexpr bool TAIL = decltype(tail_tag)::value;
            const int j = c0 + W;
            if (TAIL && j >= L) return false;
            // The column's record (a scalar load) and the row entering
            // lane 63 (a padding record past the triangle; a scalar load,
            // or an LDS read of the staged rows; the ring path reads rows
            // from LDS instead) were loaded one step ago; the next step's are loaded now, so
            // the loads overlap a step.  Steps are not interleaved by the
            // scheduler: that keeps each step's use of its records (and so
            // the wait for them) after the previous step's loads have had
            // a whole step to land.
            __builtin_amdgcn_sched_barrier(0);
            const WalkRec c = cur_c, f = cur_f;
            const int fnh = cur_nh;
            // Scalar loads may return out of order, so a wait for this step's
            // records waits for every scalar and LDS load in flight: take
            // the wait here (the empty asm needs the values), before the
            // next step's loads are issued (its memory clobber keeps them
            // below).
            if constexpr (RING || FEEDLDS) {
                asm volatile("" ::"s"(c.a), "s"(c.am), "s"(c.s), "s"(c.mu) : "memory");
            } else {
                asm volatile("" ::"s"(c.a), "s"(c.am), "s"(c.s), "s"(c.mu), "s"(f.a), "s"(f.am),
                             "s"(f.s), "s"(f.mu)
                             : "memory");
            }
            cur_c = colrec[j + 1];
            if constexpr (!RING && !FEEDLDS) cur_f = rowrec[j + 1 + dbase + BW];
            // the last step's successor is the next window's first step,
            // which reads these when the window starts
            if constexpr (W + 1 < F) {
                if constexpr (FEEDLDS) cur_f = fw[W + 1];
                cur_nh = *neghi;
            }
            // the row this lane takes at the end of the step: row
            // c0 + W + dbase + (lane+1)*K, which lane+1 finishes now; read
            // before the step's LDS writes, so the read overlaps the step
            WalkRec nr;
            if constexpr (RING) {
                // slot rbase + W, rbase a multiple of 8: a constant offset
                const double2* p =
                    reinterpret_cast<const double2*>(ringb + rbase + 32 * W + 16 * (W >> 3));
                const double2 lo = p[0], hi = p[1];
                nr.a = lo.x;
                nr.am = lo.y;
                nr.s = hi.x;
                nr.mu = hi.y;
            }
            if constexpr (W > 0) put_partials(W - 1);
            __builtin_amdgcn_sched_barrier(0);  // issue them before the step's work
            const double ahj = c.a, ahjm = c.am, shj = c.s, mhj = c.mu;
            double sc[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int sl = (S + k) % K;
                sc[k] = fma(q[k], ws[sl] * shj, -(wm[sl] * mhj));
                q[k] = fma(wam[sl], ahjm, q[k]);
                q[k] = fma(-wa[sl], ahj, q[k]);
                if constexpr (WRITEBACK) {
                    if (j == tl - 1 - (dbase + lid0 + k) && dbase + lid0 + k < blocklen)
                        qt[dbase + lid0 + k] = q[k];
                }
                // the row in slot (S + K - 1) % K entered this lane at the end
                // of the last step: its first cell here starts lm
                lm[sl] = (k == K - 1) ? sc[k] : dmax(lm[sl], sc[k]);
            }
            pcv = max_tree<K>(sc);
            // the row in slot S leaves this lane: its maximum so far, and
            // this lane as the argmax's when its own cells beat the lanes
            // above (strictly: the row met those lanes first, at smaller
            // columns); a NaN lm (only masked cells) changes nothing
            prv = dmax(rin[S], lm[S]);
            pri = (lm[S] > rin[S]) ? lane + 1 : tag[S];
            rin[S] = shift_down_f64_lo0(prv, fnh);  // -inf into lane 63
            tag[S] = shift_down_i32_zero(pri);
            if constexpr (RING) {
                wa[S] = nr.a;
                wam[S] = nr.am;
                ws[S] = nr.s;
                wm[S] = nr.mu;
            } else {
                // FEEDLDS: f is in vector registers already; else one
                // 64-bit move per double brings it there
                if constexpr (FEEDLDS) {
                    wa[S] = shift_down_f64(wa[S], f.a);
                    wam[S] = shift_down_f64(wam[S], f.am);
                    ws[S] = shift_down_f64(ws[S], f.s);
                    wm[S] = shift_down_f64(wm[S], f.mu);
                } else {
                    wa[S] = shift_down_f64(wa[S], vmov64(f.a));
                    wam[S] = shift_down_f64(wam[S], vmov64(f.am));
                    ws[S] = shift_down_f64(ws[S], vmov64(f.s));
                    wm[S] = shift_down_f64(wm[S], vmov64(f.mu));
                }
            }
            return true;
        };
        auto window = [&](int c0, auto tail_tag) __attribute__((always_inline)) {
            if constexpr (RING) {
                // The next window's new rows, band rows c0 + F + BW + x, go
                // from the staging array to slots (c0 + F + BW + x) mod RR,
                // which held rows c0 - F + x: finished in the last window.
                __builtin_amdgcn_wave_barrier();
                if (lane < F) {
                    const WalkRec t = stage[lane];
                    const int slot = (c0 + F + BW + lane) % RR;
                    ring_put(slot, t);
                    if (slot < F) ring_put(RR + slot, t);
                }
                // then the staging array takes band rows c0 + 2F + BW + x
                // (padding records past the triangle), 16 bytes per lane;
                // the staging read above has completed before the wave goes on
                __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
                if (lane < 2 * F) {
                    const char* src = reinterpret_cast<const char*>(
                                          rowrec + (dbase + c0 + 2 * F + BW + (lane >> 1))) +
                                      16 * (lane & 1);
                    __builtin_amdgcn_global_load_lds(
                        (const __attribute__((address_space(1))) void*)(src),
                        (__attribute__((address_space(3))) void*)(stage), 16, 0, 0);
                    // the current results of this window's columns and rows,
                    // for the flush's read filter (4 bytes per lane, clamped)
                    const int jj = min(c0 + (lane >> 1), L - 1);
                    const char* cs = reinterpret_cast<const char*>(colval + jj) + 4 * (lane & 1);
                    const char* rs = reinterpret_cast<const char*>(rowval + jj + dbase) + 4 * (lane & 1);
                    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(cs),
                                                     (__attribute__((address_space(3))) void*)(filt), 4, 0, 0);
                    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(rs),
                                                     (__attribute__((address_space(3))) void*)(filt + F), 4,
                                                     0, 0);
                }
                __builtin_amdgcn_wave_barrier();
                rbase = ring_off((c0 + lid0 + K) % RR);  // a byte offset
                cur_nh = *neghi;
            } else if constexpr (!FEEDLDS) {
                cur_nh = *neghi;
            } else {
                // The next window's feed rows (band rows c0 + F + BW + x,
                // padding records past the triangle) go to the other
                // buffer, 16 bytes per lane; this window's landed before
                // the last flush.
                __builtin_amdgcn_wave_barrier();
                const int nb = (c0 / F) & 1;
                if (lane < 2 * F) {
                    const char* src = reinterpret_cast<const char*>(
                                          rowrec + (dbase + c0 + F + BW + (lane >> 1))) +
                                      16 * (lane & 1);
                    __builtin_amdgcn_global_load_lds(
                        (const __attribute__((address_space(1))) void*)(src),
                        (__attribute__((address_space(3))) void*)(feed + (nb ^ 1) * F), 16, 0, 0);
                }
                fw = feed + nb * F;
                cur_f = fw[0];
                cur_nh = *neghi;
                __builtin_amdgcn_wave_barrier();
            }
            bool go = true;
            [&]<int... W>(std::integer_sequence<int, W...>) __attribute__((always_inline)) {
                ((go = go && step(c0, std::integral_constant<int, W>{}, tail_tag)), ...);
            }(std::make_integer_sequence<int, F>{});
            // the window's last step writes no successor's partials
            if constexpr (decltype(tail_tag)::value) {
                put_partials(min(F, L - c0) - 1);
            } else {
                put_partials(F - 1);
            }
        };
        // results of column c0 + (lane mod F) and of its row, clamped in bounds
        auto current = [&](int c0, uint64_t& pc, uint64_t& pr) __attribute__((always_inline)) {
            const int jj = min(c0 + (lane & (F - 1)), L - 1);
            pc = colval[jj];
            pr = rowval[jj + dbase];
        };
        int c0 = cb;  // a multiple of 2F: the first window reads feed buffer 0
        uint64_t pc = 0, pr = 0;
        if constexpr (!RING) current(cb, pc, pr);
        if constexpr (FEEDLDS) {
            // the first window's feed rows (band rows cb + BW + x) to buffer 0
            __builtin_amdgcn_wave_barrier();
            if (lane < 2 * F) {
                const char* src =
                    reinterpret_cast<const char*>(rowrec + (dbase + cb + BW + (lane >> 1))) + 16 * (lane & 1);
                __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(src),
                                                 (__attribute__((address_space(3))) void*)(feed), 16, 0, 0);
            }
            __builtin_amdgcn_wave_barrier();
        }
        cur_c = colrec[cb];
        if constexpr (!RING && !FEEDLDS) cur_f = rowrec[dbase + cb + BW];
        // let the band's set-up loads land here, so the windows never wait on
        // vector memory (whose counter would also cover the flush's atomics)
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
        // whole flush windows: no control flow inside a window
        for (; c0 + F <= L; c0 += F) {
            window(c0, std::false_type{});
            // the direct-to-LDS loads of this window (the next rows, and
            // with the ring the filter values) must have landed: the
            // compiler does not order LDS reads after them
            if constexpr (RING || FEEDLDS) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            if constexpr (RING) {
                flush(c0, F, dbase, pc, pr, [](uint64_t&, uint64_t&) {});
            } else {
                flush(c0, F, dbase, pc, pr, [&](uint64_t& a, uint64_t& b) __attribute__((always_inline)) {
                    current(c0 + F, a, b);
                });
            }
        }
        // the last partial window
        if (c0 < L) {
            window(c0, std::true_type{});
            if constexpr (RING || FEEDLDS) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            flush(c0, L - c0, dbase, pc, pr, [](uint64_t&, uint64_t&) {});
        }
        if constexpr (SEGS) {
            // A segment that ends before its band (L a multiple of K): slot k
            // of each lane holds row L + dbase + lid0 + k.  Slot K-1's row has
            // just entered the lane (no cell here yet: only rin counts); the
            // others' lm holds this lane's cells of them so far.  Their
            // maxima so far are flushed as the rows leaving lane 0 are (the
            // code names the block of the lane that holds the maximum).
            if (L < Lb) {
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const bool own = (k < K - 1) && (lm[k] > rin[k]);
                    const double v = (k < K - 1) ? dmax(rin[k], lm[k]) : rin[k];
                    const int t = own ? lane + 1 : tag[k];
                    const int jj = L + lid0 + k;  // the column at which the row leaves lane 0
                    const int i = jj + dbase;
                    if (t > 0 && i < tl && v == v) {
                        const uint64_t orw = ord_f64(v);
                        int64_t rcode;
                        if constexpr (K == 1) {
                            rcode = static_cast<int64_t>(jj) - (t - 1) + row_off;
                        } else {
                            rcode = ((static_cast<int64_t>(jj) - (t - 1) * K + row_off) << 1) | 1;
                        }
                        atomicMax(&rowval[i], static_cast<unsigned long long>(orw));
                        atomicMax(&rowkey[i], static_cast<unsigned long long>(make_key(orw, ib, rcode)));
                    }
                }
            }
        }
    }
}

// Lane l receives lane l-1's value, lane 0 lane 63's (DPP wave_ror:1).
// Every lane has a source lane, so no lane keeps an old value.
__device__ __forceinline__ double rotate_up_f64(double src) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(src), 0x13C, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(src), 0x13C, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}

// Rotating band walker: the rows stay in place and the diagonals move.  The
// band's diagonals, cells, record streams, masking and result merging are
// those of walk_body; what differs is which lane holds what.  Band row b (row
// dbase + b) sits at band position b mod BW, in lane (b mod BW) / K, slot
// b mod K, for its whole time in the band: at column j the wave holds rows
// j .. j+BW-1, and the row that has its last cell at column j (b = j, diagonal
// dbase) is replaced in place by row j+BW when the step ends.  Diagonal
// dbase + e is at position (j + e) mod BW at column j, so each QT value moves
// one position per step: within a lane by the choice of register (register t
// holds the diagonal at slot (t + j) mod K), and from lane l's last slot to
// lane l+1's first (lane 63's to lane 0's) by one DPP rotation of one double.
// A step thus moves 2 dwords between lanes where walk_body moves a row record,
// its running maximum and its lane tag.
//
// Rows.  A row's maximum never leaves its lane: slot s keeps the maximum of the
// current block of K steps (lm) and of the row's finished blocks (run), and
// the last column of the block that holds run (tag: the first such block, as
// blocks end in column order and only a strictly larger maximum replaces it).
// Blocks of slot s end at the columns j = s (mod K); the leaving row's block
// ends at its last column, where its lane alone (exec-masked LDS operations)
// writes (run, tag) for the flush and takes the entering row: run becomes a
// quiet NaN (its high word set), which every block maximum replaces, and the
// row's record comes from the LDS staging of FEEDLDS.  A row result's key
// names the K columns of the block, tag - k, and the unpack picks among them.
//
// Columns.  At column j the lane holding position j mod BW (lane wl, wrap slot
// S = j mod K) holds rows j .. j+K-1-S in slots S .. K-1 and the band's last
// rows j+BW-S .. j+BW-1 in slots 0 .. S-1; every other lane holds K
// consecutive rows, those of lane wl + r starting at j + K r - S.  Every lane
// writes its maximum at index r = (l - wl) mod 64, so the partials of a column
// lie in ascending row order, except that the wrap lane writes the maximum
// over slots S .. K-1 at index 0 and the one over slots 0 .. S-1 (the band's
// largest rows) as a 65th partial at index 64: the flush's smallest index
// holding the maximum is then the smallest row, as the keys rank exact ties,
// and the key names the K rows from that partial's first (index 64: the K rows
// ending at j+BW-1).
//
// Column filter.  Once the profile fills, a band's cells rarely beat a
// column's current result (at n = 2^22 a column has a candidate in 0.06% of
// the windows), so a step first tests its cells against that result, read by
// a scalar load one step ahead through colthr (colval again, read-only; the
// scalar cache may return an older value, which is lower, as results only
// grow): the maximum of the cells' high words (v_max3_i32) against the
// result's high word.  Only when
// some lane passes does the step build its pairs (six FP64 maxima) and write
// them, marking the column in a window mask; the flush reduces marked columns
// only, and most windows reduce none.
//
// A step has one exec-masked region (each costs a branch) and one uniform
// branch (the column filter).
template <int K>
__device__ __forceinline__ void walk_rot_body(const WalkRec* __restrict__ colrec,
                                              const WalkRec* __restrict__ rowrec, const double* qt,
                                              unsigned long long* rowval, unsigned long long* rowkey,
                                              unsigned long long* colval, unsigned long long* colkey,
                                              const unsigned long long* __restrict__ colthr,
                                              unsigned int* counter, int64_t row_off, int64_t col_off,
                                              int blocklen, int tl, int ib, int nbands) {
    // 32 steps per flush window (the other walkers use 16): with the column
    // filter most flushes only merge rows, so fewer of them save their fixed
    // cost, and one partial per lane (below) keeps the LDS for two groups per CU
    constexpr int F = 2 * kFlush;
    constexpr int BW = kWave * K;
    // a window starts at wrap slot 0 and its positions never pass the wave's end
    static_assert(F % K == 0 && BW % F == 0, "window and band geometry");
    static_assert(F == kWave / 2, "the flush reduces each column in two halves of the wave");
    constexpr int Q = kWave / 2;  // partials per flush lane
    struct RowRes {
        double v;
        int tag, pad;
    };
    // Column partials: index r = (l - wl) mod 64 holds lane l's maximum, except
    // that the wrap lane (r = 0) keeps its two halves apart, the second at
    // index 64.  Rows of 66 doubles.
    __shared__ double s_cv[kWavesPerGroup][F][kWave + 2];
    __shared__ RowRes s_rr[kWavesPerGroup][F];  // the leaving row's result per step
    __shared__ __align__(16) WalkRec s_feed[kWavesPerGroup][2 * F];

    const int wv = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const double NEG = -__builtin_inf();
    const double QNAN = __builtin_nan("");
    const int lid0 = lane * K;
    WalkRec* const feed = &s_feed[wv][0];
    // the high words of the column results (little-endian: word 1 of each)
    const uint32_t* const colthr_hi = reinterpret_cast<const uint32_t*>(colthr) + 1;

    // Flush columns [jstart, jstart+cnt) and the rows that left at those
    // columns, as walk_body's flush does; lane (cs, qq) reduces half of column
    // cs's partials and the 65th.
    auto flush = [&](int jstart, int cnt, int dbase, uint64_t& pc, uint64_t& pr, unsigned colmask, auto reload)
        __attribute__((always_inline)) {
        __builtin_amdgcn_wave_barrier();
        const int cs = lane & (F - 1), qq = lane / F;
        const int jj = jstart + lane;
        uint64_t oc = 0;
        ScoreId b{NEG, -1};
        // only the columns whose step wrote its partials (colmask) can hold a
        // candidate; the others' LDS rows are stale
        if (colmask != 0) {
            const double* const pv = &s_cv[wv][cs][qq * Q];
            double v[Q + 1];
#pragma unroll
            for (int u = 0; u < Q; ++u) v[u] = pv[u];
            v[Q] = s_cv[wv][cs][kWave];
            const double thr = unord_f64((pc >> ib) << ib);
            double mx = max_tree<Q + 1>(v);
            const bool cand = (mx >= thr) && (cs < cnt) && ((colmask >> cs) & 1u);
            if (__builtin_amdgcn_ballot_w64(cand) != 0) {
                mx = max_halves(mx);
                int id = (v[Q] == mx) ? kWave : kWave + 1;
#pragma unroll
                for (int u = Q - 1; u >= 0; --u) id = (v[u] == mx) ? qq * Q + u : id;
                id = min_halves(id);
                b = ScoreId{mx, id <= kWave ? id : -1};
                oc = ord_f64(mx);
            }
        }
        // (b of an unflagged column was taken from stale partials)
        const bool do_col = (lane < cnt) & ((colmask >> cs) & 1u) & (b.id >= 0) & (b.v == b.v) &
                            ((oc >> ib) >= (pc >> ib));
        const RowRes rr = s_rr[wv][lane & (F - 1)];
        const uint64_t orw = ord_f64(rr.v);
        // a row whose cells were all masked keeps -inf or the reset NaN
        const bool do_row = (lane < cnt) & (rr.v > NEG) & ((orw >> ib) >= (pr >> ib));
        __builtin_amdgcn_sched_barrier(0);
        reload(pc, pr);
        if (do_col) {
            // first row of the winning partial; index kWave: the K rows ending at jj+BW-1
            const int S = jj & (K - 1);
            const bool last = b.id == kWave;
            const int64_t b0 = last ? static_cast<int64_t>(jj) + BW - 1
                                    : static_cast<int64_t>(jj) + max(K * b.id - S, 0);
            const int64_t ccode = ((b0 + dbase + col_off) << 1) | (last ? 1 : 0);
            atomicMax(&colval[jj], static_cast<unsigned long long>(oc));
            atomicMax(&colkey[jj], static_cast<unsigned long long>(make_key(oc, ib, ccode)));
        }
        if (do_row) {
            const int i = jj + dbase;
            const int64_t rcode = ((static_cast<int64_t>(rr.tag) + row_off) << 1) | 1;
            atomicMax(&rowval[i], static_cast<unsigned long long>(orw));
            atomicMax(&rowkey[i], static_cast<unsigned long long>(make_key(orw, ib, rcode)));
        }
        __builtin_amdgcn_wave_barrier();
    };

    for (;;) {
        // One band per item, longest first (band 0 has the most columns):
        // waves that finish early take the remaining, shorter bands, so all
        // waves end within about one short band of each other.
        const int band = next_item(counter, lane);
        if (band >= nbands) break;
        const int dbase = band * BW;
        const int L = tl - dbase;
        double qr[K], wa[K], wam[K], ws[K], wm[K], lm[K], run[K];
        int tag[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int d = dbase + lid0 + k;  // at column 0 position e holds diagonal and row e
            qr[k] = (d < blocklen) ? qt[d] : QNAN;
            const WalkRec r = rowrec[d];
            wa[k] = r.a;
            wam[k] = r.am;
            ws[k] = r.s;
            wm[k] = r.mu;
            lm[k] = NEG;
            run[k] = NEG;
            tag[k] = -1;
        }
        WalkRec cur_c;             // the next step's column record
        uint32_t cur_t;            // and the high word of its current result (ord)
        unsigned colmask = 0;      // the window's steps that wrote their column partials
        const WalkRec* fw = feed;  // this window's entering rows
        int wl0 = 0;               // the wrap lane of the window's first column
        // this lane's rotated index, per half window (its wrap lane)
        int rot = 0;
        auto half = [&](int wl) __attribute__((always_inline)) { rot = (lane - wl) & (kWave - 1); };
        auto step = [&](int c0, auto w_tag, auto tail_tag) __attribute__((always_inline)) -> bool {
            constexpr int W = decltype(w_tag)::value;
            constexpr int S = W % K;             // the wrap slot: its row leaves after this step
            constexpr int S0 = (W + K - 1) % K;  // the slot whose block starts at this step
            constexpr bool TAIL = decltype(tail_tag)::value;
            const int j = c0 + W;
            if (TAIL && j >= L) return false;
            __builtin_amdgcn_sched_barrier(0);
            // Scalar loads may return out of order, so the wait for this
            // step's record (forced by the empty asm, which needs the
            // values) covers every scalar and LDS operation in flight; the
            // next step's record is loaded after it.
            const WalkRec c = cur_c;
            const uint32_t th = cur_t;
            asm volatile("" ::"s"(c.a), "s"(c.am), "s"(c.s), "s"(c.mu), "s"(th) : "memory");
            cur_c = colrec[j + 1];
            cur_t = colthr_hi[2 * (j + 1)];  // j + 1 <= L <= tl: the self-join's P > W entries
            if constexpr (W > 0 && S == 0) half(wl0 + W / K);
            __builtin_amdgcn_sched_barrier(0);
            const double ahj = c.a, ahjm = c.am, shj = c.s, mhj = c.mu;
            double sc[K];
            auto cell = [&](auto s_tag) __attribute__((always_inline)) {
                constexpr int s = decltype(s_tag)::value;
                constexpr int t = (s - S + K) % K;  // the register of slot s's diagonal
                sc[s] = fma(qr[t], ws[s] * shj, -(wm[s] * mhj));
                qr[t] = fma(wam[s], ahjm, qr[t]);
                qr[t] = fma(-wa[s], ahj, qr[t]);
                lm[s] = (s == S0) ? sc[s] : dmax(lm[s], sc[s]);
            };
            // the leaving row first, so that its successor's record loads
            // while the other slots compute
            cell(std::integral_constant<int, S>{});
            // The QT updates read the rows: the compiler would sink them to
            // their next use, the next step, past the masked region that
            // replaces a row, and copy that row's registers to keep them.
            // Empty asm statements that take the updated values keep them
            // here (the leaving slot's) and in the step (the others).
            asm volatile("" ::"v"(qr[0]));
            tag[S] = !(lm[S] <= run[S]) ? j : tag[S];
            run[S] = dmax(run[S], lm[S]);
            if (__builtin_expect(rot == 0, 1)) {
                s_rr[wv][W].v = run[S];
                s_rr[wv][W].tag = tag[S];
                run[S] = __hiloint2double(static_cast<int>(0xFFF80000u), __double2loint(run[S]));
                const double2* p = reinterpret_cast<const double2*>(fw + W);
                const double2 lo = p[0], hi = p[1];
                wa[S] = lo.x;
                wam[S] = lo.y;
                ws[S] = hi.x;
                wm[S] = hi.y;
            }
            [&]<int... s>(std::integer_sequence<int, s...>) __attribute__((always_inline)) {
                ((s != S ? cell(std::integral_constant<int, s>{}) : void()), ...);
            }(std::make_integer_sequence<int, K>{});
            // slot K-1's diagonal moves to the next lane's slot 0
            constexpr int tl7 = (K - 1 - S + K) % K;
            qr[tl7] = rotate_up_f64(qr[tl7]);
#pragma unroll
            for (int t = 1; t < K; ++t) asm volatile("" ::"v"(qr[t]));
            // The column's partials are needed only if some cell can reach
            // the column's current result.  A test on the cells' high words,
            // as signed integers, passes every such cell: for positive scores
            // they order as the scores do (ties included, as the result keys
            // truncate only low-word bits: ib <= 32 for P <= 2^30), a
            // negative score is a negative integer, and a masked cell's NaN
            // passes (the partials then take care of it).  A result that is
            // not positive (none yet, or the -1 it starts from) passes every
            // cell.
            const int sgn = static_cast<int>(th) >> 31;  // -1: a positive score
            const int thi = (static_cast<int>(th & 0x7FFFFFFFu) & sgn) | (~sgn & INT_MIN);
            int hv[K];
#pragma unroll
            for (int s = 0; s < K; ++s) hv[s] = __double2hiint(sc[s]);
            if (__builtin_amdgcn_ballot_w64(hi_max_tree<K>(hv) >= thi) != 0) {
                const double pav = max_tree<K - S, S>(sc);
                const double pbv = (S == 0) ? NEG : max_tree<(S > 0 ? S : 1), 0>(sc);
                s_cv[wv][W][rot] = (rot == 0) ? pav : dmax(pav, pbv);
                if (rot == 0) s_cv[wv][W][kWave] = pbv;
                colmask |= 1u << W;
            }
            return true;
        };
        // rows dbase + c0 + BW + x, x < F (padding records past the
        // triangle), to buffer nb: lane l < 2F loads half l % 2 of row l / 2
        auto stage = [&](int c0, int nb) __attribute__((always_inline)) {
            if (lane < 2 * F) {
                const char* src =
                    reinterpret_cast<const char*>(rowrec + (dbase + c0 + BW + (lane >> 1))) + 16 * (lane & 1);
                __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(src),
                                                 (__attribute__((address_space(3))) void*)(feed + nb * F), 16,
                                                 0, 0);
            }
        };
        auto window = [&](int c0, auto tail_tag) __attribute__((always_inline)) {
            // the next window's entering rows (band rows c0 + F + BW + x) to
            // the other buffer, 16 bytes per lane; this window's landed
            // before the last flush
            __builtin_amdgcn_wave_barrier();
            const int nb = (c0 / F) & 1;
            stage(c0 + F, nb ^ 1);
            fw = feed + nb * F;
            wl0 = (c0 % BW) / K;
            half(wl0);
            __builtin_amdgcn_wave_barrier();
            bool go = true;
            [&]<int... W>(std::integer_sequence<int, W...>) __attribute__((always_inline)) {
                ((go = go && step(c0, std::integral_constant<int, W>{}, tail_tag)), ...);
            }(std::make_integer_sequence<int, F>{});
        };
        auto current = [&](int c0, uint64_t& pc, uint64_t& pr) __attribute__((always_inline)) {
            const int jj = min(c0 + (lane & (F - 1)), L - 1);
            pc = colval[jj];
            pr = rowval[jj + dbase];
        };
        int c0 = 0;
        uint64_t pc = 0, pr = 0;
        current(0, pc, pr);
        // the first window's entering rows (band rows BW + x) to buffer 0
        __builtin_amdgcn_wave_barrier();
        stage(0, 0);
        __builtin_amdgcn_wave_barrier();
        cur_c = colrec[0];
        cur_t = colthr_hi[0];
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
        for (; c0 + F <= L; c0 += F) {
            colmask = 0;
            window(c0, std::false_type{});
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            flush(c0, F, dbase, pc, pr, colmask, [&](uint64_t& a, uint64_t& b) __attribute__((always_inline)) {
                current(c0 + F, a, b);
            });
        }
        if (c0 < L) {
            colmask = 0;
            window(c0, std::true_type{});
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            flush(c0, L - c0, dbase, pc, pr, colmask, [](uint64_t&, uint64_t&) {});
        }
    }
}

#define MP_WALK_PARAMS                                                                       \
    const WalkRec *__restrict__ colrec, const WalkRec *__restrict__ rowrec, double *qt,          \
        unsigned long long *rowval, unsigned long long *rowkey, unsigned long long *colval,     \
        unsigned long long *colkey, unsigned int *counter, int64_t row_off, int64_t col_off,    \
        int blocklen, int tl, int m, int ib, int nbands
#define MP_WALK_ARGS                                                                           \
    colrec, rowrec, qt, rowval, rowkey, colval, colkey, counter, row_off, col_off, blocklen, tl, \
        m, ib, nbands

// K <= 4: at most 168 VGPRs, so that the 3 groups the LDS allows per CU all fit.
// Its feed rows come from LDS staging (K = 1 at n = 65,536: 9.1 -> 7.4 ms).
template <int K, bool WRITEBACK>
__global__ __launch_bounds__(kGroup) __attribute__((amdgpu_waves_per_eu(3))) void k_walk_f64(
    MP_WALK_PARAMS) {
    walk_body<K, WRITEBACK, false, true>(MP_WALK_ARGS);
}
// Start values of the column segments: diagonal d of band b advanced from its
// seed at column 0 to each segment start s*SEG (s = 1 .. n_b - 1) by the
// walker's own recurrence, the same two FMAs in the same order, so that a
// segment continues exactly as the whole band's walk would.  One lane per
// diagonal, one wave per 64 diagonals d0 .. d0+63 of a band.  Each 64-step tile
// (columns j0 .. j0+63) stages its column records and the rows it meets, j0 + d0
// .. j0 + d0 + 127, in LDS one tile ahead; at step t lane l reads row t + l of
// the tile (consecutive lanes, consecutive entries) and the column by
// broadcast, one group of 8 steps ahead, so a step is its two FMAs.
__global__ __launch_bounds__(256) void k_seg_seeds(const WalkRec* __restrict__ colrec,
                                                   const WalkRec* __restrict__ rowrec,
                                                   const double* __restrict__ qt, double* __restrict__ ckpt,
                                                   int tl, int blocklen, int bw, int nbands, int seg_r) {
    constexpr int T = kWave;  // steps per tile
    __shared__ double2 s_row[4][2][2 * T], s_col[4][2][T];
    const int wv = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int d0 = (blockIdx.x * 4 + wv) * kWave;
    const int b = d0 / bw;
    if (b >= nbands || d0 >= blocklen) return;
    const int seg = seg_r * bw;
    const int nseg = (tl - b * bw + seg - 1) / seg;
    if (nseg <= 1) return;
    const int d = d0 + lane;
    double* const out = ckpt + seg_first_item(b, nbands, seg_r) * bw + (d - b * bw);
    const int steps = (nseg - 1) * seg;  // a multiple of T
    // (a, am) of record r: the first 16 bytes
    auto ld = [](const WalkRec* r) __attribute__((always_inline)) {
        return *reinterpret_cast<const double2*>(r);
    };
    double q = (d < blocklen) ? qt[d] : 0.0;
    // tile 0 to buffer 0
    s_row[wv][0][lane] = ld(rowrec + d0 + lane);
    s_row[wv][0][T + lane] = ld(rowrec + d0 + T + lane);
    s_col[wv][0][lane] = ld(colrec + lane);
    __builtin_amdgcn_wave_barrier();
    for (int j0 = 0; j0 < steps; j0 += T) {
        const int buf = (j0 / T) & 1;
        // the next tile's records, written to the other buffer after this tile
        double2 n0 = make_double2(0.0, 0.0), n1 = n0, nc = n0;
        const bool more = j0 + T < steps;
        if (more) {
            n0 = ld(rowrec + j0 + T + d0 + lane);
            n1 = ld(rowrec + j0 + 2 * T + d0 + lane);
            nc = ld(colrec + j0 + T + lane);
        }
        const double2* const rb = &s_row[wv][buf][lane];
        const double2* const cb = s_col[wv][buf];
        constexpr int U = 8;
        double2 rA[U], cA[U], rB[U], cB[U];
        auto rd = [&](int t0, double2(&r)[U], double2(&c)[U]) __attribute__((always_inline)) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                r[u] = rb[t0 + u];
                c[u] = cb[t0 + u];
            }
        };
        auto run = [&](const double2(&r)[U], const double2(&c)[U]) __attribute__((always_inline)) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                q = fma(r[u].y, c[u].y, q);   // + am(row) * am(column)
                q = fma(-r[u].x, c[u].x, q);  // - a(row) * a(column)
            }
        };
        rd(0, rA, cA);
#pragma unroll
        for (int t0 = 0; t0 < T; t0 += 2 * U) {
            rd(t0 + U, rB, cB);
            __builtin_amdgcn_sched_barrier(0);
            run(rA, cA);
            __builtin_amdgcn_sched_barrier(0);
            if (t0 + 2 * U < T) rd(t0 + 2 * U, rA, cA);
            __builtin_amdgcn_sched_barrier(0);
            run(rB, cB);
            __builtin_amdgcn_sched_barrier(0);
        }
        if ((j0 + T) % seg == 0 && d < blocklen) out[static_cast<int64_t>((j0 + T) / seg) * bw] = q;
        __builtin_amdgcn_wave_barrier();
        if (more) {
            s_row[wv][buf ^ 1][lane] = n0;
            s_row[wv][buf ^ 1][T + lane] = n1;
            s_col[wv][buf ^ 1][lane] = nc;
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// Column segments of k_walk_f64's bands (see walk_body); seg_r = SEG / BW.
template <int K>
__global__ __launch_bounds__(kGroup) __attribute__((amdgpu_waves_per_eu(3))) void k_walk_f64_seg(
    MP_WALK_PARAMS, int seg_r, const double* __restrict__ ckpt) {
    walk_body<K, false, false, true, true>(MP_WALK_ARGS, seg_r, ckpt);
}
// K = 8 holds twice the row records (one wave per SIMD); its rows come from the
// LDS ring, which the one group per CU this leaves has room for.
template <int K, bool WRITEBACK>
__global__ __launch_bounds__(kGroup) void k_walk_f64_wide(MP_WALK_PARAMS) {
    walk_body<K, WRITEBACK, true, false>(MP_WALK_ARGS);
}
// K = 8 with its rows shifted down the lanes by DPP instead of the LDS ring:
// 38 KB of LDS per group and at most 256 VGPRs, so two waves share each SIMD
// and hide each other's latency; it pays when there are work items for two
// waves per SIMD.  Its feed rows stay scalar loads: staged in LDS (fewer
// vector instructions) the kernel measured 8% slower at n = 2^22, m = 1024.
template <int K, bool WRITEBACK>
__global__ __launch_bounds__(kGroup) __attribute__((amdgpu_waves_per_eu(2))) void k_walk_f64_pair(
    MP_WALK_PARAMS) {
    walk_body<K, WRITEBACK, false, false>(MP_WALK_ARGS);
}
// K = 8 with rows in place and the diagonals rotating (walk_rot_body); the
// self-join only (no final-QT write-back).
template <int K>
__global__ __launch_bounds__(kGroup) __attribute__((amdgpu_waves_per_eu(2))) void k_walk_f64_rot(
    MP_WALK_PARAMS, const unsigned long long* __restrict__ colthr) {
    (void)m;
    walk_rot_body<K>(colrec, rowrec, qt, rowval, rowkey, colval, colkey, colthr, counter, row_off, col_off,
                     blocklen, tl, ib, nbands);
}
#undef MP_WALK_PARAMS
#undef MP_WALK_ARGS

}  // namespace mp

// ---------------------------------------------------------------------------
// Scratch of the allocation-free exact self-join.  The array's name is a host
// stand-in in host code: kernels get the device address that
// hipGetSymbolAddress returns for each device (kept in g_noalloc), never the
// name itself.
// ---------------------------------------------------------------------------
namespace mp {

static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
struct NoAllocSlot {
    char* ptr = nullptr;
    std::mutex use;  // held by a call while it works in the array
};
std::mutex g_noalloc_init;
NoAllocSlot g_noalloc[kMaxDevices];
}  // namespace

int noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return MP_ERR_BAD_ARG;
    std::lock_guard<std::mutex> g(g_noalloc_init);
    if (g_noalloc[device].ptr) return MP_OK;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MP_ERR_DEVICE;
    g_noalloc[device].ptr = static_cast<char*>(p);
    return MP_OK;
}

int noalloc_scratch(int device, char** ptr, std::mutex** lock) {
    if (device < 0 || device >= kMaxDevices) return MP_ERR_BAD_ARG;
    std::lock_guard<std::mutex> g(g_noalloc_init);
    if (!g_noalloc[device].ptr) return MP_ERR_DEVICE;
    *ptr = g_noalloc[device].ptr;
    *lock = &g_noalloc[device].use;
    return MP_OK;
}

}  // namespace mp

// ---------------------------------------------------------------------------
// Host-side launchers (used by api.cpp).  All take device pointers and a stream.
// ---------------------------------------------------------------------------
namespace mp {

int launch_ref_stats(const double* A, int n, int m, double* cs, double* sq, double* sInv,
                     double* mus, hipStream_t st) {
    const int P = n - m + 1;
    hipLaunchKernelGGL(k_ref_cumsum, dim3(2), dim3(64), 0, st, A, n, cs, sq);  // one wave per chain
    hipLaunchKernelGGL(k_ref_stats, dim3((P + 255) / 256), dim3(256), 0, st, cs, sq, P, m, sInv,
                       mus);
    return hipGetLastError() == hipSuccess ? MP_OK : MP_ERR_DEVICE;
}

int launch_ref_seeds(const double* A, int base, int count, int m, double* qt, hipStream_t st) {
    if (count <= 0) return MP_OK;
    hipLaunchKernelGGL(k_ref_seeds, dim3((count + 255) / 256), dim3(256), 0, st, A, base, count,
                       m, qt);
    return hipGetLastError() == hipSuccess ? MP_OK : MP_ERR_DEVICE;
}

int launch_fill_u64(unsigned long long* p, int64_t count, unsigned long long v, hipStream_t st) {
    if (count <= 0) return MP_OK;
    hipLaunchKernelGGL(k_fill_u64, dim3(static_cast<unsigned>((count + 255) / 256)), dim3(256), 0,
                       st, p, count, v);
    return hipGetLastError() == hipSuccess ? MP_OK : MP_ERR_DEVICE;
}

int launch_fill_u32(unsigned int* p, int count, unsigned int v, hipStream_t st) {
    if (count <= 0) return MP_OK;
    hipLaunchKernelGGL(k_fill_u32, dim3((count + 255) / 256), dim3(256), 0, st, p, count, v);
    return hipGetLastError() == hipSuccess ? MP_OK : MP_ERR_DEVICE;
}

int launch_unpack_f64(const unsigned long long* val, const unsigned long long* key, int P, int m,
                      int ib, int64_t index_init, int band_k, const double* A, const double* sInv,
                      const double* mus, int base, double* profile, int64_t* index, hipStream_t st) {
    const dim3 grid((P + 255) / 256), block(256);
    switch (band_k) {
        case 1:
            hipLaunchKernelGGL(k_unpack_f64<1>, grid, block, 0, st, val, key, P, m, ib, index_init, A, sInv, mus,
                               base, profile, index);
            break;
        case 2:
            hipLaunchKernelGGL(k_unpack_f64<2>, grid, block, 0, st, val, key, P, m, ib, index_init, A, sInv, mus,
                               base, profile, index);
            break;
        case 4:
            hipLaunchKernelGGL(k_unpack_f64<4>, grid, block, 0, st, val, key, P, m, ib, index_init, A, sInv, mus,
                               base, profile, index);
            break;
        case 8:
            hipLaunchKernelGGL(k_unpack_f64<8>, grid, block, 0, st, val, key, P, m, ib, index_init, A, sInv, mus,
                               base, profile, index);
            break;
        default:
            return MP_ERR_BAD_ARG;
    }
    return hipGetLastError() == hipSuccess ? MP_OK : MP_ERR_DEVICE;
}

int launch_pack_block(const double* prof, int count, int ib, unsigned long long* val,
                      unsigned long long* key, hipStream_t st) {
    if (count <= 0) return MP_OK;
    hipLaunchKernelGGL(k_pack_block, dim3((count + 255) / 256), dim3(256), 0, st, prof, count,
                       ib, val, key);
    return hipGetLastError() == hipSuccess ? MP_OK : MP_ERR_DEVICE;
}

int launch_unpack_block(const unsigned long long* val, const unsigned long long* key, int count,
                        int ib, int64_t pmin, const int64_t* idx_in, double* prof_out,
                        int64_t* idx_out, hipStream_t st) {
    if (count <= 0) return MP_OK;
    hipLaunchKernelGGL(k_unpack_block, dim3((count + 255) / 256), dim3(256), 0, st, val, key,
                       count, ib, pmin, idx_in, prof_out, idx_out);
    return hipGetLastError() == hipSuccess ? MP_OK : MP_ERR_DEVICE;
}


int launch_pack_rec(const double* A, const double* s, const double* mu, int m, int valid,
                    int total, WalkRec* out, hipStream_t st) {
    if (total <= 0) return MP_OK;
    hipLaunchKernelGGL(k_pack_rec, dim3((total + 255) / 256), dim3(256), 0, st, A, s, mu, m,
                       valid, total, out);
    return hipGetLastError() == hipSuccess ? MP_OK : MP_ERR_DEVICE;
}

int walk_f64_nbands(int blocklen, int band_k) {
    const int bw = kWave * band_k;
    return (blocklen + bw - 1) / bw;
}

int64_t walk_f64_seg_items(int blocklen, int band_k) {
    return seg_items(walk_f64_nbands(blocklen, band_k), kSegCols / (kWave * band_k));
}

int launch_walk_f64(const WalkF64Launch& a, hipStream_t st) {
    const int nbands = walk_f64_nbands(a.blocklen, a.band_k);
    if (nbands <= 0) return MP_OK;
    if (a.form == 4) {
        // the self-join only: one block of every diagonal, no write-back
        if (a.writeback || a.blocklen != a.tl || !a.ckpt) return MP_ERR_BAD_ARG;
        const int bw = kWave * a.band_k, seg_r = kSegCols / bw;
        const int64_t threads = static_cast<int64_t>(nbands) * bw;
        hipLaunchKernelGGL(k_seg_seeds, dim3(static_cast<unsigned>((threads + 255) / 256)), dim3(256), 0, st,
                           a.colrec, a.rowrec, a.qt, a.ckpt, a.tl, a.blocklen, bw, nbands, seg_r);
        if (hipGetLastError() != hipSuccess) return MP_ERR_DEVICE;
#define MP_SEG_CASE(KK)                                                                                   \
    case KK:                                                                                                \
        hipLaunchKernelGGL((k_walk_f64_seg<KK>), dim3(a.groups), dim3(kGroup), 0, st, a.colrec, a.rowrec, \
                           a.qt, a.rowval, a.rowkey, a.colval, a.colkey, a.counter, a.row_off, a.col_off,   \
                           a.blocklen, a.tl, a.m, a.ib, nbands, seg_r, static_cast<const double*>(a.ckpt)); \
        break;
        switch (a.band_k) {
            MP_SEG_CASE(1)
            MP_SEG_CASE(2)
            MP_SEG_CASE(4)
            default:
                return MP_ERR_BAD_ARG;
        }
#undef MP_SEG_CASE
        return hipGetLastError() == hipSuccess ? MP_OK : MP_ERR_DEVICE;
    }
    const dim3 grid(a.groups), block(kGroup);
#define MP_WALK_CASE(KK, KERNEL)                                                               \
    case KK:                                                                                     \
        if (a.writeback)                                                                         \
            hipLaunchKernelGGL((KERNEL<KK, true>), grid, block, 0, st, a.colrec, a.rowrec, a.qt, \
                               a.rowval, a.rowkey, a.colval, a.colkey, a.counter, a.row_off,     \
                               a.col_off, a.blocklen, a.tl, a.m, a.ib, nbands);                  \
        else                                                                                     \
            hipLaunchKernelGGL((KERNEL<KK, false>), grid, block, 0, st, a.colrec, a.rowrec,      \
                               a.qt, a.rowval, a.rowkey, a.colval, a.colkey, a.counter,          \
                               a.row_off, a.col_off, a.blocklen, a.tl, a.m, a.ib, nbands);       \
        break;
    if (a.band_k == 8 && a.form == 3) {
        // the self-join only: one block of every diagonal, and column results
        // of P > tl entries (the walker reads the one after its last column)
        if (a.writeback || a.blocklen != a.tl) return MP_ERR_BAD_ARG;
        hipLaunchKernelGGL((k_walk_f64_rot<8>), grid, block, 0, st, a.colrec, a.rowrec, a.qt, a.rowval, a.rowkey,
                           a.colval, a.colkey, a.counter, a.row_off, a.col_off, a.blocklen, a.tl, a.m, a.ib,
                           nbands, static_cast<const unsigned long long*>(a.colval));
    } else if (a.band_k == 8 && a.form == 2) {
        switch (a.band_k) { MP_WALK_CASE(8, k_walk_f64_pair) }
    } else {
        switch (a.band_k) {
            MP_WALK_CASE(1, k_walk_f64)
            MP_WALK_CASE(2, k_walk_f64)
            MP_WALK_CASE(4, k_walk_f64)
            MP_WALK_CASE(8, k_walk_f64_wide)
            default:
                return MP_ERR_BAD_ARG;
        }
    }
#undef MP_WALK_CASE
    return hipGetLastError() == hipSuccess ? MP_OK : MP_ERR_DEVICE;
}

}  // namespace mp
