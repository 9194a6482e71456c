// FP32-reseeded SCRIMP path (one series or a batch of series in one launch).
//
//   k_comp_stats   mean/sigma of every window from compensated (Neumaier,
//                  exact-product) sliding sums in FP64, restarted every 64
//                  windows of a series from a direct sum, on data shifted by
//                  the chunk's first value so that mean^2 >> sigma^2 does not
//                  cancel.
//   k_prep_terms   FP32 recurrence terms df, dg and 1/(sigma*sqrt(m)).
//   k_tiles_f32    persistent tile kernel: a tile is a band of 64*K diagonals
//                  times T columns.  Its start is seeded with an exact FP64
//                  centred dot product, the covariance then advances in FP32,
//                  cov(i+1,j+1) = cov(i,j) + df[i]*dg[j] + df[j]*dg[i].
//   k_unpack_f32r  key -> index; d^2 from the FP32 correlation or, with
//                  refine, recomputed in FP64 for the winning pair.
//
// The recurrence terms follow the mean-centred formulation of SCAMP
// (Zimmerman et al., SoCC 2019): df[i] = (A[i+m]-A[i])/2,
// dg[i] = (A[i+m]-mu[i+1]) + (A[i]-mu[i]).  Results use the reference's
// d^2 = 2m - 2*score with score = m*rho (ScrimpTrivialPar.cpp:336-340).
//
// Lazy argmax.  A cell's result is c1 = corr + 1 (>= 0) in FP32, and a step
// keeps 32-bit maxima only: the column maximum of a lane's K cells and, per
// row, the maximum over the lane's cells of that row.  The lane that holds an
// argmax is tracked exactly (a row carries the lane that last raised it; a
// column's flush reduction keeps the lane of its maximum); which of the
// lane's K diagonals it was is left open.  A result's key holds the lane's
// first candidate partner and the direction of the others (p0 + k for a
// column result, rows j + d; p0 - k for a row result, columns i - d), and the
// unpack recomputes the K candidates' correlations in FP64 and keeps the best.
#include <type_traits>
#include <utility>

#include "common.hpp"
#include "reseeded_f32.hpp"

namespace mp {

// Two adjacent doubles read by one 16-byte load (8-byte alignment suffices on
// gfx950, whose global loads need not be naturally aligned).
struct alignas(8) DoublePair {
    double a, b;
};

__device__ __forceinline__ void neu_add(double& s, double& c, double x) {
    const double t = s + x;
    if (fabs(s) >= fabs(x))
        c += (s - t) + x;
    else
        c += (x - t) + s;
    s = t;
}

// exact square: p + e == x*x
__device__ __forceinline__ void neu_add_sq(double& s, double& c, double x) {
    const double p = x * x;
    const double e = fma(x, x, -p);
    neu_add(s, c, p);
    c += e;
}
__device__ __forceinline__ void neu_sub_sq(double& s, double& c, double x) {
    const double p = x * x;
    const double e = fma(x, x, -p);
    neu_add(s, c, -p);
    c -= e;
}

__device__ __forceinline__ int find_series(const int64_t* __restrict__ prefix, int nseries,
                                           int64_t v) {
    int lo = 0, hi = nseries;  // prefix[lo] <= v < prefix[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= v)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// One thread per chunk of kStatChunk windows of one series: a direct
// compensated sum at the chunk's first window, then sliding updates.  Chunks
// start at series-local multiples of kStatChunk, so a series gets the same
// statistics alone or in any batch.  A lane's values arrive in blocks of 8
// (four 16-byte loads), the next block's loads issued before the current
// block's sums, so the dependent sums do not wait on each load in turn.  A
// block can reach past the series (at most 14 elements; their values are not
// used): every series is followed by the next one or by the 64 elements of
// padding its buffer has, and reads are clamped to that padding.
__global__ void k_comp_stats(const SeriesDesc* __restrict__ descs,
                             const int64_t* __restrict__ stat_prefix, int nseries, int64_t nchunks,
                             const double* __restrict__ A, double* __restrict__ mu,
                             double* __restrict__ sigma) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nchunks) return;
    const int s = find_series(stat_prefix, nseries, t);
    const SeriesDesc d = descs[s];
    const int i0 = static_cast<int>(t - stat_prefix[s]) * kStatChunk;
    const int i1 = min(i0 + kStatChunk, d.P);
    const double* a = A + d.off;
    const int m = d.m;
    const double dm = static_cast<double>(m);
    const int lim = d.n + 56;  // a pair read from here stays inside the series' 64-element padding
    constexpr int B = 8;
    auto load = [&](int at, double (&v)[B]) __attribute__((always_inline)) {
#pragma unroll
        for (int u = 0; u < B; u += 2) {
            const DoublePair q = *reinterpret_cast<const DoublePair*>(a + min(at + u, lim));
            v[u] = q.a;
            v[u + 1] = q.b;
        }
    };
    // data shifted by the chunk's first value: sums of small numbers
    const double shift = a[i0];
    double S = 0.0, Sc = 0.0, Q = 0.0, Qc = 0.0;
    {
        double v[B], w[B];
        load(i0, v);
        for (int k0 = 0; k0 < m; k0 += B) {
            load(i0 + k0 + B, w);
#pragma unroll
            for (int u = 0; u < B; ++u) {
                if (k0 + u < m) {
                    const double x = v[u] - shift;
                    neu_add(S, Sc, x);
                    neu_add_sq(Q, Qc, x);
                }
            }
#pragma unroll
            for (int u = 0; u < B; ++u) v[u] = w[u];
        }
    }
    auto emit = [&](int i) __attribute__((always_inline)) {
        const double sm = (S + Sc) / dm;
        const double qm = (Q + Qc) / dm;
        double var = fma(-sm, sm, qm);
        // below the rounding noise of the sums the window is flat (sigma = 0):
        // a constant window reached by sliding need not give exactly 0
        if (!(var > 64.0 * __DBL_EPSILON__ * qm)) var = 0.0;
        mu[d.off + i] = shift + sm;
        sigma[d.off + i] = sqrt(var);
    };
    emit(i0);
    // window i0 + 1 + q drops a[i0 + q] and takes a[i0 + m + q]
    const int nq = i1 - i0 - 1;
    double vo[B], vn[B], wo[B], wn[B];
    load(i0, vo);
    load(i0 + m, vn);
    for (int q0 = 0; q0 < nq; q0 += B) {
        load(i0 + q0 + B, wo);
        load(i0 + m + q0 + B, wn);
#pragma unroll
        for (int u = 0; u < B; ++u) {
            if (q0 + u < nq) {
                const double xo = vo[u] - shift, xn = vn[u] - shift;
                neu_add(S, Sc, xn);
                neu_add(S, Sc, -xo);
                neu_add_sq(Q, Qc, xn);
                neu_sub_sq(Q, Qc, xo);
                emit(i0 + 1 + q0 + u);
            }
        }
#pragma unroll
        for (int u = 0; u < B; ++u) {
            vo[u] = wo[u];
            vn[u] = wn[u];
        }
    }
}

__global__ void k_prep_terms(const SeriesDesc* __restrict__ descs,
                             const int64_t* __restrict__ entry_prefix, int nseries,
                             int64_t nentries, const double* __restrict__ A,
                             const double* __restrict__ mu, const double* __restrict__ sigma,
                             TermRec* __restrict__ terms) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= nentries) return;
    const int s = find_series(entry_prefix, nseries, e);
    const SeriesDesc d = descs[s];
    const int i = static_cast<int>(e - entry_prefix[s]);
    const double* a = A + d.off;
    const int64_t o = d.off + i;
    const double sg = sigma[o];
    const float iv = (sg > 0.0) ? static_cast<float>(1.0 / (sg * sqrt(static_cast<double>(d.m)))) : 0.0f;
    float f = 0.0f, g = 0.0f;
    if (i + 1 < d.P) {
        f = static_cast<float>(0.5 * (a[i + d.m] - a[i]));
        g = static_cast<float>((a[i + d.m] - mu[o + 1]) + (a[i] - mu[o]));
    }
    terms[o] = TermRec{f, g, iv, 0.0f};
}

__global__ void k_init_keys(const SeriesDesc* __restrict__ descs,
                            const int64_t* __restrict__ entry_prefix, int nseries,
                            int64_t nentries, unsigned long long* __restrict__ keys) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= nentries) return;
    const int s = find_series(entry_prefix, nseries, e);
    const SeriesDesc d = descs[s];
    const int i = static_cast<int>(e - entry_prefix[s]);
    keys[d.off + i] = (static_cast<unsigned long long>(__float_as_uint(d.init_c1)) << 32) |
                      0xFFFFFFFFull;
}

// Global key: FP32 bits of c1 above, 0xFFFFFFFE - code below (a smaller code
// wins an exact tie; 0xFFFFFFFF is the initial "no partner" word).  code =
// p0 << 1 | dir (dir = 1: the candidates run p0, p0 - 1, ...).
__device__ __forceinline__ unsigned long long global_key(float c1, int64_t code) {
    const uint32_t bits = (c1 > 0.0f) ? __float_as_uint(c1) : 0u;
    return (static_cast<unsigned long long>(bits) << 32) |
           static_cast<unsigned long long>(0xFFFFFFFEu - static_cast<uint32_t>(code));
}

// Maximum of two c1 values.  c1 is a finite FP32 number (never NaN), so the
// plain instruction needs none of the quieting a generic fmaxf inserts.
__device__ __forceinline__ float c1_max(float a, float b) {
    float r;
    asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// A column result: its maximum c1 and the smallest lane that holds it.
struct C1Lane {
    float v;
    int lane;
};
// c1_max() and min() over lanes l, l^16, l^32 and l^48, in every lane.  Each
// gfx950 row swap leaves a lane with its own dword and the dword of lane l^32
// (l^16) in two registers, in some order; both operations are symmetric.
__device__ __forceinline__ float c1_max_quarters(float x) {
    auto step = [](float v, auto swap) __attribute__((always_inline)) {
        const auto a = swap(__float_as_uint(v));
        return c1_max(__uint_as_float(a[0]), __uint_as_float(a[1]));
    };
    x = step(x, [](unsigned v) { return __builtin_amdgcn_permlane32_swap(v, v, false, false); });
    return step(x, [](unsigned v) { return __builtin_amdgcn_permlane16_swap(v, v, false, false); });
}
__device__ __forceinline__ int c1_min_quarters(int x) {
    auto step = [](int v, auto swap) __attribute__((always_inline)) {
        const auto a = swap(static_cast<unsigned>(v));
        return min(static_cast<int>(a[0]), static_cast<int>(a[1]));
    };
    x = step(x, [](unsigned v) { return __builtin_amdgcn_permlane32_swap(v, v, false, false); });
    return step(x, [](unsigned v) { return __builtin_amdgcn_permlane16_swap(v, v, false, false); });
}
// Maximum of three c1 values in one instruction.
__device__ __forceinline__ float c1_max3(float a, float b, float c) {
    float r;
    asm("v_max3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}
// Maximum of N consecutive c1 values starting at K0: three-input maxima where
// they fit (8 values: 3 v_max3 + 1 v_max, where pairwise maxima take 7).
template <int N, int K0 = 0>
__device__ __forceinline__ float c1_max_tree(const float* c) {
    if constexpr (N == 1) {
        return c[K0];
    } else if constexpr (N == 2) {
        return c1_max(c[K0], c[K0 + 1]);
    } else if constexpr (N == 3) {
        return c1_max3(c[K0], c[K0 + 1], c[K0 + 2]);
    } else {
        constexpr int A = (N + 2) / 3, B = (N - A + 1) / 2;
        return c1_max3(c1_max_tree<A, K0>(c), c1_max_tree<B, K0 + A>(c), c1_max_tree<N - A - B, K0 + A + B>(c));
    }
}

// ACC64 = false: FP32 recurrence with FP32 accumulation.  ACC64 = true: each
// diagonal's covariance is an FP64 sum covd plus an FP32 partial cov of the
// current flush window's increments, folded into covd when the next window
// starts; a cell's correlation is taken from fl32(covd) and the partial.  The
// FP32 partial spans 16 steps, so its rounding stays about m/16 times below
// that of a covariance accumulated in FP32, while a step costs FP32
// instructions only (the fold adds three FP64-class ones per 16 cells).
//
// Per step the wave needs two uniform records: the column's terms, a scalar
// load, and the terms of the row entering lane 63, a broadcast LDS read from a
// staging array that a direct-to-LDS load fills one flush window ahead (two
// buffers, alternating by window).  The LDS read lands in the registers that
// the DPP shift of the rows then writes, so lane 63's fill costs no vector
// instruction.  Both are issued one step ahead; scalar loads may return out of
// order, so the wait for them is a wait for every scalar and LDS operation in
// flight, and the step takes it first, before it issues anything.  The step's
// column partial and the row leaving lane 0 are written to LDS by the next
// step, after that wait, so the wait never covers LDS writes issued just
// before it.  fterms is terms again: the staging loads read through it, so that
// they are known not to write what the scalar loads read.
template <int K, bool ACC64>
__device__ __forceinline__ void tiles_body(
    const SeriesDesc* __restrict__ descs, const int64_t* __restrict__ tile_prefix,
    const int64_t* __restrict__ band_prefix, int nseries, const double* __restrict__ A,
    const double* __restrict__ mu, const TermRec* __restrict__ terms, const TermRec* __restrict__ fterms,
    unsigned long long* keys, unsigned int* counter, int64_t ntiles, int T) {
    using Rec = TermRec;
    constexpr int F = kFlush;
    constexpr int BW = kWave * K;
    static_assert(F % K == 0, "flush period must be a multiple of K");
    // column partials (F rows of 64, padded to 65) and, before a tile's steps,
    // the staging of its FP64 seeds
    constexpr int kBufDoubles = (BW > F * (kWave + 1) / 2) ? BW : F * (kWave + 1) / 2;
    __shared__ double s_buf[kWavesPerGroup][kBufDoubles];
    __shared__ float s_rk[kWavesPerGroup][2 * F + kWave];  // [0, F) row exits; the rest scratch
    __shared__ int s_rl[kWavesPerGroup][2 * F + kWave];    // their lanes + 1 (0: none)
    __shared__ Rec s_feed[kWavesPerGroup][2 * F];           // rows entering lane 63

    const int wv = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const int lid0 = lane * K;
    double* seeds = &s_buf[wv][0];
    Rec* const feed = &s_feed[wv][0];
    float* const ck_base = reinterpret_cast<float*>(&s_buf[wv][0]);  // [F][kWave + 1]

    for (;;) {
        unsigned int tv = 0;
        if (lane == 0) tv = atomicAdd(counter, 1u);
        const int64_t tile = static_cast<int64_t>(__builtin_amdgcn_readfirstlane(tv));
        if (tile >= ntiles) break;
        const int s = find_series(tile_prefix, nseries, tile);
        const SeriesDesc d = descs[s];
        const int64_t u = tile - tile_prefix[s] + d.tile_base;
        const int64_t* bp = band_prefix + d.band_off;
        const int b = find_series(bp, d.nbands, u);
        const int q = static_cast<int>(u - bp[b]);
        const int dbase = b * BW;
        const int tl = d.tl;
        const int m = d.m;
        const int L = tl - dbase;
        const int lm = band_main_cols(L, K);
        const int nmain = (lm + T - 1) / T;
        const bool edge = q >= nmain;  // the band's last tile: columns [lm, L)
        const int cb = edge ? lm : q * T;
        const int ce = edge ? L : min(cb + T, lm);
        const int64_t roff = d.off + d.base;  // row (vertical) views
        const int64_t coff = d.off;           // column (horizontal) views
        // The whole tile is specialised on EDGE (main tiles: whole windows,
        // no cell tests), so that each specialisation holds its own registers.
        auto run_tile = [&](auto edge_tag) __attribute__((always_inline)) {
        constexpr bool EDGE = decltype(edge_tag)::value;

        // ---- FP64 seeds on the matrix cores ----
        // seed[d] = sum_t A[r0 + d + t] * c[t] with c[t] = A[cb + t] - mu[cb]
        // (centred column window), for the band's diagonals d = 256h + 16 d1 +
        // d0.  With p = 4s + t0 and c = 0 outside [0, m):
        //   D_h[d0][d1] = sum_s sum_t0 A[r0 + 256h + d0 + 4s + t0] * c[4s + t0 - 16 d1],
        // a sum over s of 16x4 by 4x16 FP64 products whose first operand does
        // not depend on d1 and second not on d0 (v_mfma_f64_16x16x4f64; lane l
        // supplies A[l % 16][l / 16] and B[l / 16][l % 16], and holds
        // D[4r + l / 16][l % 16] in its r-th result).  Element indices past the
        // last one a valid diagonal reads are clamped: they only reach the
        // sums of diagonals that lie outside the triangle.
        {
            typedef double v4d __attribute__((ext_vector_type(4)));
            constexpr int H = K / 4;  // blocks of 256 diagonals
            static_assert(K % 4 == 0, "bands are whole 256-diagonal blocks");
            v4d acc[H];
#pragma unroll
            for (int h = 0; h < H; ++h) acc[h] = v4d{0.0, 0.0, 0.0, 0.0};
            const double muc = mu[coff + cb];
            const double* av = A + roff;
            const double* ah = A + coff + cb;
            const int r0 = cb + dbase;
            const int lim = tl + m - 2;  // last element of the last valid row's window
            const int l16 = lane % 16, t0 = lane / 16;
            // Steps in blocks of U, all loads of a block issued before its
            // products.  Steps past (m + 240) / 4 read columns past m, whose
            // operand is 0, so the step count is rounded up to whole blocks.
            constexpr int U = 8;
            const int nsteps = ((m + 15 * 16 + 3) / 4 + U - 1) / U * U;
            for (int st0 = 0; st0 < nsteps; st0 += U) {
                double bv[U], a[U][H];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int st = st0 + u;
                    const int ib = 4 * st + t0 - 16 * l16;
                    const double raw = ah[min(max(ib, 0), m - 1)];
                    bv[u] = (ib >= 0 && ib < m) ? raw - muc : 0.0;
#pragma unroll
                    for (int h = 0; h < H; ++h) a[u][h] = av[min(r0 + 256 * h + l16 + 4 * st + t0, lim)];
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
#pragma unroll
                    for (int h = 0; h < H; ++h)
                        acc[h] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u][h], bv[u], acc[h], 0, 0, 0);
                }
            }
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int h = 0; h < H; ++h) {
#pragma unroll
                for (int r = 0; r < 4; ++r) seeds[256 * h + 16 * l16 + 4 * r + t0] = acc[h][r];
            }
            __builtin_amdgcn_wave_barrier();
        }
        // Row records: recurrence terms and 1/norm.  Rows past the triangle
        // (row >= tl) are clamped to its last row: every cell of such a row lies
        // outside the triangle, and the windows that reach them mask those cells.
        float cov[K], wdf[K], wdg[K];
        double covd[ACC64 ? K : 1];
        float chi[ACC64 ? K : 1];  // fl32(covd), refreshed at each fold
        float wiv[K];
        // per row slot: rin = the row's maximum over the lanes above this one,
        // tag = the lane that holds it + 1 (0: none), lm = the maximum over
        // this lane's cells of the row so far.  rin = 0 means none: a value
        // c1 <= 0 is never merged, so it needs no lane.
        float rin[K], lm[K];
        int tag[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if constexpr (ACC64) {
                covd[k] = seeds[lid0 + k];
                chi[k] = static_cast<float>(covd[k]);
                cov[k] = 0.0f;
            } else {
                cov[k] = static_cast<float>(seeds[lid0 + k]);
            }
            const Rec t = terms[roff + min(cb + dbase + lid0 + k, tl - 1)];
            wdf[k] = t.df;
            wdg[k] = t.dg;
            wiv[k] = t.inv;
            rin[k] = 0.0f;
            lm[k] = -1.0f;
            tag[k] = 0;
        }
        __builtin_amdgcn_wave_barrier();
        // lane 0 stores the row leaving the band at s_rk[wv][fslot]; lane l > 0
        // stores into s_rk[wv][F + l + fslot], a scratch slot nobody reads
        float* const rk_out = &s_rk[wv][(lane == 0) ? 0 : F + lane];
        int* const rl_out = &s_rl[wv][(lane == 0) ? 0 : F + lane];
        float* const ck_out = ck_base + lane;

        // The filter: a partial is merged only if it beats the key read when
        // the window started (lane x < F holds those of column c0 + x and of its
        // row).  Keys only grow, so a stale read only lets an unneeded atomic
        // through.
        unsigned long long pc = 0, pr = 0;
        // (a main tile's windows are whole: no clamp, and the window's base is
        // uniform, so the loads take a scalar base and a fixed lane offset)
        const int lane_f = lane & (F - 1);
        auto load_filter = [&](int c0, auto edge_tag) __attribute__((always_inline)) {
            if constexpr (decltype(edge_tag)::value) {
                const int jj = min(c0 + lane_f, ce - 1);
                pc = keys[coff + jj];
                pr = keys[roff + jj + dbase];
            } else {
                pc = (keys + (coff + c0))[lane_f];
                pr = (keys + (roff + c0 + dbase))[lane_f];
            }
        };
        // Column results hold p0 = the partner of the lane's first diagonal
        // (candidates p0 + k); row results p0 = the column of that diagonal
        // (candidates p0 - k).
        auto flush = [&](int jstart, int cnt) __attribute__((always_inline)) {
            __builtin_amdgcn_wave_barrier();
            // lane (cs, qq) reduces a quarter of column cs's 64 partials, those
            // of lanes 16 qq .. 16 qq + 15 (for cs >= cnt a stale quarter, whose
            // result is not used)
            const int cs = lane & (F - 1), qq = lane / F;
            const float* part = ck_base + cs * (kWave + 1) + qq * (kWave / 4);
            float pv[kWave / 4];
#pragma unroll
            for (int v = 0; v < kWave / 4; ++v) pv[v] = part[v];
            // A column can pass the filter only if one of its partials reaches
            // the c1 of its current key (the key's high word; c1 >= 0 there, so
            // the order of the words is the order of the values).  The
            // reduction runs only when some partial of the window can pass.
            const float thr = __uint_as_float(static_cast<uint32_t>(pc >> 32));
            const float qmx = c1_max_tree<kWave / 4>(pv);  // this quarter's maximum
            const bool cand = (qmx >= thr) && (cs < cnt);
            C1Lane b{-1.0f, 0};
            if (__builtin_amdgcn_ballot_w64(cand) != 0) {
                // the column's maximum, then the smallest lane holding it: this
                // quarter's first, then the smallest over the quarters
                const float mx = c1_max_quarters(qmx);
                int id = kWave;
#pragma unroll
                for (int v = kWave / 4 - 1; v >= 0; --v) id = (pv[v] == mx) ? qq * (kWave / 4) + v : id;
                b = C1Lane{mx, c1_min_quarters(id)};
            }
            if (lane < cnt && b.v > 0.0f && b.v >= thr) {
                const int jj = jstart + lane;
                const int64_t p0 = static_cast<int64_t>(d.base) + jj + dbase + b.lane * K;
                const unsigned long long gk = global_key(b.v, p0 << 1);
                if (gk > pc) atomicMax(&keys[coff + jj], gk);
            }
            if (lane < cnt) {
                const float rkv = s_rk[wv][lane];
                const int rl = s_rl[wv][lane] - 1;
                // (the c1 of the row's current key first: most rows cannot
                // pass, and then no lane builds a key)
                if (rl >= 0 && rkv > 0.0f && rkv >= __uint_as_float(static_cast<uint32_t>(pr >> 32))) {
                    const int jj = jstart + lane;
                    const int64_t p0 = static_cast<int64_t>(jj) - rl * K;
                    const unsigned long long gk = global_key(rkv, (p0 << 1) | 1);
                    if (gk > pr) atomicMax(&keys[roff + jj + dbase], gk);
                }
            }
            __builtin_amdgcn_wave_barrier();
        };
        // The rows entering lane 63 at columns c0 .. c0 + F - 1, clamped like
        // the others, to feed buffer buf: 16 bytes per lane.
        auto stage_feed = [&](int c0, int buf) __attribute__((always_inline)) {
            if (lane < F) {
                const Rec* src = fterms + roff + min(c0 + lane + dbase + BW, tl - 1);
                __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(src),
                                                 (__attribute__((address_space(3))) void*)(feed + buf * F), 16, 0,
                                                 0);
            }
        };
        Rec cur_c, cur_f;  // the next step's column and feed records
        const Rec* fw = feed;  // this window's feed rows
        float pck = -1.0f, prk = -1.0f;  // the last step's column partial and row exit
        int prl = 0;
        auto put_partials = [&](int w) __attribute__((always_inline)) {
            ck_out[w * (kWave + 1)] = pck;
            rk_out[w] = prk;
            rl_out[w] = prl;
        };

        // One column j = c0 + W of the band; W is the step's place in the flush
        // window, so the row slot (W mod K) and the LDS offsets are constants.
        // EDGE: the window reaches past the triangle; a cell (j, lid) lies in it
        // iff j + lid < L, and a cell outside gets the key -1 (never a winner:
        // valid keys are corr + 1 >= 0).  TAIL: the window may end before F steps.
        // Row maxima two steps at a time: an even step (the first of a pair;
        // K and F are even, so a pair never spans windows) leaves its cells of
        // rows that stay in the lane in pend[k], and the odd step folds each
        // with its own cell of the row in one v_max3 (9 maxima per pair where
        // one per cell took 14).
        float pend[K];
#pragma unroll
        for (int k = 0; k < K; ++k) pend[k] = -1.0f;
        auto step = [&](int c0, auto w_tag, auto edge_tag, auto tail_tag) __attribute__((always_inline)) -> bool {
            constexpr int W = decltype(w_tag)::value;
            constexpr int S = W % K;
            constexpr bool ODD = (W & 1) != 0;
            static_assert(K % 2 == 0 && F % 2 == 0, "steps pair up within a window");
            constexpr bool EDGE = decltype(edge_tag)::value;
            constexpr bool TAIL = decltype(tail_tag)::value;
            const int j = c0 + W;
            if (TAIL && j >= ce) return false;
            __builtin_amdgcn_sched_barrier(0);
            const Rec c = cur_c, f = cur_f;
            asm volatile("" ::"s"(c.df), "s"(c.dg), "s"(c.inv) : "memory");
            cur_c = terms[coff + j + 1];
            // the window's last step: the next window reads its first row
            if constexpr (W + 1 < F) cur_f = fw[W + 1];
            if constexpr (W > 0) put_partials(W - 1);
            __builtin_amdgcn_sched_barrier(0);
            const float dfh = c.df, dgh = c.dg;
            const float ivh = c.inv;
            float c1[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int sl = (S + k) % K;
                // the same three roundings in both forms: of the covariance to
                // FP32, of its product with 1/norm_i, and of the final FMA
                if constexpr (ACC64 && W == 0) {
                    // the window's first step: cov = 0 (just folded), and
                    // fma(chi, 1/norm_i, 0) is the product (1/norm_i is finite)
                    c1[k] = fmaf(chi[k] * wiv[sl], ivh, 1.0f);
                } else if constexpr (ACC64) {
                    c1[k] = fmaf(fmaf(chi[k], wiv[sl], cov[k] * wiv[sl]), ivh, 1.0f);
                } else {
                    c1[k] = fmaf(cov[k] * wiv[sl], ivh, 1.0f);
                }
                if constexpr (EDGE) c1[k] = (lid0 + k < L - j) ? c1[k] : -1.0f;
                cov[k] = fmaf(wdf[sl], dgh, cov[k]);
                cov[k] = fmaf(dfh, wdg[sl], cov[k]);
                // the row in slot (S + K - 1) % K entered this lane at the end of
                // the last step: its first cell here starts lm; the row in slot
                // S (k = 0) leaves now, its last cell taken below.  The others
                // wait in pend (even step) or fold the pending cell of the
                // last step, which was their cell k + 1 (odd step; k = K - 2:
                // that one was the row's first, already in lm).
                if (k == K - 1) {
                    lm[sl] = c1[k];
                } else if (k > 0) {
                    if constexpr (!ODD) {
                        pend[k] = c1[k];
                    } else if (k == K - 2) {
                        lm[sl] = c1_max(lm[sl], c1[k]);
                    } else {
                        lm[sl] = c1_max3(lm[sl], pend[k + 1], c1[k]);
                    }
                }
            }
            pck = c1_max_tree<K>(c1);
            // the row in slot S leaves this lane: its maximum so far (with its
            // cells not yet in lm: this step's c1[0] and, after an even step,
            // the last step's pend[1]), and this lane as the argmax's when its
            // cells beat the lanes above (strictly: the row met those lanes
            // first, at smaller columns)
            if constexpr (!ODD) {
                prk = c1_max3(rin[S], lm[S], c1[0]);
            } else {
                prk = c1_max3(rin[S], c1_max(lm[S], pend[1]), c1[0]);
            }
            prl = (prk > rin[S]) ? lane + 1 : tag[S];
            rin[S] = shift_down_f32_zero(prk);
            tag[S] = shift_down_i32_zero(prl);
            wdf[S] = shift_down_f32(wdf[S], f.df);
            wdg[S] = shift_down_f32(wdg[S], f.dg);
            wiv[S] = shift_down_f32(wiv[S], f.inv);
            return true;
        };
        auto window = [&](int c0, auto edge_tag, auto tail_tag) __attribute__((always_inline)) {
            if constexpr (ACC64) {  // fold the last window's partials
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    covd[k] += static_cast<double>(cov[k]);
                    cov[k] = 0.0f;
                    chi[k] = static_cast<float>(covd[k]);
                }
            }
            load_filter(c0, edge_tag);
            // the next window's feed rows to the other buffer; this window's
            // landed before the last flush
            const int buf = ((c0 - cb) / F) & 1;
            __builtin_amdgcn_wave_barrier();
            stage_feed(c0 + F, buf ^ 1);
            fw = feed + buf * F;
            cur_f = fw[0];
            __builtin_amdgcn_wave_barrier();
            bool go = true;
            [&]<int... W>(std::integer_sequence<int, W...>) __attribute__((always_inline)) {
                ((go = go && step(c0, std::integral_constant<int, W>{}, edge_tag, tail_tag)), ...);
            }(std::make_integer_sequence<int, F>{});
            // the window's last step has no successor to write its partials
            if constexpr (decltype(tail_tag)::value) {
                const int cnt = min(F, ce - c0);
                put_partials(cnt - 1);
                // a window cut after an even step: that step's pending cells
                // go into lm now (its slot is a compile-time constant per case)
                if (cnt & 1) {
                    const int s_last = (cnt - 1) % K;
                    [&]<int... I>(std::integer_sequence<int, I...>) __attribute__((always_inline)) {
                        auto fold = [&](auto s_tag) __attribute__((always_inline)) {
                            constexpr int SL = decltype(s_tag)::value;
#pragma unroll
                            for (int k = 1; k < K - 1; ++k) lm[(SL + k) % K] = c1_max(lm[(SL + k) % K], pend[k]);
                        };
                        ((s_last == 2 * I ? fold(std::integral_constant<int, 2 * I>{}) : void()), ...);
                    }(std::make_integer_sequence<int, K / 2>{});
                }
            } else {
                put_partials(F - 1);
            }
        };
        cur_c = terms[coff + cb];
        stage_feed(cb, 0);
        for (int c0 = cb; c0 < ce; c0 += F) {
            // the direct-to-LDS loads in flight (this window's feed rows)
            // must have landed: the compiler does not order LDS reads after them
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            window(c0, edge_tag, edge_tag);
            flush(c0, EDGE ? min(F, ce - c0) : F);
        }
        // rows still inside the band at the tile's end: slot sl holds row
        // ce + dbase + lid0 + ((sl - (ce - cb)) mod K); its maximum so far is
        // max(rin, lm), from this lane when lm is larger.  The slot the last
        // step refilled holds a row with no cell in this lane yet: its lm still
        // belongs to the row that left.
        const int refilled = ((ce - cb - 1) % K + K) % K;
#pragma unroll
        for (int sl = 0; sl < K; ++sl) {
            const int k = ((sl - (ce - cb)) % K + K) % K;
            const int i = ce + dbase + lid0 + k;
            const bool mine = (sl != refilled) && (lm[sl] > rin[sl]);
            const float v = mine ? lm[sl] : rin[sl];
            const int rl = mine ? lane : tag[sl] - 1;
            if (i < tl && rl >= 0 && v > 0.0f) {
                const int64_t p0 = static_cast<int64_t>(i) - dbase - rl * K;
                const unsigned long long gk = global_key(v, (p0 << 1) | 1);
                if (gk > keys[roff + i]) atomicMax(&keys[roff + i], gk);
            }
        }
        };
        if (edge)
            run_tile(std::true_type{});
        else
            run_tile(std::false_type{});
    }
}

#define MP_TILE_PARAMS                                                                          \
    const SeriesDesc *__restrict__ descs, const int64_t *__restrict__ tile_prefix,              \
        const int64_t *__restrict__ band_prefix, int nseries, const double *__restrict__ A,     \
        const double *__restrict__ mu, const TermRec *__restrict__ terms,                       \
        const TermRec *__restrict__ fterms, unsigned long long *keys, unsigned int *counter,    \
        int64_t ntiles, int T
#define MP_TILE_ARGS \
    descs, tile_prefix, band_prefix, nseries, A, mu, terms, fterms, keys, counter, ntiles, T

// K <= 8: at most 128 VGPRs, so that 4 waves share each SIMD.
template <int K, bool ACC64>
__global__ __launch_bounds__(kGroup) __attribute__((amdgpu_waves_per_eu(4))) void k_tiles_f32(
    MP_TILE_PARAMS) {
    tiles_body<K, ACC64>(MP_TILE_ARGS);
}
// K = 16 holds twice the row records; capping its registers would spill.
template <int K, bool ACC64>
__global__ __launch_bounds__(kGroup) void k_tiles_f32_wide(MP_TILE_PARAMS) {
    tiles_body<K, ACC64>(MP_TILE_ARGS);
}
#undef MP_TILE_PARAMS
#undef MP_TILE_ARGS

// The winning key of an entry names the K candidate partners of one lane's
// diagonals; their correlations are recomputed in FP64 from centred dot
// products and the best (the smaller partner on an exact tie) is the index.
// d^2 = 2m(1 - rho) with that rho (refine) or with the FP32 key's (no refine).
// One wave per chunk of 64 consecutive entries, one entry per lane: the series
// of the chunk's first entry is found once, and a lane whose entry lies in a
// later series advances from there.  The candidates' windows overlap: a lane
// walks its window x_t = A[i+t] - mu_i in blocks of K positions, holds the
// 2K-1 values y = A[lo+t+v] - mu_lo that the K candidates lo .. lo+K-1 need
// there (K of them loaded, K-1 kept from the previous block) and does K*K
// FMAs.  Loads past the last candidate's window (kcount <
// K near the end of a series) are skipped, so no read leaves the series.
constexpr int kUnpackChunk = 64;

template <int K>
__global__ __launch_bounds__(256) void k_unpack_f32r(
    const SeriesDesc* __restrict__ descs, const int64_t* __restrict__ entry_prefix, int nseries,
    int64_t nentries, const double* __restrict__ A, const double* __restrict__ mu,
    const double* __restrict__ sigma, const unsigned long long* __restrict__ keys, int refine,
    int64_t index_init, double* __restrict__ profile, int64_t* __restrict__ index) {
    const int lane = threadIdx.x % kWave;
    const int64_t wave = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
    const int64_t nwaves = static_cast<int64_t>(gridDim.x) * blockDim.x / kWave;
    for (int64_t e0 = wave * kUnpackChunk; e0 < nentries; e0 += nwaves * kUnpackChunk) {
        const int64_t e = e0 + lane;
        const bool live = e < nentries;
        int s = find_series(entry_prefix, nseries, e0);
        if (live) {
            while (s + 1 < nseries && entry_prefix[s + 1] <= e) ++s;
        }
        const SeriesDesc d = descs[s];
        const int i = live ? static_cast<int>(e - entry_prefix[s]) : 0;
        const int64_t ai = d.off + i;
        const unsigned long long key = live ? keys[ai] : 0xFFFFFFFFull;
        const uint32_t low = static_cast<uint32_t>(key);
        const double dm = static_cast<double>(d.m);
        double d2 = 0.0;
        int64_t idx = 0;
        if (!live) {
        } else if (low == 0xFFFFFFFFu || !(sigma[ai] > 0.0)) {
            // nothing beat the initial score -1.0, or a flat window (zero
            // variance), whose every score is NaN in FP64 and never wins
            d2 = 2.0 * dm + 2.0;
            idx = index_init;
        } else {
            const int64_t code = 0xFFFFFFFEu - low;
            const int64_t p0 = code >> 1;
            const int lo0 = static_cast<int>((code & 1) ? p0 - (K - 1) : p0);
            const int lo = lo0 < 0 ? 0 : lo0;  // candidates lo .. lo+kcount-1, within [lo0, lo0+K)
            const int kcount = static_cast<int>(min<int64_t>(K, static_cast<int64_t>(d.P) - lo));
            const float c1 = __uint_as_float(static_cast<uint32_t>(key >> 32));
            d2 = 2.0 * dm * (2.0 - static_cast<double>(c1));
            idx = -2;  // no candidate cell: reported, never dereferenced
            if (kcount > 0) {
                const int m = d.m;
                const double mi = mu[ai], ml = mu[d.off + lo];
                const double* __restrict__ x = A + ai;
                const double* __restrict__ y = A + d.off + lo;
                const int ylim = m + kcount - 1;  // y positions the candidates read
                double acc[K], sx = 0.0;
#pragma unroll
                for (int k = 0; k < K; ++k) acc[k] = 0.0;
                // ys slides along y: a block loads its last K values, the first
                // K-1 are the previous block's last ones
                double ys[2 * K - 1];
#pragma unroll
                for (int v = 0; v < K - 1; ++v) ys[v] = (v < ylim) ? y[v] - ml : 0.0;
                int t = 0;
                // with all K candidates every read of a block lies within the
                // series (t + 2K - 2 < m + K - 1): two values per 16-byte load
                const bool full = kcount == K;
                for (; t + K <= m; t += K) {
                    double xs[K];
#pragma unroll
                    for (int u = 0; u < K; u += 2) {
                        const DoublePair q = *reinterpret_cast<const DoublePair*>(x + t + u);
                        xs[u] = q.a - mi;
                        xs[u + 1] = q.b - mi;
                    }
#pragma unroll
                    for (int u = 0; u < K; ++u) sx += xs[u];
                    if (full) {
#pragma unroll
                        for (int v = K - 1; v < 2 * K - 1; v += 2) {
                            const DoublePair q = *reinterpret_cast<const DoublePair*>(y + t + v);
                            ys[v] = q.a - ml;
                            ys[v + 1] = q.b - ml;
                        }
                    } else {
#pragma unroll
                        for (int v = K - 1; v < 2 * K - 1; ++v) ys[v] = (t + v < ylim) ? y[t + v] - ml : 0.0;
                    }
#pragma unroll
                    for (int k = 0; k < K; ++k) {
#pragma unroll
                        for (int u = 0; u < K; ++u) acc[k] = fma(xs[u], ys[u + k], acc[k]);
                    }
#pragma unroll
                    for (int v = 0; v < K - 1; ++v) ys[v] = ys[v + K];
                }
                for (; t < m; ++t) {
                    const double xv = x[t] - mi;
                    sx += xv;
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        if (k < kcount) acc[k] = fma(xv, y[t + k] - ml, acc[k]);
                    }
                }
                double best_rho = 0.0;
                int best = -1, first = -1;  // first: a candidate cell, even of a flat window
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const int p = lo + k;
                    const int gap = p > i ? p - i : i - p;
                    if (k < kcount && p >= lo0 && gap >= d.base) {
                        if (first < 0) first = p;
                        const double sp = sigma[d.off + p];
                        // centred dot: sum x_t (A[p+t] - mu_p) = acc - (mu_p - mu_lo) * sum x_t
                        const double rho = (acc[k] - (mu[d.off + p] - ml) * sx) / (dm * sigma[ai] * sp);
                        if (sp > 0.0 && (best < 0 || rho > best_rho)) {
                            best = p;
                            best_rho = rho;
                        }
                    }
                }
                if (best >= 0) {
                    idx = best;
                    if (refine) d2 = 2.0 * dm * (1.0 - best_rho);
                } else if (first >= 0) {
                    idx = first;  // only flat partners (correlation 0 in FP32): the FP32 value stands
                }
            }
        }
        if (live) {
            profile[d.out_off + i] = d2;
            index[d.out_off + i] = idx;
        }
    }
}

__device__ __forceinline__ int find_stage(const StageDesc* d, int count, int64_t chunk) {
    int lo = 0, hi = count;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (d[mid].chunk_base <= chunk)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// pinned series (then its band prefix, as int64 bits) -> device pools
__global__ __launch_bounds__(256) void k_stage_in(const StageDesc* __restrict__ descs, int count,
                                                  const double* __restrict__ pinned,
                                                  double* __restrict__ A,
                                                  int64_t* __restrict__ bandpool) {
    const int64_t chunk = blockIdx.x;
    const StageDesc d = descs[find_stage(descs, count, chunk)];
    const int64_t e0 = (chunk - d.chunk_base) * kStageChunk;
    const int64_t total = static_cast<int64_t>(d.n) + d.nb;
    const int64_t e1 = min(e0 + kStageChunk, total);
    const int64_t* pin_bands = reinterpret_cast<const int64_t*>(pinned + d.pin_off + d.n);
    for (int64_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
        if (e < d.n)
            A[d.dev_off + e] = pinned[d.pin_off + e];
        else
            bandpool[d.band_dev_off + (e - d.n)] = pin_bands[e - d.n];
    }
}

// device profile/index -> pinned (profile, then index as int64 bits)
__global__ __launch_bounds__(256) void k_stage_out(const StageDesc* __restrict__ descs, int count,
                                                   const double* __restrict__ prof,
                                                   const int64_t* __restrict__ idx,
                                                   double* __restrict__ pinned) {
    const int64_t chunk = blockIdx.x;
    const StageDesc d = descs[find_stage(descs, count, chunk)];
    const int64_t e0 = (chunk - d.chunk_base) * kStageChunk;
    const int64_t e1 = min(e0 + kStageChunk, static_cast<int64_t>(d.n));
    int64_t* pin_idx = reinterpret_cast<int64_t*>(pinned + d.pin_off + d.n);
    for (int64_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
        pinned[d.pin_off + e] = prof[d.dev_off + e];
        pin_idx[e] = idx[d.dev_off + e];
    }
}

int launch_stage_in(const StageDesc* d, int count, int64_t nchunks, const double* pinned,
                    double* A, int64_t* bandpool, hipStream_t st) {
    if (count <= 0 || nchunks <= 0) return MP_OK;
    hipLaunchKernelGGL(k_stage_in, dim3(static_cast<unsigned>(nchunks)), dim3(256), 0, st, d, count,
                       pinned, A, bandpool);
    return hipGetLastError() == hipSuccess ? MP_OK : MP_ERR_DEVICE;
}

int launch_stage_out(const StageDesc* d, int count, int64_t nchunks, const double* prof,
                     const int64_t* idx, double* pinned, hipStream_t st) {
    if (count <= 0 || nchunks <= 0) return MP_OK;
    hipLaunchKernelGGL(k_stage_out, dim3(static_cast<unsigned>(nchunks)), dim3(256), 0, st, d,
                       count, prof, idx, pinned);
    return hipGetLastError() == hipSuccess ? MP_OK : MP_ERR_DEVICE;
}

// ---------------------------------------------------------------------------
int launch_f32r_prep(const F32rBatch& b, hipStream_t st) {
    if (b.nstat > 0)
        hipLaunchKernelGGL(k_comp_stats, dim3(static_cast<unsigned>((b.nstat + 63) / 64)), dim3(64), 0, st,
                           b.descs, b.stat_prefix, b.nseries, b.nstat, b.A, b.mu, b.sigma);
    const unsigned g = static_cast<unsigned>((b.nentries + 255) / 256);
    hipLaunchKernelGGL(k_prep_terms, dim3(g), dim3(256), 0, st, b.descs, b.entry_prefix,
                       b.nseries, b.nentries, b.A, b.mu, b.sigma, static_cast<TermRec*>(b.terms));
    hipLaunchKernelGGL(k_init_keys, dim3(g), dim3(256), 0, st, b.descs, b.entry_prefix,
                       b.nseries, b.nentries, b.keys);
    return hipGetLastError() == hipSuccess ? MP_OK : MP_ERR_DEVICE;
}

int launch_f32r_tiles(const F32rBatch& b, int groups, hipStream_t st) {
    if (b.ntiles <= 0) return MP_OK;
#define MP_TILE_CASE(KK, KERNEL)                                                                \
    case KK:                                                                                    \
        if (b.accum64)                                                                          \
            hipLaunchKernelGGL((KERNEL<KK, true>), dim3(groups), dim3(kGroup), 0, st,           \
                               b.descs, b.tile_prefix, b.band_prefix, b.nseries, b.A, b.mu,     \
                               static_cast<const TermRec*>(b.terms),                            \
                               static_cast<const TermRec*>(b.terms), b.keys, b.counter,         \
                               b.ntiles, b.tile_len);                                           \
        else                                                                                    \
            hipLaunchKernelGGL((KERNEL<KK, false>), dim3(groups), dim3(kGroup), 0, st,          \
                               b.descs, b.tile_prefix, b.band_prefix, b.nseries, b.A, b.mu,     \
                               static_cast<const TermRec*>(b.terms),                            \
                               static_cast<const TermRec*>(b.terms), b.keys, b.counter,         \
                               b.ntiles, b.tile_len);                                           \
        break;
    switch (b.band_k) {
        MP_TILE_CASE(4, k_tiles_f32)
        MP_TILE_CASE(8, k_tiles_f32)
        MP_TILE_CASE(16, k_tiles_f32_wide)
        default:
            return MP_ERR_BAD_ARG;
    }
#undef MP_TILE_CASE
    return hipGetLastError() == hipSuccess ? MP_OK : MP_ERR_DEVICE;
}

int launch_f32r_unpack(const F32rBatch& b, int refine, int64_t index_init, double* profile,
                       int64_t* index, hipStream_t st) {
    if (b.nentries <= 0) return MP_OK;
    const int64_t waves = (b.nentries + kUnpackChunk - 1) / kUnpackChunk;
    const int64_t gw = (waves + 3) / 4;
    const unsigned g = static_cast<unsigned>(gw < 65536 ? gw : 65536);
#define MP_UNPACK_CASE(KK)                                                                              \
    case KK:                                                                                            \
        hipLaunchKernelGGL(k_unpack_f32r<KK>, dim3(g), dim3(256), 0, st, b.descs, b.entry_prefix,       \
                           b.nseries, b.nentries, b.A, b.mu, b.sigma, b.keys, refine, index_init,       \
                           profile, index);                                                             \
        break;
    switch (b.band_k) {
        MP_UNPACK_CASE(4)
        MP_UNPACK_CASE(8)
        MP_UNPACK_CASE(16)
        default:
            return MP_ERR_BAD_ARG;
    }
#undef MP_UNPACK_CASE
    return hipGetLastError() == hipSuccess ? MP_OK : MP_ERR_DEVICE;
}

}  // namespace mp
