// FP32-reseeded SCRIMP path (one series or a batch of series in one launch).
//
//   k_comp_stats   mean/sigma of every window from compensated (Neumaier,
//                  exact-product) sliding sums in FP64, restarted every 256
//                  windows from a direct sum, on data shifted by the chunk's
//                  first value so that mean^2 >> sigma^2 does not cancel.
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
// Register trick: a cell's result is the double whose high word is the FP32
// bits of corr+1 (>= 0) and whose low word is the cell's local diagonal id
// K*lane+k.  Doubles with a non-negative high word order like that word and
// then the low word, so a single v_max_f64 keeps the row (or column) maximum
// together with the diagonal that produced it.  The partner follows from the
// diagonal: column j = i - d for a row, row i = j + d for a column.
#include <type_traits>

#include "common.hpp"
#include "reseeded_f32.hpp"

namespace mp {

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

constexpr int kStatChunk = 256;

__global__ void k_comp_stats(const SeriesDesc* __restrict__ descs,
                             const int64_t* __restrict__ entry_prefix, int nseries,
                             int64_t nentries, const double* __restrict__ A,
                             double* __restrict__ mu, double* __restrict__ sigma) {
    const int64_t e0 = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) * kStatChunk;
    if (e0 >= nentries) return;
    const int64_t e1 = min(e0 + kStatChunk, nentries);
    int64_t e = e0;
    while (e < e1) {
        const int s = find_series(entry_prefix, nseries, e);
        const SeriesDesc d = descs[s];
        const int64_t send = entry_prefix[s + 1];
        const int i0 = static_cast<int>(e - entry_prefix[s]);
        const int i1 = static_cast<int>(min(e1, send) - entry_prefix[s]);
        const double* a = A + d.off;
        const int m = d.m;
        const double dm = static_cast<double>(m);
        const double shift = a[i0];
        double S = 0.0, Sc = 0.0, Q = 0.0, Qc = 0.0;
        for (int t = 0; t < m; ++t) {
            const double x = a[i0 + t] - shift;
            neu_add(S, Sc, x);
            neu_add_sq(Q, Qc, x);
        }
        for (int i = i0; i < i1; ++i) {
            if (i > i0) {
                const double xo = a[i - 1] - shift, xn = a[i - 1 + m] - shift;
                neu_add(S, Sc, xn);
                neu_add(S, Sc, -xo);
                neu_add_sq(Q, Qc, xn);
                neu_sub_sq(Q, Qc, xo);
            }
            const double sm = (S + Sc) / dm;
            double var = fma(-sm, sm, (Q + Qc) / dm);
            if (!(var > 0.0)) var = 0.0;
            mu[d.off + i] = shift + sm;
            sigma[d.off + i] = sqrt(var);
        }
        e = min(e1, send);
    }
}

__global__ void k_prep_terms(const SeriesDesc* __restrict__ descs,
                             const int64_t* __restrict__ entry_prefix, int nseries,
                             int64_t nentries, const double* __restrict__ A,
                             const double* __restrict__ mu, const double* __restrict__ sigma,
                             float* __restrict__ df, float* __restrict__ dg,
                             float* __restrict__ inv) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= nentries) return;
    const int s = find_series(entry_prefix, nseries, e);
    const SeriesDesc d = descs[s];
    const int i = static_cast<int>(e - entry_prefix[s]);
    const double* a = A + d.off;
    const int64_t o = d.off + i;
    const double sg = sigma[o];
    inv[o] = (sg > 0.0) ? static_cast<float>(1.0 / (sg * sqrt(static_cast<double>(d.m)))) : 0.0f;
    if (i + 1 < d.P) {
        df[o] = static_cast<float>(0.5 * (a[i + d.m] - a[i]));
        dg[o] = static_cast<float>((a[i + d.m] - mu[o + 1]) + (a[i] - mu[o]));
    } else {
        df[o] = 0.0f;
        dg[o] = 0.0f;
    }
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

__device__ __forceinline__ unsigned long long global_key(double k, int64_t partner) {
    float c1 = __int_as_float(__double2hiint(k));
    const uint32_t bits = (c1 > 0.0f) ? __float_as_uint(c1) : 0u;
    return (static_cast<unsigned long long>(bits) << 32) |
           static_cast<unsigned long long>(0xFFFFFFFEu - static_cast<uint32_t>(partner));
}

template <int K>
__global__ __launch_bounds__(kGroup) void k_tiles_f32(
    const SeriesDesc* __restrict__ descs, const int64_t* __restrict__ tile_prefix,
    const int64_t* __restrict__ band_prefix, int nseries, const double* __restrict__ A,
    const double* __restrict__ mu, const float* __restrict__ df, const float* __restrict__ dg,
    const float* __restrict__ inv, unsigned long long* keys, unsigned int* counter,
    int64_t ntiles, int T) {
    constexpr int F = kFlush;
    constexpr int BW = kWave * K;
    static_assert(F % K == 0, "flush period must be a multiple of K");
    static_assert(BW * sizeof(double) <= F * (kWave + 1) * sizeof(double), "seed staging fits");
    __shared__ double s_ck[kWavesPerGroup][F][kWave + 1];  // column partials; seed staging
    __shared__ double s_rk[kWavesPerGroup][F];

    const int wv = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const int lid0 = lane * K;
    double* seeds = &s_ck[wv][0][0];

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
        const int cb = q * T;
        const int ce = min(cb + T, L);
        const int64_t roff = d.off + d.base;  // row (vertical) views
        const int64_t coff = d.off;           // column (horizontal) views
        // all cells of lane l, diagonal k at column j are valid iff j + k <= lim0
        const int lim0 = tl - 1 - dbase - lid0;
        const int dtop = dbase + BW - 1;
        const int lim_top = (dtop < tl) ? (tl - 1 - dtop) : -1;

        // ---- FP64 seeds, strided over lanes for coalesced loads ----
        {
            double acc[K];
#pragma unroll
            for (int k = 0; k < K; ++k) acc[k] = 0.0;
            const double muc = mu[coff + cb];
            const double* av = A + roff;
            const double* ah = A + coff + cb;
            int rows[K];
#pragma unroll
            for (int k = 0; k < K; ++k) rows[k] = min(cb + dbase + lane + kWave * k, tl - 1);
            for (int t = 0; t < m; ++t) {
                const double c = ah[t] - muc;
#pragma unroll
                for (int k = 0; k < K; ++k) acc[k] = fma(av[rows[k] + t], c, acc[k]);
            }
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int k = 0; k < K; ++k) seeds[lane + kWave * k] = acc[k];
            __builtin_amdgcn_wave_barrier();
        }
        float cov[K], wdf[K], wdg[K], wiv[K];
        double rk[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            cov[k] = static_cast<float>(seeds[lid0 + k]);
            const int r = min(cb + dbase + lid0 + k, tl - 1);
            wdf[k] = df[roff + r];
            wdg[k] = dg[roff + r];
            wiv[k] = inv[roff + r];
            rk[k] = 0.0;
        }
        __builtin_amdgcn_wave_barrier();

        auto flush = [&](int jstart, int cnt) {
            __builtin_amdgcn_wave_barrier();
            const int cs = lane & (F - 1), qq = lane / F;
            double bk = 0.0;
            if (cs < cnt) {
#pragma unroll 4
                for (int v = 0; v < kWave / 4; ++v) bk = fmax(bk, s_ck[wv][cs][qq * (kWave / 4) + v]);
            }
#pragma unroll
            for (int off = F; off < kWave; off <<= 1) bk = fmax(bk, __shfl_xor(bk, off));
            if (lane < cnt && __double2hiint(bk) != 0) {
                const int jj = jstart + lane;
                const int lid = __double2loint(bk);
                const int64_t partner = static_cast<int64_t>(d.base) + jj + dbase + lid;
                atomicMax(&keys[coff + jj], global_key(bk, partner));
            }
            if (lane < cnt) {
                const double rkv = s_rk[wv][lane];
                if (__double2hiint(rkv) != 0) {
                    const int jj = jstart + lane;
                    const int lid = __double2loint(rkv);
                    atomicMax(&keys[roff + jj + dbase], global_key(rkv, jj - lid));
                }
            }
            __builtin_amdgcn_wave_barrier();
        };

        auto macro = [&](int c0, auto full_tag) {
            constexpr bool FULL = decltype(full_tag)::value;
#pragma unroll
            for (int st = 0; st < K; ++st) {
                const int j = c0 + st;
                if (!FULL && j >= ce) break;
                const float dfh = df[coff + j], dgh = dg[coff + j], ivh = inv[coff + j];
                const int ifeed = min(j + dbase + BW, tl - 1);
                const float fdf = df[roff + ifeed], fdg = dg[roff + ifeed], fiv = inv[roff + ifeed];
                double ck = 0.0;
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const int sl = (st + k) % K;
                    const float c1 = fmaf(cov[k] * wiv[sl], ivh, 1.0f);
                    double key = __hiloint2double(__float_as_int(c1), lid0 + k);
                    cov[k] = fmaf(wdf[sl], dgh, cov[k]);
                    cov[k] = fmaf(dfh, wdg[sl], cov[k]);
                    if constexpr (!FULL) key = (j + k <= lim0) ? key : 0.0;
                    rk[sl] = fmax(rk[sl], key);
                    ck = fmax(ck, key);
                }
                const int fslot = (j - cb) & (F - 1);
                s_ck[wv][fslot][lane] = ck;
                if (lane == 0) s_rk[wv][fslot] = rk[st];
                {
                    const uint64_t rb = static_cast<uint64_t>(__double_as_longlong(rk[st]));
                    const int lo = shift_down_i32_zero(static_cast<int>(rb));
                    const int hi = shift_down_i32_zero(static_cast<int>(rb >> 32));
                    rk[st] = __hiloint2double(hi, lo);
                }
                wdf[st] = shift_down_f32(wdf[st], fdf);
                wdg[st] = shift_down_f32(wdg[st], fdg);
                wiv[st] = shift_down_f32(wiv[st], fiv);
                if (fslot == F - 1 || (!FULL && j == ce - 1)) flush(j - fslot, fslot + 1);
            }
        };

        for (int c0 = cb; c0 < ce; c0 += K) {
            if (c0 + K - 1 < lim_top && c0 + K < ce)
                macro(c0, std::integral_constant<bool, true>{});
            else
                macro(c0, std::integral_constant<bool, false>{});
        }
        // rows still inside the band at the tile's end: slot sl holds row
        // ce + dbase + lid0 + ((sl - ce) mod K)
#pragma unroll
        for (int sl = 0; sl < K; ++sl) {
            const int k = ((sl - ce) % K + K) % K;
            const int i = ce + dbase + lid0 + k;
            if (i < tl && __double2hiint(rk[sl]) != 0) {
                const int lid = __double2loint(rk[sl]);
                atomicMax(&keys[roff + i], global_key(rk[sl], i - dbase - lid));
            }
        }
    }
}

// One wave per profile entry.  d^2 = 2m(1 - rho): rho from the FP32 key, or
// recomputed in FP64 from the centred dot product of the winning pair.
__global__ __launch_bounds__(256) void k_unpack_f32r(
    const SeriesDesc* __restrict__ descs, const int64_t* __restrict__ entry_prefix, int nseries,
    int64_t nentries, const double* __restrict__ A, const double* __restrict__ mu,
    const double* __restrict__ sigma, const unsigned long long* __restrict__ keys, int refine,
    int64_t index_init, double* __restrict__ profile, int64_t* __restrict__ index) {
    const int lane = threadIdx.x % kWave;
    const int64_t wave = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
    const int64_t nwaves = static_cast<int64_t>(gridDim.x) * blockDim.x / kWave;
    for (int64_t e = wave; e < nentries; e += nwaves) {
        const int s = find_series(entry_prefix, nseries, e);
        const SeriesDesc d = descs[s];
        const int i = static_cast<int>(e - entry_prefix[s]);
        const unsigned long long key = keys[d.off + i];
        const uint32_t low = static_cast<uint32_t>(key);
        const double dm = static_cast<double>(d.m);
        double d2;
        int64_t idx;
        if (low == 0xFFFFFFFFu) {
            d2 = 2.0 * dm + 2.0;  // nothing beat the initial score -1.0
            idx = index_init;
        } else if (0xFFFFFFFEu - low >= static_cast<uint32_t>(d.P)) {
            d2 = __builtin_nan("");  // a partner outside the profile: reported, never dereferenced
            idx = -2;
        } else {
            const int p = static_cast<int>(0xFFFFFFFEu - low);
            idx = p;
            const float c1 = __uint_as_float(static_cast<uint32_t>(key >> 32));
            d2 = 2.0 * dm * (2.0 - static_cast<double>(c1));
            if (refine) {
                const double* a = A + d.off;
                const double mi = mu[d.off + i], mp = mu[d.off + p];
                double acc = 0.0;
                for (int t = lane; t < d.m; t += kWave) acc = fma(a[i + t] - mi, a[p + t] - mp, acc);
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
                const double den = dm * sigma[d.off + i] * sigma[d.off + p];
                if (den > 0.0) d2 = 2.0 * dm * (1.0 - acc / den);
            }
        }
        if (lane == 0) {
            profile[d.out_off + i] = d2;
            index[d.out_off + i] = idx;
        }
    }
}

// ---------------------------------------------------------------------------
int launch_f32r_prep(const F32rBatch& b, hipStream_t st) {
    const int64_t nchunks = (b.nentries + kStatChunk - 1) / kStatChunk;
    hipLaunchKernelGGL(k_comp_stats, dim3(static_cast<unsigned>((nchunks + 127) / 128)), dim3(128),
                       0, st, b.descs, b.entry_prefix, b.nseries, b.nentries, b.A, b.mu, b.sigma);
    const unsigned g = static_cast<unsigned>((b.nentries + 255) / 256);
    hipLaunchKernelGGL(k_prep_terms, dim3(g), dim3(256), 0, st, b.descs, b.entry_prefix,
                       b.nseries, b.nentries, b.A, b.mu, b.sigma, b.df, b.dg, b.inv);
    hipLaunchKernelGGL(k_init_keys, dim3(g), dim3(256), 0, st, b.descs, b.entry_prefix,
                       b.nseries, b.nentries, b.keys);
    return hipGetLastError() == hipSuccess ? MP_OK : MP_ERR_DEVICE;
}

int launch_f32r_tiles(const F32rBatch& b, int groups, hipStream_t st) {
    if (b.ntiles <= 0) return MP_OK;
#define MP_TILE_CASE(KK)                                                                        \
    case KK:                                                                                    \
        hipLaunchKernelGGL(k_tiles_f32<KK>, dim3(groups), dim3(kGroup), 0, st, b.descs,         \
                           b.tile_prefix, b.band_prefix, b.nseries, b.A, b.mu, b.df, b.dg,      \
                           b.inv, b.keys, b.counter, b.ntiles, b.tile_len);                     \
        break;
    switch (b.band_k) {
        MP_TILE_CASE(4)
        MP_TILE_CASE(8)
        MP_TILE_CASE(16)
        default:
            return MP_ERR_BAD_ARG;
    }
#undef MP_TILE_CASE
    return hipGetLastError() == hipSuccess ? MP_OK : MP_ERR_DEVICE;
}

int launch_f32r_unpack(const F32rBatch& b, int refine, int64_t index_init, double* profile,
                       int64_t* index, hipStream_t st) {
    const int64_t waves = b.nentries;
    const int64_t gw = (waves + 3) / 4;
    const unsigned g = static_cast<unsigned>(gw < 65536 ? gw : 65536);
    hipLaunchKernelGGL(k_unpack_f32r, dim3(g), dim3(256), 0, st, b.descs, b.entry_prefix,
                       b.nseries, b.nentries, b.A, b.mu, b.sigma, b.keys, refine, index_init,
                       profile, index);
    return hipGetLastError() == hipSuccess ? MP_OK : MP_ERR_DEVICE;
}

}  // namespace mp
