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
#include <type_traits>
#include <utility>

#include "common.hpp"

namespace mp {

// One lane scans the whole series in the reference's sequential order (the
// dependent FP64 add chain is the critical path; the loads are uniform and go
// through the scalar cache, the stores are single-lane vector stores).
__global__ __launch_bounds__(64) void k_ref_cumsum(const double* __restrict__ A, int n,
                                                  double* __restrict__ cs,
                                                  double* __restrict__ sq) {
    if (threadIdx.x != 0 || n <= 0) return;
    double c = A[0];
    double s = A[0] * A[0];
    cs[0] = c;
    sq[0] = s;
#pragma unroll 16
    for (int i = 1; i < n; ++i) {
        const double a = A[i];
        c = a + c;
        s = fma(a, a, s);
        cs[i] = c;
        sq[i] = s;
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

__global__ void k_unpack_f64(const unsigned long long* __restrict__ val,
                             const unsigned long long* __restrict__ key, int P, int m, int ib,
                             int64_t index_init, double* __restrict__ profile,
                             int64_t* __restrict__ index) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const double score = unord_f64(val[i]);
    profile[i] = 2.0 * static_cast<double>(m) - 2.0 * score;
    const int64_t p = key_partner(key[i], ib);
    index[i] = (p < 0) ? index_init : p;
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

// (value, local diagonal id) merge: the larger score wins; exact ties keep the
// larger id so that merging is symmetric and deterministic.  NaN never wins.
struct ScoreId {
    double v;
    int id;
};
__device__ __forceinline__ ScoreId better(ScoreId a, ScoreId b) {
    const bool take = b.v > a.v || (b.v == a.v && b.id > a.id);
    return take ? b : a;
}
// strict version for the per-step tree (fixed operand order keeps it deterministic)
__device__ __forceinline__ ScoreId better_strict(ScoreId a, ScoreId b) { return (b.v > a.v) ? b : a; }

template <int K, int N>
struct ColTree {  // maximum of N consecutive cells starting at cell K0 (pairwise tree)
    template <int K0>
    __device__ __forceinline__ static ScoreId reduce(const double* sc, int lid0) {
        if constexpr (N == 1) {
            return ScoreId{sc[K0], lid0 + K0};
        } else {
            const ScoreId a = ColTree<K, N / 2>::template reduce<K0>(sc, lid0);
            const ScoreId b = ColTree<K, N - N / 2>::template reduce<K0 + N / 2>(sc, lid0);
            return better_strict(a, b);
        }
    }
};

// WRITEBACK: store each diagonal's final QT (the block drop-in's tmpQ); the
// self-join does not need it and skips the per-cell end-of-diagonal test.
template <int K, bool WRITEBACK>
__global__ __launch_bounds__(kGroup) void k_walk_f64(
    const double* __restrict__ Ah, const double* __restrict__ sh, const double* __restrict__ mh,
    const double* __restrict__ Av, const double* __restrict__ sv, const double* __restrict__ mv,
    double* qt, unsigned long long* rowval, unsigned long long* rowkey,
    unsigned long long* colval, unsigned long long* colkey, unsigned int* counter,
    int64_t row_off, int64_t col_off, int blocklen, int tl, int m, int ib, int nbands) {
    constexpr int F = kFlush;
    static_assert(F % K == 0, "flush period must be a multiple of K");
    __shared__ double s_cv[kWavesPerGroup][F][kWave + 1];
    __shared__ int s_ci[kWavesPerGroup][F][kWave + 1];
    __shared__ double s_rv[kWavesPerGroup][F + kWave];  // [F, F+64): per-lane dummy slots
    __shared__ int s_ri[kWavesPerGroup][F + kWave];

    const int wv = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const double NEG = -__builtin_inf();
    const double QNAN = __builtin_nan("");
    const int nitems = (nbands + 1) / 2;

    // Flush the results of columns [jstart, jstart+cnt) and of the rows that
    // left lane 0 at those columns.
    auto flush = [&](int jstart, int cnt, int dbase) {
        __builtin_amdgcn_wave_barrier();
        const int cs = lane & (F - 1), qq = lane / F;
        ScoreId b{NEG, -1};
        if (cs < cnt) {
#pragma unroll 2
            for (int u = 0; u < kWave / 4; ++u) {
                const ScoreId c{s_cv[wv][cs][qq * (kWave / 4) + u], s_ci[wv][cs][qq * (kWave / 4) + u]};
                b = better(b, c);
            }
        }
#pragma unroll
        for (int off = F; off < kWave; off <<= 1) {
            const ScoreId o{__shfl_xor(b.v, off), __shfl_xor(b.id, off)};
            b = better(b, o);
        }
        if (lane < cnt && b.id >= 0) {
            const int jj = jstart + lane;
            const uint64_t o = ord_f64(b.v);
            atomicMax(&colval[jj], static_cast<unsigned long long>(o));
            atomicMax(&colkey[jj], static_cast<unsigned long long>(
                                       make_key(o, ib, static_cast<int64_t>(jj) + dbase + b.id + col_off)));
        }
        if (lane < cnt) {
            const double v = s_rv[wv][lane];
            const int id = s_ri[wv][lane];
            if (id >= 0) {
                const int jj = jstart + lane;
                const int i = jj + dbase;
                const uint64_t o = ord_f64(v);
                atomicMax(&rowval[i], static_cast<unsigned long long>(o));
                atomicMax(&rowkey[i], static_cast<unsigned long long>(
                                          make_key(o, ib, static_cast<int64_t>(jj - id) + row_off)));
            }
        }
        __builtin_amdgcn_wave_barrier();
    };

    for (;;) {
        const int item = next_item(counter, lane);
        if (item >= nitems) break;
        // long/short pairing of bands (the reference pairs partitions r and 2p-1-r,
        // ScrimpTrivialPar.cpp:216-224) keeps the work per item even.
        for (int h = 0; h < 2; ++h) {
            const int band = (h == 0) ? item : nbands - 1 - item;
            if (h == 1 && band == item) break;
            const int dbase = band * kWave * K;
            const int L = tl - dbase;
            const int lid0 = lane * K;
            // Cells outside the triangle are masked by NaN: a row past the last
            // one gets sInv = NaN, a diagonal past the block gets QT = NaN, and a
            // NaN score never wins a comparison (nor does it in the reference).
            double q[K], wa[K], wam[K], ws[K], wm[K], rv[K];
            int ri[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int d = dbase + lid0 + k;
                q[k] = (d < blocklen) ? qt[d] : QNAN;
                const int r = min(d, tl - 1);
                wa[k] = Av[r];
                wam[k] = Av[r + m];
                ws[k] = (d < tl) ? sv[r] : QNAN;
                wm[k] = mv[r];
                rv[k] = NEG;
                ri[k] = -1;
            }

            // One step: column j for all K diagonals of every lane.  TAIL steps may
            // run past the band's last column and stop there.
            auto step = [&](int j, auto s_tag, auto tail_tag) -> bool {
                constexpr int S = decltype(s_tag)::value;
                constexpr bool TAIL = decltype(tail_tag)::value;
                if (TAIL && j >= L) return false;
                const double ahj = Ah[j], ahjm = Ah[j + m], shj = sh[j], mhj = mh[j];
                const int inext = j + dbase + kWave * K;  // row entering lane 63
                const int ifeed = min(inext, tl - 1);
                const double fa = Av[ifeed], fam = Av[ifeed + m], fsl = sv[ifeed], fm = mv[ifeed];
                // ifeed is clamped, so the load is always in bounds; select afterwards
                const double fs = (inext < tl) ? fsl : QNAN;
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
                    if (sc[k] > rv[sl]) {  // NaN never wins
                        rv[sl] = sc[k];
                        ri[sl] = lid0 + k;
                    }
                }
                const ScoreId col = ColTree<K, K>::template reduce<0>(sc, lid0);
                const int fslot = j & (F - 1);
                s_cv[wv][fslot][lane] = col.v;
                s_ci[wv][fslot][lane] = (col.v > NEG) ? col.id : -1;
                // lane 0's slot S holds the row leaving the band; the other lanes
                // write their copy to a private dummy slot (no branch)
                const int rsel = (lane == 0) ? fslot : F + lane;
                s_rv[wv][rsel] = rv[S];
                s_ri[wv][rsel] = ri[S];
                rv[S] = shift_down_f64(rv[S], NEG);
                ri[S] = shift_down_i32(ri[S], -1);
                wa[S] = shift_down_f64(wa[S], fa);
                wam[S] = shift_down_f64(wam[S], fam);
                ws[S] = shift_down_f64(ws[S], fs);
                wm[S] = shift_down_f64(wm[S], fm);
                return true;
            };
            auto macro = [&](int c0, auto tail_tag) -> bool {
                bool go = true;
                [&]<int... S>(std::integer_sequence<int, S...>) {
                    ((go = go && step(c0 + S, std::integral_constant<int, S>{}, tail_tag)), ...);
                }(std::make_integer_sequence<int, K>{});
                return go;
            };
            int c0 = 0;
            // whole flush windows: no control flow inside a window
            for (; c0 + F <= L; c0 += F) {
#pragma unroll 1
                for (int w = 0; w < F; w += K) macro(c0 + w, std::false_type{});
                flush(c0, F, dbase);
            }
            // the last partial window
            if (c0 < L) {
                const int start = c0;
                for (; c0 < L; c0 += K) macro(c0, std::true_type{});
                flush(start, L - start, dbase);
            }
        }
    }
}

}  // namespace mp

// ---------------------------------------------------------------------------
// Host-side launchers (used by api.cpp).  All take device pointers and a stream.
// ---------------------------------------------------------------------------
namespace mp {

int launch_ref_stats(const double* A, int n, int m, double* cs, double* sq, double* sInv,
                     double* mus, hipStream_t st) {
    const int P = n - m + 1;
    hipLaunchKernelGGL(k_ref_cumsum, dim3(1), dim3(64), 0, st, A, n, cs, sq);
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
                      int ib, int64_t index_init, double* profile, int64_t* index,
                      hipStream_t st) {
    hipLaunchKernelGGL(k_unpack_f64, dim3((P + 255) / 256), dim3(256), 0, st, val, key, P, m, ib,
                       index_init, profile, index);
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

struct WalkF64Launch {
    const double *Ah, *sh, *mh, *Av, *sv, *mv;
    double* qt;
    unsigned long long *rowval, *rowkey, *colval, *colkey;
    unsigned int* counter;
    int64_t row_off, col_off;
    int blocklen, tl, m, ib, writeback, band_k, groups;
};

int walk_f64_nbands(int blocklen, int band_k) {
    const int bw = kWave * band_k;
    return (blocklen + bw - 1) / bw;
}

int launch_walk_f64(const WalkF64Launch& a, hipStream_t st) {
    const int nbands = walk_f64_nbands(a.blocklen, a.band_k);
    if (nbands <= 0) return MP_OK;
    const dim3 grid(a.groups), block(kGroup);
#define MP_WALK_CASE(KK)                                                                         \
    case KK:                                                                                     \
        if (a.writeback)                                                                         \
            hipLaunchKernelGGL((k_walk_f64<KK, true>), grid, block, 0, st, a.Ah, a.sh, a.mh,     \
                               a.Av, a.sv, a.mv, a.qt, a.rowval, a.rowkey, a.colval, a.colkey,   \
                               a.counter, a.row_off, a.col_off, a.blocklen, a.tl, a.m, a.ib,     \
                               nbands);                                                          \
        else                                                                                     \
            hipLaunchKernelGGL((k_walk_f64<KK, false>), grid, block, 0, st, a.Ah, a.sh, a.mh,    \
                               a.Av, a.sv, a.mv, a.qt, a.rowval, a.rowkey, a.colval, a.colkey,   \
                               a.counter, a.row_off, a.col_off, a.blocklen, a.tl, a.m, a.ib,     \
                               nbands);                                                          \
        break;
    switch (a.band_k) {
        MP_WALK_CASE(1)
        MP_WALK_CASE(2)
        MP_WALK_CASE(4)
        MP_WALK_CASE(8)
        default:
            return MP_ERR_BAD_ARG;
    }
#undef MP_WALK_CASE
    return hipGetLastError() == hipSuccess ? MP_OK : MP_ERR_DEVICE;
}

}  // namespace mp
