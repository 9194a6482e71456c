// PreSCRIMP on gfx950: the anytime phase of scrimppp_orig
// (ScrimpppOrig.cpp:130-283).  Every sampled row i = 0, step, 2*step, ...
// (step = floor(stride*m)) gets its full distance profile, which updates the
// row's own entry with its minimum and every column j with d(i, j); then the
// diagonal through (i, best partner) is walked step-1 cells forward and back.
// Every update is a minimum, so the reference's shuffled sequential order only
// decides ties: here all rows run at once and merge through the exact path's
// 64-bit (score, key) atomics, and the walks run after all rows.
//
//   k_pre_rows  sliding dot products QT(i, j) of 16 sampled rows against 64
//               columns at a time on the FP64 matrix cores, with the score,
//               the row/column maxima and the merge fused into the epilogue
//   k_pre_walk  one lane per sampled row: the +-(step-1) diagonal walks
//
// Scores are score = QT*sInv_i*sInv_j - mus_i*mus_j (d2 = 2m - 2*score, the
// exact path's statistics); the reference computes the same d2 from window
// means and sigmas, and its QT by FFT.  A pair is excluded when
// i - m/4 < j < i + m/4 (ScrimpppOrig.cpp:191-200).
#include "common.hpp"
#include "prescrimp.hpp"

namespace mp {

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

// (score, partner): the larger score wins, the smaller partner on exact ties
// (the reference's strict < over ascending j); NaN (an excluded or invalid
// pair) ranks below every number.
struct SP {
    double s;
    int p;
};
__device__ __forceinline__ SP pick(SP a, SP b) {
    const bool take = b.s > a.s || (a.s != a.s) || (b.s == a.s && b.p < a.p);
    return take ? b : a;
}

// pick() over lanes l, l^16, l^32, l^48 (gfx950 row swaps; pick is symmetric
// up to NaN pairs, which are discarded)
__device__ __forceinline__ SP pick_quarters(SP x) {
    auto pair = [](SP v, auto swap) __attribute__((always_inline)) {
        const auto lo = swap(static_cast<unsigned>(__double2loint(v.s)));
        const auto hi = swap(static_cast<unsigned>(__double2hiint(v.s)));
        const auto id = swap(static_cast<unsigned>(v.p));
        return pick(SP{__hiloint2double(static_cast<int>(hi[0]), static_cast<int>(lo[0])), static_cast<int>(id[0])},
                    SP{__hiloint2double(static_cast<int>(hi[1]), static_cast<int>(lo[1])), static_cast<int>(id[1])});
    };
    x = pair(x, [](unsigned v) { return __builtin_amdgcn_permlane32_swap(v, v, false, false); });
    return pair(x, [](unsigned v) { return __builtin_amdgcn_permlane16_swap(v, v, false, false); });
}

// pick() over the 16 lanes of a DPP row (row_ror by 8, 4, 2, 1): every lane of
// the row ends with the row's result
template <int ROT>
__device__ __forceinline__ SP ror_pick(SP x) {
    constexpr int ctl = 0x120 + ROT;  // row_ror:ROT
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x.s), ctl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x.s), ctl, 0xF, 0xF, false);
    const int id = __builtin_amdgcn_update_dpp(0, x.p, ctl, 0xF, 0xF, false);
    return pick(x, SP{__hiloint2double(hi, lo), id});
}

__device__ __forceinline__ void merge(unsigned long long* val, unsigned long long* key, int64_t e,
                                      double score, int64_t partner, int ib) {
    const uint64_t o = ord_f64(score);
    atomicMax(&val[e], static_cast<unsigned long long>(o));
    atomicMax(&key[e], static_cast<unsigned long long>(make_key(o, ib, partner)));
}

constexpr int kColsPerChunk = 64;  // 4 MFMA column blocks of 16
constexpr int kU = 4;              // steps per group (their loads are issued together)
static_assert(kU == 4, "the column window rolls by one 16-column sub-block (4 steps) per group");

// One wave: sampled rows 16*rb .. 16*rb+15 against columns [c0, c0 + ncols).
// MFMA v_mfma_f64_16x16x4f64: lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16] and holds D[4r + l / 16][l % 16] in its r-th result; here
// A[row][k] = A[i_row + 4s + k], B[k][col] = A[j_col + 4s + k], D = QT.
__global__ __launch_bounds__(256) void k_pre_rows(const double* __restrict__ A,
                                                  const double* __restrict__ sInv,
                                                  const double* __restrict__ mus, int P, int m,
                                                  int step, int nrows, int excl, int ib,
                                                  int cols_per_wave, int ncolwaves,
                                                  unsigned long long* val, unsigned long long* key) {
    const int lane = threadIdx.x % kWave;
    const int64_t wave = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
    const int rb = static_cast<int>(wave / ncolwaves);
    const int cw = static_cast<int>(wave % ncolwaves);
    if (rb * 16 >= nrows) return;
    const int c0 = cw * cols_per_wave;
    const int c1 = min(P, c0 + cols_per_wave);
    const int l16 = lane % 16, t0 = lane / 16;

    // this lane's A-operand row, and the rows 4r + t0 its results belong to
    const int arow = rb * 16 + l16;
    const int ai = (arow < nrows) ? arow * step : 0;
    int ri[4];
    bool rv[4];
    double rs[4], rm[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int row = rb * 16 + 4 * r + t0;
        rv[r] = row < nrows;
        ri[r] = rv[r] ? row * step : 0;
        rs[r] = sInv[ri[r]];
        rm[r] = mus[ri[r]];
    }
    SP rbest[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) rbest[r] = SP{__builtin_nan(""), 0};

    const int nsteps = (m + 3) / 4;
    const int nsteps_u = (nsteps + kU - 1) / kU * kU;  // steps past m read 0 (A-operand)
    for (int cb = c0; cb < c1; cb += kColsPerChunk) {
        int cj[4];
        double cs[4], cm[4];
        unsigned long long cur[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            cj[c] = min(cb + 16 * c + l16, P - 1);
            cs[c] = sInv[cj[c]];
            cm[c] = mus[cj[c]];
            cur[c] = val[cj[c]] >> ib;  // the column filter, read before the products
        }
        v4d acc[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[c] = v4d{0.0, 0.0, 0.0, 0.0};
        // Column operand of sub-block c at step s is X(s + 4c), X(u) =
        // A[cb + l16 + 4u + t0]: a window of 16 X values rolls by 4 per group
        // of 4 steps, so a step needs one new column load instead of four.
        const int xb = cb + l16 + t0;
        const int nlast = P + m - 2;  // last element a valid pair reads
        double xw[16];
#pragma unroll
        for (int k = 0; k < 12; ++k) xw[k] = A[min(xb + 4 * k, nlast)];
        for (int s0 = 0; s0 < nsteps_u; s0 += kU) {
            double a[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                xw[12 + e] = A[min(xb + 4 * (s0 + 12 + e), nlast)];
                const int t = 4 * (s0 + e) + t0;
                const double x = A[ai + min(t, m - 1)];
                a[e] = (t < m) ? x : 0.0;
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[e], xw[e + 4 * c], acc[c], 0, 0, 0);
            }
#pragma unroll
            for (int k = 0; k < 12; ++k) xw[k] = xw[k + 4];
        }
        // epilogue: result (r, c) is QT(row 4r + t0, column cb + 16c + l16)
        SP cbest[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) cbest[c] = SP{__builtin_nan(""), 0};
#pragma unroll
        for (int r = 0; r < 4; ++r) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int j = cb + 16 * c + l16;
                const int i = ri[r];
                const bool ok = rv[r] && j < P && !(j > i - excl && j < i + excl);
                const double sc = fma(acc[c][r], rs[r] * cs[c], -(rm[r] * cm[c]));
                const double s = ok ? sc : __builtin_nan("");
                rbest[r] = pick(rbest[r], SP{s, j});
                cbest[c] = pick(cbest[c], SP{s, i});
            }
        }
        // columns: the four lane quarters hold the other rows of the block.
        // A column result is merged only if its truncated score reaches the
        // truncated current value (values only grow; comparing the truncated
        // scores the keys hold keeps the competing keys independent of timing)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const SP b = pick_quarters(cbest[c]);
            const int j = cb + 16 * c + l16;
            if (t0 == 0 && j < P && b.s == b.s && (ord_f64(b.s) >> ib) >= cur[c]) merge(val, key, j, b.s, b.p, ib);
        }
    }
    // rows: the 16 lanes of a DPP row hold the other columns
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        SP b = rbest[r];
        b = ror_pick<8>(b);
        b = ror_pick<4>(b);
        b = ror_pick<2>(b);
        b = ror_pick<1>(b);
        if (l16 == r && rv[r] && b.s == b.s) merge(val, key, ri[r], b.s, b.p, ib);
    }
}

// One lane per sampled row: walk the diagonal through (i, j = the row's best
// partner) step-1 cells forward and back (ScrimpppOrig.cpp:218-281).  QT of
// the start cell is recomputed in FP64 (the reference recovers it from the
// row's distance); the walk then updates it like the reference:
// QT += A[i+k+m-1]*A[j+k+m-1] - A[i+k-1]*A[j+k-1].
__global__ void k_pre_walk(const double* __restrict__ A, const double* __restrict__ sInv,
                           const double* __restrict__ mus, int P, int m, int step, int nrows,
                           int ib, unsigned long long* val, unsigned long long* key) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const int i = r * step;
    const int64_t pj = key_partner(key[i], ib);
    if (pj < 0) return;
    const int j = static_cast<int>(pj);
    double z0 = 0.0;
    for (int t = 0; t < m; ++t) z0 = fma(A[j + t], A[i + t], z0);
    double z = z0;
    for (int k = 1; k < step && i + k < P && j + k < P; ++k) {
        z = fma(A[i + k + m - 1], A[j + k + m - 1], fma(-A[i + k - 1], A[j + k - 1], z));
        const double s = fma(z, sInv[i + k] * sInv[j + k], -(mus[i + k] * mus[j + k]));
        if (s == s) {
            merge(val, key, i + k, s, j + k, ib);
            merge(val, key, j + k, s, i + k, ib);
        }
    }
    z = z0;
    for (int k = 1; k < step && i - k >= 0 && j - k >= 0; ++k) {
        z = fma(A[i - k], A[j - k], fma(-A[i - k + m], A[j - k + m], z));
        const double s = fma(z, sInv[i - k] * sInv[j - k], -(mus[i - k] * mus[j - k]));
        if (s == s) {
            merge(val, key, i - k, s, j - k, ib);
            merge(val, key, j - k, s, i - k, ib);
        }
    }
}

}  // namespace

int launch_prescrimp(const PreScrimpLaunch& a, hipStream_t st) {
    if (a.nrows <= 0) return MP_OK;
    const int rblocks = (a.nrows + 15) / 16;
    const int ncolwaves = (a.P + a.cols_per_wave - 1) / a.cols_per_wave;
    const int64_t waves = static_cast<int64_t>(rblocks) * ncolwaves;
    const int64_t groups = (waves + 3) / 4;
    if (groups > 0x7FFFFFFF) return MP_ERR_TOO_LONG;
    hipLaunchKernelGGL(k_pre_rows, dim3(static_cast<unsigned>(groups)), dim3(256), 0, st, a.A, a.sInv,
                       a.mus, a.P, a.m, a.step, a.nrows, a.excl, a.ib, a.cols_per_wave, ncolwaves,
                       a.val, a.key);
    if (hipGetLastError() != hipSuccess) return MP_ERR_DEVICE;
    hipLaunchKernelGGL(k_pre_walk, dim3(static_cast<unsigned>((a.nrows + 255) / 256)), dim3(256), 0, st,
                       a.A, a.sInv, a.mus, a.P, a.m, a.step, a.nrows, a.ib, a.val, a.key);
    return hipGetLastError() == hipSuccess ? MP_OK : MP_ERR_DEVICE;
}

}  // namespace mp
