// PreSCRIMP on gfx950: the anytime phase of scrimppp_orig
// (ScrimpppOrig.cpp:130-283).  Every sampled row i = 0, step, 2*step, ...
// (step = floor(stride*m)) gets its full distance profile, which updates the
// row's own entry with its minimum and every column j with d(i, j); then the
// diagonal through (i, best partner) is walked step-1 cells forward and back.
// Every update is a minimum, so the reference's shuffled sequential order only
// decides ties: here all rows run at once and merge through the exact path's
// 64-bit (score, key) atomics, and the walks run after all rows.
//
//   k_pre_rows  sliding dot products QT(i, j) of 32 sampled rows against 64
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

#include <type_traits>

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

// pick() within a lane, whose candidates arrive in increasing partner order:
// a strict comparison keeps the first maximum, so no tie test is needed
__device__ __forceinline__ SP pick_ordered(SP a, SP b) {
    const bool take = b.s > a.s || (a.s != a.s);
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

// buffer resource word 3 for raw (untyped) buffer loads on gfx9-class chips
constexpr int kBufWord3 = 0x00020000;
constexpr int kColsPerChunk = 64;  // 4 MFMA column blocks of 16
constexpr int kU = 4;              // steps per group (their loads are issued together)
static_assert(kU == 4, "the column window rolls by one 16-column sub-block (4 steps) per group");

// One wave: RB blocks of 16 sampled rows (rows 16*RB*rb ..) against columns
// [c0, c0 + ncols); the blocks share the column operands.
// MFMA v_mfma_f64_16x16x4f64: lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16] and holds D[4r + l / 16][l % 16] in its r-th result; here
// A[row][k] = A[i_row + 4s + k], B[k][col] = A[j_col + 4s + k], D = QT.
// Two waves per SIMD (RB = 2 takes 246 VGPRs): one wave's matrix-core work
// covers the other's epilogue.
template <int RB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_pre_rows(
    const double* __restrict__ A, const double* __restrict__ sInv, const double* __restrict__ mus,
    int P, int m, int step, int nrows, int excl, int ib, int cols_per_wave, int ncolwaves,
    unsigned long long* val, unsigned long long* key) {
    const int lane = threadIdx.x % kWave;
    const int64_t wave = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
    // wave-uniform (the buffer resources below must live in scalar registers)
    const int rb = __builtin_amdgcn_readfirstlane(static_cast<int>(wave / ncolwaves));
    const int cw = __builtin_amdgcn_readfirstlane(static_cast<int>(wave % ncolwaves));
    const int row0 = rb * 16 * RB;
    if (row0 >= nrows) return;
    const int c0 = cw * cols_per_wave;
    const int c1 = min(P, c0 + cols_per_wave);
    const int l16 = lane % 16, t0 = lane / 16;

    // Operands come through buffer loads: one resource for the rows' windows
    // (based at the block's first row), one for the column stream (based at
    // the wave's first column), a 32-bit per-lane offset and an immediate per
    // step.  A column read past the end of the series is out of the resource's
    // range and returns 0; it only feeds columns >= P or padded steps.
    const int64_t n = static_cast<int64_t>(P) + m - 1;
    const int64_t rbase = static_cast<int64_t>(row0) * step;
    auto nrec = [](int64_t elems) __attribute__((always_inline)) {
        const int64_t b = elems * 8;
        return b < 0x7FFFFFF0 ? static_cast<int>(b) : 0x7FFFFFF0;
    };
    // wave-uniform base addresses and sizes; a resource is built at each load
    // from them (scalar moves), which keeps it in scalar registers
    auto uni = [](const double* p) __attribute__((always_inline)) {
        const uint64_t v = reinterpret_cast<uint64_t>(p);
        const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v));
        const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v >> 32));
        return reinterpret_cast<double*>((static_cast<uint64_t>(hi) << 32) | lo);
    };
    double* const prow = uni(A + rbase);
    double* const pcol = uni(A + c0);
    const int brow = __builtin_amdgcn_readfirstlane(nrec(n - rbase));
    const int bcol = __builtin_amdgcn_readfirstlane(nrec(n - c0));
    auto bload = [](double* p, int bytes, int off) __attribute__((always_inline)) {
        const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(p, 0, bytes, kBufWord3);
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
    };
    // this lane's A-operand byte offset in each block (a row past the last
    // one reads the block's first row; its results are never merged)
    int aoff[RB];
    SP rbest[RB][4];
#pragma unroll
    for (int b = 0; b < RB; ++b) {
        const int arow = 16 * b + l16;
        aoff[b] = ((row0 + arow < nrows) ? arow * step + t0 : t0) * 8;
#pragma unroll
        for (int r = 0; r < 4; ++r) rbest[b][r] = SP{__builtin_nan(""), 0};
    }
    const int xoff = (l16 + t0) * 8;  // X(u) of chunk cb at xoff + 8 (cb - c0 + 4u)

    const int nsteps = (m + 3) / 4;
    const int nsteps_u = (nsteps + kU - 1) / kU * kU;  // steps past m read 0 (A-operand)
    const int G = m / 16;                                // groups of 4 steps inside the window
    // Column operand of sub-block c at step s is X(s + 4c), X(u) =
    // A[cb + l16 + 4u + t0]: a window of 16 X values rolls by 4 per group of 4
    // steps, so a step needs one new column load instead of four.  The loads
    // of a group are issued one group ahead (the first group of a chunk and
    // the chunk's first 12 window values before the previous chunk's
    // epilogue), so the matrix cores do not wait for them.
    double xw[16];
    double an[RB][4], xn[4];
    auto load_group = [&](int cbx, int s0) __attribute__((always_inline)) {
        const int xo = xoff + 8 * (cbx - c0 + 4 * (s0 + 12));
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            xn[e] = bload(pcol, bcol, xo + 32 * e);
#pragma unroll
            for (int b = 0; b < RB; ++b) an[b][e] = bload(prow, brow, aoff[b] + 32 * (s0 + e));
        }
    };
    auto load_window = [&](int cbx) __attribute__((always_inline)) {
        const int xo = xoff + 8 * (cbx - c0);
#pragma unroll
        for (int k = 0; k < 12; ++k) xw[k] = bload(pcol, bcol, xo + 32 * k);
    };
    load_window(c0);
    load_group(c0, 0);
    for (int cb = c0; cb < c1; cb += kColsPerChunk) {
        v4d acc[RB][4];
#pragma unroll
        for (int b = 0; b < RB; ++b)
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[b][c] = v4d{0.0, 0.0, 0.0, 0.0};
        // MASKED: the group reaches past the window (m % 16 != 0); its
        // A-operand elements t >= m are zeroed
        auto group = [&](int s0, auto mask_tag) __attribute__((always_inline)) {
            double a[RB][4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                xw[12 + e] = xn[e];
                const bool in = !decltype(mask_tag)::value || 4 * (s0 + e) + t0 < m;
#pragma unroll
                for (int b = 0; b < RB; ++b) a[b][e] = in ? an[b][e] : 0.0;
            }
            if (s0 + kU < nsteps_u) load_group(cb, s0 + kU);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
#pragma unroll
                for (int b = 0; b < RB; ++b)
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        acc[b][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[b][e], xw[e + 4 * c], acc[b][c], 0, 0, 0);
            }
#pragma unroll
            for (int k = 0; k < 12; ++k) xw[k] = xw[k + 4];
        };
        for (int g = 0; g < G; ++g) group(4 * g, std::false_type{});
        if (4 * G < nsteps_u) group(4 * G, std::true_type{});
        // the next chunk's first loads, issued before this chunk's epilogue
        load_window(cb + kColsPerChunk);
        load_group(cb + kColsPerChunk, 0);
        // the column statistics, and the column filter: the columns' current
        // values, read after the products (values only grow, so a later read
        // filters more)
        double cs[4], cm[4];
        unsigned long long cur[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int cj = min(cb + 16 * c + l16, P - 1);
            cs[c] = sInv[cj];
            cm[c] = mus[cj];
            cur[c] = val[cj] >> ib;
        }
        // epilogue: result (b, r, c) is QT(row 16b + 4r + t0, column cb + 16c + l16)
        SP cbest[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) cbest[c] = SP{__builtin_nan(""), 0};
#pragma unroll
        for (int b = 0; b < RB; ++b) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = row0 + 16 * b + 4 * r + t0;
                const bool rv = row < nrows;
                const int i = rv ? row * step : 0;
                const double rs = sInv[i], rm = mus[i];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int j = cb + 16 * c + l16;
                    const bool ok = rv && j < P && !(j > i - excl && j < i + excl);
                    const double sc = fma(acc[b][c][r], rs * cs[c], -(rm * cm[c]));
                    const double s = ok ? sc : __builtin_nan("");
                    // j rises along a lane's columns and i along its rows
                    rbest[b][r] = pick_ordered(rbest[b][r], SP{s, j});
                    cbest[c] = pick_ordered(cbest[c], SP{s, i});
                }
            }
        }
        // columns: the four lane quarters hold the other rows of the blocks.
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
    for (int b = 0; b < RB; ++b) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            SP x = rbest[b][r];
            x = ror_pick<8>(x);
            x = ror_pick<4>(x);
            x = ror_pick<2>(x);
            x = ror_pick<1>(x);
            const int row = row0 + 16 * b + 4 * r + t0;
            if (l16 == r && row < nrows && x.s == x.s) merge(val, key, static_cast<int64_t>(row) * step, x.s, x.p, ib);
        }
    }
}

// One lane per sampled row: walk the diagonal through (i, j = the row's best
// partner) step-1 cells forward and back (ScrimpppOrig.cpp:218-281).  QT of
// the start cell is recomputed in FP64 (the reference recovers it from the
// row's distance); the walk then updates it like the reference:
// QT += A[i+k+m-1]*A[j+k+m-1] - A[i+k-1]*A[j+k-1].
// Each sampled row's best partner once every row's distance profile has been
// merged and before any walk starts: a walk merges into rows that other walks
// read, so reading the keys inside k_pre_walk could start a walk from another
// row's (near-tied) update instead of the row's own minimum, which is where the
// reference starts it (ScrimpppOrig.cpp:215-217).
__global__ void k_pre_partners(const unsigned long long* __restrict__ key, int step, int nrows, int ib,
                               int* __restrict__ partner) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    partner[r] = static_cast<int>(key_partner(key[static_cast<int64_t>(r) * step], ib));
}

__global__ void k_pre_walk(const double* __restrict__ A, const double* __restrict__ sInv,
                           const double* __restrict__ mus, int P, int m, int step, int nrows,
                           int ib, const int* __restrict__ partner, unsigned long long* val,
                           unsigned long long* key) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const int i = r * step;
    const int j = partner[r];
    if (j < 0) return;
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
    // two blocks of 16 rows per wave: twice the MFMAs per column load (three
    // no longer fit in the VGPRs)
    constexpr int RB = 2;
    const int rblocks = (a.nrows + 16 * RB - 1) / (16 * RB);
    const int ncolwaves = (a.P + a.cols_per_wave - 1) / a.cols_per_wave;
    const int64_t waves = static_cast<int64_t>(rblocks) * ncolwaves;
    const int64_t groups = (waves + 3) / 4;
    if (groups > 0x7FFFFFFF) return MP_ERR_TOO_LONG;
    // the kernel's buffer offsets are 32-bit byte offsets from the block's
    // first row and from the wave's first column
    const int64_t row_span = (static_cast<int64_t>(16 * RB) * a.step + a.m + 16) * 8;
    const int64_t col_span = (static_cast<int64_t>(a.cols_per_wave) + kColsPerChunk + a.m + 128) * 8;
    if (row_span >= 0x7FFFFFF0 || col_span >= 0x7FFFFFF0) return MP_ERR_TOO_LONG;
    hipLaunchKernelGGL(k_pre_rows<RB>, dim3(static_cast<unsigned>(groups)), dim3(256), 0, st, a.A, a.sInv,
                       a.mus, a.P, a.m, a.step, a.nrows, a.excl, a.ib, a.cols_per_wave, ncolwaves,
                       a.val, a.key);
    if (hipGetLastError() != hipSuccess) return MP_ERR_DEVICE;
    const dim3 rgrid(static_cast<unsigned>((a.nrows + 255) / 256));
    hipLaunchKernelGGL(k_pre_partners, rgrid, dim3(256), 0, st, a.key, a.step, a.nrows, a.ib, a.partner);
    hipLaunchKernelGGL(k_pre_walk, rgrid, dim3(256), 0, st, a.A, a.sInv, a.mus, a.P, a.m, a.step, a.nrows,
                       a.ib, a.partner, a.val, a.key);
    return hipGetLastError() == hipSuccess ? MP_OK : MP_ERR_DEVICE;
}

}  // namespace mp
