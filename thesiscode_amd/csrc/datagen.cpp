// Synthetic input generators for the SCRIMP self-join workloads.
//
// Every series is produced by a counter-based Philox4x32-10 generator
// (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy as 1, 2, 3",
// SC'11) followed by the Box-Muller transform.  Everything is plain C++ with
// glibc's log/sqrt/sin/cos, so the same binary produces the same bits on the
// build container and on the GPU host (same image, same libm).
//
// Stream layout (key = 64-bit seed, counter = {pair lo, pair hi, stream, 0}):
//   normal #t of stream s  -> pair t/2 of stream s, first or second output.
//   uniform #t of stream s -> 53-bit draw from counter {t, 0, s, 1}.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "scrimp_hip.h"

namespace {

struct Philox4x32 {
    uint32_t v[4];
};

inline void mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
    const uint64_t p = static_cast<uint64_t>(a) * b;
    hi = static_cast<uint32_t>(p >> 32);
    lo = static_cast<uint32_t>(p);
}

// Ten rounds of Philox4x32 with the published multipliers and Weyl key bumps.
Philox4x32 philox10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint64_t seed) {
    uint32_t k0 = static_cast<uint32_t>(seed), k1 = static_cast<uint32_t>(seed >> 32);
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo(0xD2511F53u, c0, hi0, lo0);
        mulhilo(0xCD9E8D57u, c2, hi1, lo1);
        const uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return {{c0, c1, c2, c3}};
}

constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double kInv2p53 = 1.0 / 9007199254740992.0;

// 53-bit uniform in [0,1) from two 32-bit words.
inline double u53(uint32_t a, uint32_t b) {
    return (static_cast<double>(a >> 5) * 67108864.0 + static_cast<double>(b >> 6)) * kInv2p53;
}

inline void normal_pair(uint64_t seed, uint32_t stream, uint64_t pair, double& z0, double& z1) {
    const Philox4x32 x = philox10(static_cast<uint32_t>(pair), static_cast<uint32_t>(pair >> 32),
                                  stream, 0u, seed);
    const double u1 = 1.0 - u53(x.v[0], x.v[1]);  // (0, 1]
    const double u2 = u53(x.v[2], x.v[3]);        // [0, 1)
    const double r = std::sqrt(-2.0 * std::log(u1));
    z0 = r * std::cos(kTwoPi * u2);
    z1 = r * std::sin(kTwoPi * u2);
}

}  // namespace

extern "C" {

void mp_gen_philox4x32_10(const uint32_t ctr[4], uint64_t key, uint32_t out[4]) {
    const Philox4x32 x = philox10(ctr[0], ctr[1], ctr[2], ctr[3], key);
    for (int i = 0; i < 4; ++i) out[i] = x.v[i];
}

void mp_gen_normals(uint64_t seed, uint32_t stream, int64_t first, int64_t count, double* out) {
    for (int64_t i = 0; i < count; ++i) {
        const uint64_t t = static_cast<uint64_t>(first + i);
        double z0, z1;
        normal_pair(seed, stream, t >> 1, z0, z1);
        out[i] = (t & 1u) ? z1 : z0;
    }
}

void mp_gen_uniforms(uint64_t seed, uint32_t stream, int64_t first, int64_t count, double* out) {
    for (int64_t i = 0; i < count; ++i) {
        const uint64_t t = static_cast<uint64_t>(first + i);
        const Philox4x32 x = philox10(static_cast<uint32_t>(t), static_cast<uint32_t>(t >> 32),
                                      stream, 1u, seed);
        out[i] = u53(x.v[0], x.v[1]);
    }
}

// x[0] = 0, x[t] = x[t-1] + N(0,1); the step for x[t] is normal #(t-1) of the stream.
void mp_gen_random_walk(uint64_t seed, uint32_t stream, int64_t n, double* out) {
    if (n <= 0) return;
    out[0] = 0.0;
    double acc = 0.0;
    for (int64_t t = 1; t < n; t += 2) {
        double z0, z1;
        normal_pair(seed, stream, static_cast<uint64_t>(t - 1) >> 1, z0, z1);
        acc += z0;
        out[t] = acc;
        if (t + 1 < n) {
            acc += z1;
            out[t + 1] = acc;
        }
    }
}

// Precision workload: random walk (stream 0) + three sinusoids of amplitude 5
// (periods U[200,5000], phases U[0,2pi), uniforms 0..5 of stream 1) + 20 Gaussian
// bumps of width m (std m/6 over an m-point support, centred), height 8 sigma where
// sigma is the population std of the host window before any bump is added;
// bump b starts at floor(u_b (n-m)) with u_b uniform #b of stream 2.
// meta (optional, 26 doubles): periods[3], phases[3], starts[20].
int mp_gen_precision_series(uint64_t seed, int64_t n, int32_t m, double* out, double* meta) {
    if (n <= 0 || m <= 0 || m > n) return MP_ERR_BAD_WINDOW;
    mp_gen_random_walk(seed, 0u, n, out);
    double u[6];
    mp_gen_uniforms(seed, 1u, 0, 6, u);
    double per[3], phs[3];
    for (int k = 0; k < 3; ++k) {
        per[k] = 200.0 + 4800.0 * u[k];
        phs[k] = kTwoPi * u[3 + k];
    }
    for (int64_t t = 0; t < n; ++t) {
        double s = 0.0;
        for (int k = 0; k < 3; ++k) s += 5.0 * std::sin(kTwoPi * static_cast<double>(t) / per[k] + phs[k]);
        out[t] += s;
    }
    double ub[20];
    mp_gen_uniforms(seed, 2u, 0, 20, ub);
    int64_t start[20];
    double height[20];
    for (int b = 0; b < 20; ++b) {
        start[b] = static_cast<int64_t>(std::floor(ub[b] * static_cast<double>(n - m)));
        double mu = 0.0;
        for (int t = 0; t < m; ++t) mu += out[start[b] + t];
        mu /= m;
        double var = 0.0;
        for (int t = 0; t < m; ++t) {
            const double d = out[start[b] + t] - mu;
            var += d * d;
        }
        height[b] = 8.0 * std::sqrt(var / m);
    }
    const double sd = static_cast<double>(m) / 6.0, ctr = 0.5 * static_cast<double>(m - 1);
    for (int b = 0; b < 20; ++b) {
        for (int t = 0; t < m; ++t) {
            const double z = (static_cast<double>(t) - ctr) / sd;
            out[start[b] + t] += height[b] * std::exp(-0.5 * z * z);
        }
    }
    if (meta) {
        for (int k = 0; k < 3; ++k) { meta[k] = per[k]; meta[3 + k] = phs[k]; }
        for (int b = 0; b < 20; ++b) meta[6 + b] = static_cast<double>(start[b]);
    }
    return MP_OK;
}

}  // extern "C"
