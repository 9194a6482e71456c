// Shared device helpers for the SCRIMP HIP kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "scrimp_hip.h"

namespace mp {

constexpr int kWave = 64;
constexpr int kWavesPerGroup = 4;
constexpr int kGroup = kWave * kWavesPerGroup;
constexpr int kFlush = 16;  // steps between flushes of row/column results

// Monotone map of a double to uint64: a < b  <=>  ord(a) < ord(b) (NaN excluded).
__device__ __forceinline__ uint64_t ord_f64(double x) {
    const uint64_t b = static_cast<uint64_t>(__double_as_longlong(x));
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ __forceinline__ double unord_f64(uint64_t o) {
    const uint64_t b = (o >> 63) ? (o & 0x7FFFFFFFFFFFFFFFull) : ~o;
    double d;
    __builtin_memcpy(&d, &b, sizeof d);
    return d;
}
__host__ __device__ __forceinline__ uint64_t ord_f64_host(double x) {
    uint64_t b;
    __builtin_memcpy(&b, &x, sizeof b);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

// Key of the exact path: the top (64-ib) bits of ord(score) with the partner
// index in the low ib bits, stored as (2^ib - 2 - partner) so that among equal
// truncated scores the smallest partner wins; 2^ib - 1 is the "no partner"
// sentinel, which never loses an exact tie.
__host__ __device__ __forceinline__ uint64_t make_key(uint64_t ord, int ib, int64_t partner) {
    const uint64_t mask = (1ull << ib) - 1ull;
    return (ord & ~mask) | (mask - 1ull - static_cast<uint64_t>(partner));
}
__host__ __device__ __forceinline__ int64_t key_partner(uint64_t key, int ib) {
    const uint64_t mask = (1ull << ib) - 1ull;
    const uint64_t low = key & mask;
    return (low == mask) ? -1 : static_cast<int64_t>(mask - 1ull - low);
}

// Lane l receives lane l+1's value (DPP wave_shl:1); lane 63 keeps `fill`.
__device__ __forceinline__ int shift_down_i32(int src, int fill) {
    return __builtin_amdgcn_update_dpp(fill, src, 0x130, 0xF, 0xF, false);
}
__device__ __forceinline__ double shift_down_f64(double src, double fill) {
    const uint64_t s = static_cast<uint64_t>(__double_as_longlong(src));
    const uint64_t f = static_cast<uint64_t>(__double_as_longlong(fill));
    const int lo = shift_down_i32(static_cast<int>(s), static_cast<int>(f));
    const int hi = shift_down_i32(static_cast<int>(s >> 32), static_cast<int>(f >> 32));
    return __longlong_as_double(static_cast<long long>(
        (static_cast<uint64_t>(static_cast<uint32_t>(hi)) << 32) | static_cast<uint32_t>(lo)));
}
// Lane 63 receives 0 bits (bound_ctrl), everything else lane l+1's value.
__device__ __forceinline__ int shift_down_i32_zero(int src) {
    return __builtin_amdgcn_update_dpp(0, src, 0x130, 0xF, 0xF, true);
}
__device__ __forceinline__ float shift_down_f32(float src, float fill) {
    return __int_as_float(shift_down_i32(__float_as_int(src), __float_as_int(fill)));
}
// A wave-uniform double (scalar registers) into vector registers with one
// 64-bit move, where the compiler would issue two 32-bit moves: the fill value
// of a shift_down_f64, whose DPP moves then overwrite lanes 0..62 in place.
__device__ __forceinline__ double vmov64(double s) {
    double d;
    asm("v_mov_b64 %0, %1" : "=v"(d) : "s"(s));
    return d;
}
// Lane 63 receives +0.0.
__device__ __forceinline__ float shift_down_f32_zero(float src) {
    return __int_as_float(shift_down_i32_zero(__float_as_int(src)));
}
// Lane 63 receives the double whose high word is fill_hi and low word 0 (for
// -inf: fill_hi = 0xFFF00000).  The low word needs no fill register.
__device__ __forceinline__ double shift_down_f64_lo0(double src, int fill_hi) {
    const uint64_t s = static_cast<uint64_t>(__double_as_longlong(src));
    const int lo = shift_down_i32_zero(static_cast<int>(s));
    const int hi = shift_down_i32(static_cast<int>(s >> 32), fill_hi);
    return __longlong_as_double(static_cast<long long>(
        (static_cast<uint64_t>(static_cast<uint32_t>(hi)) << 32) | static_cast<uint32_t>(lo)));
}

__device__ __forceinline__ int uniform_i32(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Grab the next work item for the whole wave from a global counter.
__device__ __forceinline__ int next_item(unsigned int* counter, int lane) {
    unsigned int v = 0;
    if (lane == 0) v = atomicAdd(counter, 1u);
    return uniform_i32(static_cast<int>(v));
}

}  // namespace mp

#define MP_HIP_CHECK(expr)                                        \
    do {                                                          \
        hipError_t mp_err_ = (expr);                              \
        if (mp_err_ != hipSuccess) return MP_ERR_DEVICE;          \
    } while (0)
