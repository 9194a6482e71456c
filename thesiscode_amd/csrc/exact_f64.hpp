// Host-side launchers of the exact FP64 path (defined in exact_f64.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstddef>
#include <mutex>

namespace mp {

// One window's record as the band walker reads it: A[i], A[i+m], 1/sigma and
// mu*sqrt(m)/sigma.  Rows past the triangle are padding records with
// 1/sigma = NaN, which masks every cell they touch.
struct alignas(32) WalkRec {
    double a, am, s, mu;
};
// padding records needed past the last row: one band of the widest walker and
// three of its 32-step windows (the ring form stages rows up to BW + 3F - 2
// past its last column)
constexpr int kWalkRecPad = 64 * 8 + 3 * 32;

// Device bytes the exact self-join carves from its scratch for n points, P
// windows and W walked diagonals (each buffer 256-byte aligned): A (n+64),
// prefix sums (2n), 1/sigma and mu*sqrt(m)/sigma (P+64 each), seeds (W+1),
// window records (P + kWalkRecPad), result values and keys (P each), counter.
constexpr size_t exact_aligned(int64_t count, size_t elem) {
    return (static_cast<size_t>(count) * elem + 255) & ~size_t(255);
}
constexpr size_t exact_scratch_bytes(int64_t n, int64_t P, int64_t W) {
    return exact_aligned(n + 64, 8) + 2 * exact_aligned(n, 8) + 2 * exact_aligned(P + 64, 8) +
           exact_aligned(W + 1, 8) + exact_aligned(P + kWalkRecPad, sizeof(WalkRec)) +
           2 * exact_aligned(P, 8) + exact_aligned(1, 4);
}

// The allocation-free exact self-join works in one file-scope device array of
// this library, sized for series of up to kNoAllocMaxPoints points (2^19, the
// longest request of the serving workload): about 50 MB that every GPU the
// library loads on holds until the process ends.
constexpr int64_t kNoAllocMaxPoints = int64_t(1) << 19;
constexpr size_t kNoAllocScratchBytes = exact_scratch_bytes(kNoAllocMaxPoints, kNoAllocMaxPoints, kNoAllocMaxPoints);
// Resolve the array's address on `device` (which must be current), once per
// device; later calls return the kept address.  MP_OK or MP_ERR_DEVICE.
int noalloc_scratch_init(int device);
// The array's device address on `device` (after noalloc_scratch_init) and the
// lock that serialises the calls using it.
int noalloc_scratch(int device, char** ptr, std::mutex** lock);

int launch_ref_stats(const double* A, int n, int m, double* cs, double* sq, double* sInv,
                     double* mus, hipStream_t st);
int launch_ref_seeds(const double* A, int base, int count, int m, double* qt, hipStream_t st);
int launch_fill_u64(unsigned long long* p, int64_t count, unsigned long long v, hipStream_t st);
int launch_fill_u32(unsigned int* p, int count, unsigned int v, hipStream_t st);
// profile/index from the merged (value, key) arrays.  band_k: the band width of
// the walker whose keys these are (1: keys hold partners; 2, 4, 8: keys hold
// candidate codes, resolved against A/sInv/mus, base = exclusion + 1).
int launch_unpack_f64(const unsigned long long* val, const unsigned long long* key, int P, int m,
                      int ib, int64_t index_init, int band_k, const double* A, const double* sInv,
                      const double* mus, int base, double* profile, int64_t* index, hipStream_t st);
int launch_pack_block(const double* prof, int count, int ib, unsigned long long* val,
                      unsigned long long* key, hipStream_t st);
int launch_unpack_block(const unsigned long long* val, const unsigned long long* key, int count,
                        int ib, int64_t pmin, const int64_t* idx_in, double* prof_out,
                        int64_t* idx_out, hipStream_t st);
// out[i] = {A[i], A[i+m], s[i], mu[i]} for i < valid, padding records up to total
int launch_pack_rec(const double* A, const double* s, const double* mu, int m, int valid,
                    int total, WalkRec* out, hipStream_t st);

// Column segments.  On short series the bands are too few to occupy the GPU and
// the walk lasts as long as its longest band.  The DPP-row walker (K <= 4) can
// then cut band b (L_b = tl - b*BW columns) at every multiple of SEG = r*BW
// columns, into n_b = ceil(L_b / SEG) = ceil((nbands - b) / r) segments, and
// hand them out band by band: item i is segment s of the band b with
// seg_first_item(b) <= i < seg_first_item(b + 1), s = i - seg_first_item(b).
// A segment s > 0 starts from its diagonals' running dot products at column
// s*SEG, which k_seg_seeds computes beforehand by the walker's own recurrence
// and stores at ckpt[i * BW + (d - b*BW)].  A row whose cells a segment
// boundary splits gets one partial result per segment, merged by the result
// keys as the parts of a row that two bands cover are.
constexpr int kSegCols = 2048;                    // SEG, a multiple of every BW <= 256
constexpr int64_t kSegMaxW = int64_t(1) << 18;    // longest walk cut into segments
// sum over u = 1 .. U of ceil(u / r)
__host__ __device__ constexpr int64_t seg_ceil_sum(int64_t U, int64_t r) {
    const int64_t q = U / r, t = U % r;
    return r * q * (q + 1) / 2 + t * (q + 1);
}
__host__ __device__ constexpr int64_t seg_first_item(int64_t b, int64_t nbands, int64_t r) {
    return seg_ceil_sum(nbands, r) - seg_ceil_sum(nbands - b, r);
}
__host__ __device__ constexpr int64_t seg_items(int64_t nbands, int64_t r) { return seg_ceil_sum(nbands, r); }
// the band of item i: the largest b with seg_first_item(b) <= i
__host__ __device__ inline int seg_band_of(int64_t i, int nbands, int r) {
    int lo = 0, hi = nbands - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (seg_first_item(mid, nbands, r) <= i)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}
// device bytes of the segment start values for W diagonals at band width K
constexpr size_t seg_ckpt_bytes(int64_t W, int band_k) {
    const int64_t bw = 64 * band_k;
    const int64_t nb = (W + bw - 1) / bw;
    return exact_aligned(seg_items(nb, kSegCols / bw) * bw, 8);
}

struct WalkF64Launch {
    const WalkRec* colrec;  // column j: colrec[j], j < tl
    const WalkRec* rowrec;  // row r: rowrec[r], r < tl + kWalkRecPad (padding past tl)
    double* qt;
    unsigned long long *rowval, *rowkey, *colval, *colkey;
    unsigned int* counter;
    int64_t row_off, col_off;
    int blocklen, tl, m, ib, writeback, band_k, groups;
    int form;               // band_k = 8: 1 rows from an LDS ring, one wave per SIMD
                            // (k_walk_f64_wide); 2 rows by DPP, two waves per SIMD
                            // (k_walk_f64_pair); 3 rows in place, diagonals rotating
                            // by DPP, two waves per SIMD (k_walk_f64_rot, no write-back);
                            // 4 (band_k <= 4) column segments (k_walk_f64_seg, no write-back)
    double* ckpt;           // form 4: segment start values (seg_ckpt_bytes)
};
// form 4: the number of segments (work items) of the launch
int64_t walk_f64_seg_items(int blocklen, int band_k);
int walk_f64_nbands(int blocklen, int band_k);
int launch_walk_f64(const WalkF64Launch& a, hipStream_t st);

}  // namespace mp
