// Host-side launcher of the PreSCRIMP kernels (defined in prescrimp.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mp {

struct PreScrimpLaunch {
    const double* A;         // series, n points
    const double* sInv;      // 1/sigma per window (exact path's statistics)
    const double* mus;       // mu*sqrt(m)/sigma per window
    int P, m;                // windows, window length
    int step;                // floor(stride*m) >= 1: sampled rows 0, step, 2*step, ...
    int nrows;               // ceil(P / step)
    int excl;                // m/4: partners with |i - j| < excl are excluded
    int ib;                  // partner bits of the merge keys
    int cols_per_wave;       // a multiple of 64
    unsigned long long *val, *key;  // P entries each, filled with ord(-inf)
    int* partner;            // nrows entries: each sampled row's best partner before the walks
};

int launch_prescrimp(const PreScrimpLaunch& a, hipStream_t st);

}  // namespace mp
