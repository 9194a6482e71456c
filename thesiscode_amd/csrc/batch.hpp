// Host-side plan of a batch of FP32-reseeded self-joins.
#pragma once

#include <stdint.h>

#include <vector>

#include "reseeded_f32.hpp"

namespace mp {

struct BatchPlan {
    int nseries = 0;
    int band_k = 8, tile_len = 1024;
    std::vector<SeriesDesc> descs;
    std::vector<int64_t> eprefix;  // prefix of profile lengths
    std::vector<int64_t> sprefix;  // prefix of statistics chunks
    std::vector<int64_t> tprefix;  // prefix of tiles per series
    std::vector<int64_t> bprefix;  // per series: prefix of tiles over its bands
    int64_t total_n = 0, total_P = 0, ntiles = 0;
    double cells = 0.0;
    int build(int nseries, const int64_t* lengths, const int32_t* windows, int excl, int band_k,
              int tile_len);
};

}  // namespace mp
