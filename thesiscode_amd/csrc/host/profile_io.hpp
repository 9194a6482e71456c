// Reference file formats for time series and matrix profiles, with plain
// POSIX I/O instead of MPI-IO.
//
// Binary files start with a 100-byte header whose first 16 bytes are
// {uint32 format version, uint32 data offset (=100), uint64 element count}
// (scrimppp/inc/biniobase.h:13-21, scrimppp/src/biniobase.cpp:130-150).
//   time series: format 2, count doubles            (scrimppp/src/bintsfile.cpp:21-23)
//   profile:     format 1, count {double d2, uint64 index} records
//                                                   (scrimppp/inc/binproffile.h:11-16, 51)
// ASCII: series = whitespace-separated doubles (ScrimpSequ.hpp:48-62);
// profile = one "sqrt(|d2|) index" line per entry with 17 significant digits
// (ScrimpSequ.cpp:265-277).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "matprof_algo.hpp"

namespace matrix_profile {

constexpr uint32_t kTsFormat = 2;
constexpr uint32_t kProfileFormat = 1;
constexpr uint32_t kHeaderSize = 100;

// Resolves BY_SUFFIX from the input file name (".bin" / ".ascii", main.cpp:240-252);
// throws std::runtime_error when the suffix says neither.
Scrimppp_params::FILETYPE resolve_filetype(const Scrimppp_params& params);

// Default output name of main.cpp:255-267.
std::string default_output_name(const Scrimppp_params& params);

std::vector<double> read_time_series(const std::string& path, Scrimppp_params::FILETYPE type);
void write_time_series_bin(const std::string& path, const std::vector<double>& ts);

void store_matrix_profile(const double* profile, const int64_t* index, int64_t length,
                          const std::string& path, Scrimppp_params::FILETYPE type);
void read_matrix_profile_bin(const std::string& path, std::vector<double>& profile,
                             std::vector<int64_t>& index);

}  // namespace matrix_profile
