#include "profile_io.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <sstream>
#include <stdexcept>

namespace matrix_profile {

namespace {

struct Header {
    uint32_t fmt;
    uint32_t offset;
    uint64_t count;
};
static_assert(sizeof(Header) == 16, "header layout");

std::string basename_of(const std::string& path) {
    const auto pos = path.find_last_of('/');
    return pos == std::string::npos ? path : path.substr(pos + 1);
}

Header read_header(std::FILE* f, uint32_t expected_fmt) {
    Header h{};
    if (std::fread(&h, sizeof h, 1, f) != 1) throw std::runtime_error("could not read the binary file header");
    // biniobase.cpp:73-78: both checks throw
    if (h.fmt != expected_fmt || h.offset != kHeaderSize)
        throw std::runtime_error(
            "The binary input file uses a format different from the one the program was built for.");
    return h;
}

void write_header(std::FILE* f, uint32_t fmt, uint64_t count) {
    unsigned char buf[kHeaderSize];
    std::memset(buf, 0, sizeof buf);
    const Header h{fmt, kHeaderSize, count};
    std::memcpy(buf, &h, sizeof h);
    if (std::fwrite(buf, 1, sizeof buf, f) != sizeof buf) throw std::runtime_error("header write failed");
}

struct File {
    std::FILE* f;
    File(const std::string& p, const char* mode) : f(std::fopen(p.c_str(), mode)) {
        if (!f) throw std::runtime_error("Could not open file " + p);
    }
    ~File() {
        if (f) std::fclose(f);
    }
};

}  // namespace

Scrimppp_params::FILETYPE resolve_filetype(const Scrimppp_params& params) {
    if (params.filetype != Scrimppp_params::BY_SUFFIX) return params.filetype;
    if (params.time_series_filename.rfind(".bin") != std::string::npos) return Scrimppp_params::BIN;
    if (params.time_series_filename.rfind(".ascii") != std::string::npos) return Scrimppp_params::ASCII;
    throw std::runtime_error(
        "file type could not be determined from suffix. Specify explicitly or rename the file...");
}

std::string default_output_name(const Scrimppp_params& params) {
    std::stringstream s;
    s << params.algo_name << "_new_MatrixProfile_" << params.query_window_len
      << basename_of(params.time_series_filename) << ".matprof";
    const auto t = resolve_filetype(params);
    if (t == Scrimppp_params::BIN) s << ".bin";
    if (t == Scrimppp_params::ASCII) s << ".ascii";
    return s.str();
}

std::vector<double> read_time_series(const std::string& path, Scrimppp_params::FILETYPE type) {
    std::vector<double> a;
    if (type == Scrimppp_params::ASCII) {
        std::ifstream in(path);
        if (!in.is_open()) throw std::runtime_error("Could not open input file");
        double v;
        while (in >> v) a.push_back(v);
        return a;
    }
    if (type != Scrimppp_params::BIN) throw std::runtime_error("Unhandled file type detected.");
    File f(path, "rb");
    const Header h = read_header(f.f, kTsFormat);
    a.resize(h.count);
    if (std::fseek(f.f, h.offset, SEEK_SET) != 0) throw std::runtime_error("seek failed");
    if (h.count && std::fread(a.data(), sizeof(double), h.count, f.f) != h.count)
        throw std::runtime_error("time series file is shorter than its header says");
    return a;
}

void write_time_series_bin(const std::string& path, const std::vector<double>& ts) {
    File f(path, "wb");
    write_header(f.f, kTsFormat, ts.size());
    if (!ts.empty() && std::fwrite(ts.data(), sizeof(double), ts.size(), f.f) != ts.size())
        throw std::runtime_error("time series write failed");
}

void store_matrix_profile(const double* profile, const int64_t* index, int64_t length,
                          const std::string& path, Scrimppp_params::FILETYPE type) {
    if (path.empty()) throw std::runtime_error("Empty output file name specified!");
    if (type == Scrimppp_params::ASCII) {
        std::ofstream out(path);
        if (!out.is_open()) throw std::runtime_error("Could not open output file " + path);
        for (int64_t i = 0; i < length; ++i) {
            out << std::setprecision(std::numeric_limits<double>::digits10 + 2)
                << std::sqrt(std::fabs(profile[i])) << ' '
                << std::setprecision(std::numeric_limits<int>::digits10 + 1) << index[i] << '\n';
        }
        return;
    }
    if (type != Scrimppp_params::BIN) throw std::runtime_error("Unhandled file type detected.");
    File f(path, "wb");
    write_header(f.f, kProfileFormat, static_cast<uint64_t>(length));
    std::vector<unsigned char> buf(16 * 4096);
    for (int64_t s = 0; s < length; s += 4096) {
        const int64_t c = std::min<int64_t>(4096, length - s);
        for (int64_t k = 0; k < c; ++k) {
            const uint64_t idx = static_cast<uint64_t>(index[s + k]);  // binproffile.h:12 unsigned long
            std::memcpy(&buf[16 * k], &profile[s + k], 8);
            std::memcpy(&buf[16 * k + 8], &idx, 8);
        }
        if (std::fwrite(buf.data(), 16, static_cast<size_t>(c), f.f) != static_cast<size_t>(c))
            throw std::runtime_error("profile write failed");
    }
}

void read_matrix_profile_bin(const std::string& path, std::vector<double>& profile,
                             std::vector<int64_t>& index) {
    File f(path, "rb");
    const Header h = read_header(f.f, kProfileFormat);
    if (h.count > static_cast<uint64_t>(std::numeric_limits<int64_t>::max()))
        throw std::runtime_error("the length of the stored matrix profile exceeds the index types maximum");
    profile.resize(h.count);
    index.resize(h.count);
    if (std::fseek(f.f, h.offset, SEEK_SET) != 0) throw std::runtime_error("seek failed");
    std::vector<unsigned char> rec(16);
    for (uint64_t i = 0; i < h.count; ++i) {
        if (std::fread(rec.data(), 16, 1, f.f) != 1) throw std::runtime_error("profile file truncated");
        std::memcpy(&profile[i], &rec[0], 8);
        uint64_t u;
        std::memcpy(&u, &rec[8], 8);
        index[i] = static_cast<int64_t>(u);
    }
}

}  // namespace matrix_profile
