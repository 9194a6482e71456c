// Converters between the reference's file formats (scrimppp/cvt_src/):
//   cvt_ts_hip ascii-to-bin INPUT [OUTPUT]       time series ASCII -> BIN format 2
//                                                (cvt_ts.cpp:58-82)
//   cvt_matprof_hip bin-to-ascii INPUT [OUTPUT]  profile BIN format 1 -> ASCII
//                                                (cvt_matprof.cpp:67-83)
// The output defaults to "converted.bin" / "converted.ascii" (cvt_cli.cpp:78-89).
// The reverse directions are not implemented in the reference either
// (cvt_ts.cpp:85-87 throws, cvt_matprof.cpp:63-65 is empty); they are provided here.
// One binary serves both names: the tool is chosen from argv[0].
#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "profile_io.hpp"

using namespace matrix_profile;

namespace {

int convert(const std::string& tool, const std::string& mode, const std::string& in, std::string out) {
    std::ifstream probe(in);
    if (!probe.good()) throw std::runtime_error("Input file not existing!");
    const bool to_bin = mode == "ascii-to-bin";
    if (!to_bin && mode != "bin-to-ascii") throw std::invalid_argument("unknown conversion mode " + mode);
    if (out.empty()) out = to_bin ? "converted.bin" : "converted.ascii";
    if (tool.find("matprof") != std::string::npos) {
        std::vector<double> p;
        std::vector<int64_t> i;
        if (to_bin) {
            // ASCII profiles hold sqrt(|d2|): squaring restores d2 (the sign is lost, as in the reference)
            std::ifstream f(in);
            double d;
            long long k;
            while (f >> d >> k) {
                p.push_back(d * d);
                i.push_back(k);
            }
            store_matrix_profile(p.data(), i.data(), static_cast<int64_t>(p.size()), out, Scrimppp_params::BIN);
        } else {
            read_matrix_profile_bin(in, p, i);
            store_matrix_profile(p.data(), i.data(), static_cast<int64_t>(p.size()), out, Scrimppp_params::ASCII);
        }
    } else {
        if (to_bin) {
            write_time_series_bin(out, read_time_series(in, Scrimppp_params::ASCII));
        } else {
            const std::vector<double> ts = read_time_series(in, Scrimppp_params::BIN);
            std::ofstream f(out);
            f.precision(17);
            for (double v : ts) f << v << "\n";
        }
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    const std::string tool = argv[0];
    if (argc < 3 || !std::strcmp(argv[1], "--help")) {
        std::cerr << "usage: " << tool << " ascii-to-bin|bin-to-ascii INPUT [OUTPUT]\n";
        return 1;
    }
    try {
        return convert(tool, argv[1], argv[2], argc > 3 ? argv[3] : "");
    } catch (const std::exception& e) {
        std::cerr << "Runtime ERROR: " << e.what() << "\n";
        return 1;
    }
}
