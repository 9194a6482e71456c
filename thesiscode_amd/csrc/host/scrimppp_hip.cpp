// scrimppp_hip: command line of the HIP self-join with the reference's flags
// and positional order (scrimppp/src/main.cpp:176-278):
//   scrimppp_hip [--algo] NAME [--window-length] M [--prescrimp-stride] S [--input-file] FILE
//                [--output-file O] [--blocklen B] [--filetype BIN|ASCII|BY_SUFFIX]
//                [--dist-io true|false] [--chunksize C] [--dbg]
//                [--band-k K] [--tile-len T] [--no-refine]       (HIP options)
// An unknown algorithm lists the registered ones and exits with 1
// (main.cpp:53-61); the runtime is printed like main.cpp:92-98.
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "matprof_algo.hpp"
#include "profile_io.hpp"

namespace matrix_profile {
void register_hip_algorithms();
}

using namespace matrix_profile;

namespace {

void usage() {
    std::cout << "usage: scrimppp_hip ALGO WINDOW_LENGTH PRESCRIMP_STRIDE INPUT_FILE [options]\n"
                 "  --algo NAME            algorithm: "
              << MatProfImplFactory::getInstance().get_available_implementations()
              << "\n  --window-length M      size m of the matching window\n"
                 "  --prescrimp-stride S   PreSCRIMP step floor(S*M) of scrimp_hip_orig (other algorithms ignore it)\n"
                 "  --input-file FILE      time series (.bin format 2 or .ascii)\n"
                 "  --output-file FILE     result (default <algo>_new_MatrixProfile_<m><file>.matprof.<type>)\n"
                 "  --blocklen B           accepted for compatibility (bands are sized on the device)\n"
                 "  --filetype T           BIN, ASCII or BY_SUFFIX\n"
                 "  --dist-io BOOL, --chunksize C, --dbg   accepted for compatibility\n"
                 "  --band-k K, --tile-len T, --no-refine  HIP kernel options\n";
}

bool parse_bool(const std::string& v) {
    return v == "1" || v == "true" || v == "TRUE" || v == "True" || v == "yes";
}

int parse_args(int argc, char** argv, Scrimppp_params& p) {
    std::vector<std::string> pos;
    bool have_m = false;
    p.output_filename.clear();
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto val = [&](const char* name) -> std::string {
            const std::string pre = std::string(name) + "=";
            if (a.rfind(pre, 0) == 0) return a.substr(pre.size());
            if (i + 1 >= argc) throw std::invalid_argument(std::string("missing value for ") + name);
            return argv[++i];
        };
        auto is = [&](const char* name) { return a == name || a.rfind(std::string(name) + "=", 0) == 0; };
        if (a == "--help" || a == "-h") {
            usage();
            return 1;
        } else if (is("--algo")) {
            p.algo_name = val("--algo");
        } else if (is("--window-length")) {
            p.query_window_len = std::stoi(val("--window-length"));
            have_m = true;
        } else if (is("--prescrimp-stride")) {
            p.prescrimp_stride = std::stoi(val("--prescrimp-stride"));
        } else if (is("--input-file")) {
            p.time_series_filename = val("--input-file");
        } else if (is("--output-file")) {
            p.output_filename = val("--output-file");
        } else if (is("--blocklen")) {
            p.block_length = std::stoi(val("--blocklen"));
        } else if (is("--filetype")) {
            const std::string t = val("--filetype");
            if (t == "BIN") p.filetype = Scrimppp_params::BIN;
            else if (t == "ASCII") p.filetype = Scrimppp_params::ASCII;
            else if (t == "BY_SUFFIX" || t == "SUFFIX") p.filetype = Scrimppp_params::BY_SUFFIX;
            else throw std::invalid_argument("unknown --filetype " + t);
        } else if (is("--dist-io")) {
            p.use_distributed_io = parse_bool(val("--dist-io"));
        } else if (is("--chunksize")) {
            p.writing_chunk_size = std::stoi(val("--chunksize"));
        } else if (a == "--dbg") {
        } else if (is("--band-k")) {
            p.band_k = std::stoi(val("--band-k"));
        } else if (is("--tile-len")) {
            p.tile_len = std::stoi(val("--tile-len"));
        } else if (a == "--no-refine") {
            p.refine = false;
        } else if (!a.empty() && a[0] == '-' && a.size() > 1 && !std::isdigit(static_cast<unsigned char>(a[1]))) {
            throw std::invalid_argument("unknown option " + a);
        } else {
            pos.push_back(a);
        }
    }
    // positional order: algo winlen stride infile (main.cpp:216)
    size_t k = 0;
    if (p.algo_name.empty() && k < pos.size()) p.algo_name = pos[k++];
    if (!have_m && k < pos.size()) {
        p.query_window_len = std::stoi(pos[k++]);
        have_m = true;
    }
    if (k < pos.size() && p.prescrimp_stride == 0) p.prescrimp_stride = std::stoi(pos[k++]);
    if (p.time_series_filename.empty() && k < pos.size()) p.time_series_filename = pos[k++];
    if (p.algo_name.empty() || !have_m || p.time_series_filename.empty())
        throw std::invalid_argument("the option '--algo', '--window-length' and '--input-file' are required");
    p.filetype = resolve_filetype(p);
    if (p.output_filename.empty()) p.output_filename = default_output_name(p);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    register_hip_algorithms();
    Scrimppp_params params;
    try {
        if (parse_args(argc, argv, params) != 0) return 0;
    } catch (const std::exception& e) {
        std::cerr << "ERROR parsing commandline arguments: " << e.what() << "\n";
        usage();
        return 1;
    }
    std::unique_ptr<IMatProfAlgo> algo;
    try {
        algo = MatProfImplFactory::getInstance().create(params.algo_name);
    } catch (const std::out_of_range&) {
        std::cerr << "Invalid algorithm specified: " << params.algo_name << ". Available: "
                  << MatProfImplFactory::getInstance().get_available_implementations() << "\n";
        return 1;
    }
    try {
        const auto t0 = std::chrono::steady_clock::now();
        algo->initialize(params);
        const auto t1 = std::chrono::steady_clock::now();
        algo->compute_matrix_profile(params);
        const auto t2 = std::chrono::steady_clock::now();
        algo->log_info();
        std::clog << "runtime: " << std::chrono::duration<double>(t2 - t1).count()
                  << " s (initialization " << std::chrono::duration<double>(t1 - t0).count() << " s)\n";
    } catch (const std::exception& e) {
        std::cerr << "Runtime ERROR: " << e.what() << "\n";
        return 2;
    }
    return 0;
}
