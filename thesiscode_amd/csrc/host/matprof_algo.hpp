// Algorithm-plugin interface of the SCRIMP host layer.
//
// Same names, argument meaning and error behaviour as the reference's plugin
// API (scrimppp/inc/MatProfAlgo.hpp:15-42, scrimppp/inc/ScrimpppParams.hpp:6-28,
// scrimppp/src/MatProfAlgo.cpp:35-37), so a caller selects "scrimp_hip_f64"
// exactly as it selects "scrimp_triv_par": the factory throws
// std::out_of_range for an unknown name, algorithms throw std::runtime_error
// or std::invalid_argument on bad input and write their own output file.
#pragma once

#include <functional>
#include <map>
#include <memory>
#include <string>

namespace matrix_profile {

struct Scrimppp_params {
    std::string algo_name;
    std::string time_series_filename;
    std::string output_filename = "matprof";
    int query_window_len = 0;
    int prescrimp_stride = 0;
    int block_length = 500;
    enum FILETYPE { BY_SUFFIX, ASCII, BIN } filetype = BY_SUFFIX;
    bool use_distributed_io = true;
    int writing_chunk_size = 0;
    // options of the HIP algorithms (no reference counterpart): 0 = default
    int band_k = 0;
    int tile_len = 0;
    bool refine = true;
};

class IMatProfAlgo {
public:
    virtual ~IMatProfAlgo() = default;
    virtual void initialize(const Scrimppp_params& params) = 0;
    virtual void compute_matrix_profile(const Scrimppp_params& params) = 0;
    virtual void log_info() {}
};

class MatProfImplFactory {
public:
    using Creator = std::unique_ptr<IMatProfAlgo> (*)();
    static MatProfImplFactory& getInstance();
    bool register_impl(const std::string& name, Creator creator_fn);
    std::string get_available_implementations() const;
    std::unique_ptr<IMatProfAlgo> create(const std::string& algo_name) const;  // throws std::out_of_range

private:
    MatProfImplFactory() = default;
    std::map<std::string, Creator> _registrations;
};

template <class impl>
class FactoryRegistration {
public:
    explicit FactoryRegistration(const std::string& name) {
        MatProfImplFactory::getInstance().register_impl(
            name, []() { return std::unique_ptr<IMatProfAlgo>(new impl()); });
    }
};

}  // namespace matrix_profile
