// Factory and the HIP-backed algorithm plugins.
//
// "scrimp_hip_f64"  -- the exact FP64 self-join: drop-in for scrimp_triv_par /
//                      scrimp_vert_blocked results (index init -1; the
//                      "_seqinit" variant uses 0 like scrimp_vert_blocked).
// "scrimp_hip_f32r" -- FP32 recurrence with FP64 statistics, re-seeding and
//                      refinement.
// Each plugin reads the series in the reference's formats, calls the C ABI
// and writes the profile in the reference's formats, like the CPU algorithms
// (ScrimpTrivialPar.cpp:142-368).
#include <chrono>
#include <cmath>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <vector>

#include "matprof_algo.hpp"
#include "profile_io.hpp"
#include "scrimp_hip.h"

namespace matrix_profile {

MatProfImplFactory& MatProfImplFactory::getInstance() {
    static MatProfImplFactory inst;
    return inst;
}

bool MatProfImplFactory::register_impl(const std::string& name, Creator creator_fn) {
    if (_registrations.count(name)) throw std::logic_error("algorithm registered twice: " + name);
    _registrations.emplace(name, creator_fn);
    return true;
}

std::string MatProfImplFactory::get_available_implementations() const {
    std::stringstream ss;
    for (const auto& kv : _registrations) ss << kv.first << " ";
    return ss.str();
}

std::unique_ptr<IMatProfAlgo> MatProfImplFactory::create(const std::string& algo_name) const {
    return _registrations.at(algo_name)();
}

namespace {

void throw_status(int rc, const char* where) {
    if (rc == MP_OK) return;
    std::stringstream s;
    s << where << ": " << mp_status_string(rc) << " (" << rc << ")";
    if (rc == MP_ERR_SHORT_SERIES || rc == MP_ERR_BAD_WINDOW)
        throw std::invalid_argument(s.str());
    throw std::runtime_error(s.str());
}

class ScrimpHipBase : public IMatProfAlgo {
public:
    void initialize(const Scrimppp_params&) override {}

    void compute_matrix_profile(const Scrimppp_params& params) override {
        const auto t0 = std::chrono::steady_clock::now();
        const auto type = resolve_filetype(params);
        std::vector<double> A = read_time_series(params.time_series_filename, type);
        const auto t1 = std::chrono::steady_clock::now();
        const int64_t n = static_cast<int64_t>(A.size());
        const int m = params.query_window_len;
        if (n - m + 1 <= 0)
            throw std::runtime_error("parameters chosen badly: resulting matrix profile of length 0. done.");
        const int64_t P = n - m + 1;
        std::vector<double> prof(P);
        std::vector<int64_t> idx(P);
        mp_opts o;
        mp_opts_default(&o);
        o.band_k = params.band_k;
        o.tile_len = params.tile_len;
        o.refine = params.refine ? 1 : 0;
        o.index_init = index_init();
        throw_status(run(A.data(), n, m, prof.data(), idx.data(), &o, &_timing), name());
        const auto t2 = std::chrono::steady_clock::now();
        const std::string out =
            params.output_filename.empty() ? default_output_name(params) : params.output_filename;
        store_matrix_profile(prof.data(), idx.data(), P, out, type);
        const auto t3 = std::chrono::steady_clock::now();
        _io_s = std::chrono::duration<double>((t1 - t0) + (t3 - t2)).count();
        _wall_s = std::chrono::duration<double>(t2 - t1).count();
        _P = P;
        _excl = m / 4;
    }

    void log_info() override {
        const double dev_s = _timing.total_ms * 1e-3;
        const double tri = static_cast<double>(_P - _excl);
        std::clog << std::setprecision(6) << name() << ": device time " << dev_s << " s (kernel "
                  << _timing.kernel_ms * 1e-3 << " s), call time " << _wall_s << " s, I/O time "
                  << _io_s << " s\n"
                  << "cells evaluated: " << _timing.cells << ", throughput " << _timing.cells / dev_s
                  << " cells/second\n"
                  << "throughput evaluations: " << tri * tri / dev_s
                  << " matrix entries/second (the reference's (P-ez)^2 count)\n";
    }

protected:
    virtual int run(const double* A, int64_t n, int m, double* prof, int64_t* idx, const mp_opts* o,
                    mp_timing* tm) = 0;
    virtual const char* name() const = 0;
    virtual int index_init() const { return -1; }

    mp_timing _timing{};
    double _io_s = 0.0, _wall_s = 0.0;
    int64_t _P = 0, _excl = 0;
};

class ScrimpHipF64 : public ScrimpHipBase {
protected:
    int run(const double* A, int64_t n, int m, double* p, int64_t* i, const mp_opts* o,
            mp_timing* tm) override {
        return mp_selfjoin_f64(A, n, m, p, i, o, tm);
    }
    const char* name() const override { return "scrimp_hip_f64"; }
};

class ScrimpHipF64SeqInit : public ScrimpHipF64 {
protected:
    const char* name() const override { return "scrimp_hip_f64_seqinit"; }
    int index_init() const override { return 0; }
};

class ScrimpHipF32r : public ScrimpHipBase {
protected:
    int run(const double* A, int64_t n, int m, double* p, int64_t* i, const mp_opts* o,
            mp_timing* tm) override {
        return mp_selfjoin_f32r(A, n, m, p, i, o, tm);
    }
    const char* name() const override { return "scrimp_hip_f32r"; }
};

}  // namespace

// explicit registration (static registration objects of a static library could
// be dropped by the linker); called by the CLI before the factory is used
void register_hip_algorithms() {
    static FactoryRegistration<ScrimpHipF64> r1("scrimp_hip_f64");
    static FactoryRegistration<ScrimpHipF64SeqInit> r2("scrimp_hip_f64_seqinit");
    static FactoryRegistration<ScrimpHipF32r> r3("scrimp_hip_f32r");
}

}  // namespace matrix_profile
