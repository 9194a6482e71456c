// Factory and the HIP-backed algorithm plugins.
//
// "scrimp_hip_f64"  -- the exact FP64 self-join: drop-in for scrimp_triv_par /
//                      scrimp_vert_blocked results (index init -1; the
//                      "_seqinit" variant uses 0 like scrimp_vert_blocked).
// "scrimp_hip_f32r" -- FP32 recurrence with FP64 statistics, re-seeding and
//                      refinement.
// "scrimp_hip_orig" -- drop-in for scrimppp_orig (ScrimpppOrig.cpp:37-402):
//                      PreSCRIMP written to <out>_prescrimp, then the full
//                      join, whose profile starts from PreSCRIMP's.
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

// scrimppp_orig reads its series as ASCII text and writes both profiles as
// ASCII whatever the file names say (ScrimpppOrig.cpp:56-70, 303-310,
// 392-399); its step is floor(prescrimp_stride * m) with the integer
// prescrimp_stride parameter (ScrimpppOrig.cpp:51).  Its SCRIMP phase starts
// from the PreSCRIMP profile and lowers it (strict <), so the final profile is
// the element-wise minimum of PreSCRIMP's and the full join's.
class ScrimpHipOrig : public IMatProfAlgo {
public:
    void initialize(const Scrimppp_params&) override {}

    void compute_matrix_profile(const Scrimppp_params& params) override {
        const auto t0 = std::chrono::steady_clock::now();
        std::vector<double> A = read_time_series(params.time_series_filename, Scrimppp_params::ASCII);
        const int64_t n = static_cast<int64_t>(A.size());
        const int m = params.query_window_len;
        if (n < m) throw std::runtime_error("ERROR: Time series is shorter than the window length, can not proceed");
        const int64_t P = n - m + 1;
        const std::string out =
            params.output_filename.empty() ? default_output_name(params) : params.output_filename;
        std::vector<double> pre(P), prof(P);
        std::vector<int64_t> pidx(P), idx(P);
        const auto t1 = std::chrono::steady_clock::now();
        throw_status(mp_prescrimp_f64(A.data(), n, m, static_cast<double>(params.prescrimp_stride), pre.data(),
                                      pidx.data(), &_pre),
                     "scrimp_hip_orig (PreSCRIMP)");
        const auto t2 = std::chrono::steady_clock::now();
        store_matrix_profile(pre.data(), pidx.data(), P, out + "_prescrimp", Scrimppp_params::ASCII);
        const auto t3 = std::chrono::steady_clock::now();
        mp_opts o;
        mp_opts_default(&o);
        o.index_init = -1;
        throw_status(mp_selfjoin_f64(A.data(), n, m, prof.data(), idx.data(), &o, &_full), "scrimp_hip_orig");
        for (int64_t i = 0; i < P; ++i) {
            if (idx[i] < 0 || pre[i] < prof[i]) {  // unreached by any diagonal, or PreSCRIMP's is smaller
                prof[i] = pre[i];
                idx[i] = pidx[i];
            }
        }
        const auto t4 = std::chrono::steady_clock::now();
        store_matrix_profile(prof.data(), idx.data(), P, out, Scrimppp_params::ASCII);
        const auto t5 = std::chrono::steady_clock::now();
        _wall_s = std::chrono::duration<double>((t2 - t1) + (t4 - t3)).count();
        _io_s = std::chrono::duration<double>((t1 - t0) + (t3 - t2) + (t5 - t4)).count();
    }

    void log_info() override {
        std::clog << std::setprecision(6) << "scrimp_hip_orig: PreSCRIMP device time " << _pre.total_ms * 1e-3
                  << " s (" << _pre.work_items << " sampled rows), full join device time " << _full.total_ms * 1e-3
                  << " s, call time " << _wall_s << " s, I/O time " << _io_s << " s\n";
    }

private:
    mp_timing _pre{}, _full{};
    double _io_s = 0.0, _wall_s = 0.0;
};

}  // namespace

// explicit registration (static registration objects of a static library could
// be dropped by the linker); called by the CLI before the factory is used
void register_hip_algorithms() {
    static FactoryRegistration<ScrimpHipF64> r1("scrimp_hip_f64");
    static FactoryRegistration<ScrimpHipF64SeqInit> r2("scrimp_hip_f64_seqinit");
    static FactoryRegistration<ScrimpHipF32r> r3("scrimp_hip_f32r");
    static FactoryRegistration<ScrimpHipOrig> r4("scrimp_hip_orig");
}

}  // namespace matrix_profile
