// Batched FP32-reseeded self-joins: many independent series of uneven length
// and window in one persistent launch, captured once in a hipGraph and
// replayed; plus the autotuner that picks band width and tile length.
//
// Layout: the series are concatenated back to back (series s starts at the
// prefix sum of the lengths); the profile and index outputs are concatenated
// the same way over the profile lengths P_s = n_s - m_s + 1.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <new>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "batch.hpp"
#include "ctx.hpp"
#include "reseeded_f32.hpp"
#include "scrimp_hip.h"

namespace mp {
int launch_fill_u32(unsigned int* p, int count, unsigned int v, hipStream_t st);
}

#define MP_TRY(expr)                        \
    do {                                    \
        int mp_rc_ = (expr);                \
        if (mp_rc_ != MP_OK) return mp_rc_; \
    } while (0)
#define MP_HIP(expr)                                    \
    do {                                                \
        if ((expr) != hipSuccess) return MP_ERR_DEVICE; \
    } while (0)

namespace {

size_t al(size_t count, size_t elem) { return (count * elem + 255) & ~size_t(255); }

}  // namespace

namespace mp {

int BatchPlan::build(int nseries_, const int64_t* lengths, const int32_t* windows, int excl_opt,
                     int band_k_, int tile_len_) {
    if (nseries_ <= 0) return MP_ERR_BAD_ARG;
    if (band_k_ != 4 && band_k_ != 8 && band_k_ != 16) return MP_ERR_BAD_ARG;
    if (tile_len_ <= 0 || tile_len_ % 16 != 0) return MP_ERR_BAD_ARG;
    nseries = nseries_;
    band_k = band_k_;
    tile_len = tile_len_;
    descs.assign(nseries, SeriesDesc{});
    eprefix.assign(nseries + 1, 0);
    sprefix.assign(nseries + 1, 0);
    tprefix.assign(nseries + 1, 0);
    bprefix.clear();
    total_n = 0;
    cells = 0.0;
    const int BW = 64 * band_k;
    for (int s = 0; s < nseries; ++s) {
        const int64_t n = lengths[s];
        const int m = windows[s];
        if (m < 1) return MP_ERR_BAD_WINDOW;
        if (n < m) return MP_ERR_SHORT_SERIES;
        if (n > (int64_t(1) << 30)) return MP_ERR_TOO_LONG;
        const int excl = excl_opt >= 0 ? excl_opt : m / 4;
        SeriesDesc& d = descs[s];
        d.off = total_n;
        d.out_off = eprefix[s];
        d.n = static_cast<int>(n);
        d.m = m;
        d.P = static_cast<int>(n - m + 1);
        d.base = excl + 1;
        d.tl = std::max(0, d.P - d.base);
        d.nbands = (d.tl + BW - 1) / BW;
        d.band_off = static_cast<int>(bprefix.size());
        d.init_c1 = static_cast<float>(1.0 - 1.0 / static_cast<double>(m));
        int64_t acc = 0;
        bprefix.push_back(0);
        for (int b = 0; b < d.nbands; ++b) {
            const int64_t L = static_cast<int64_t>(d.tl) - static_cast<int64_t>(b) * BW;
            acc += band_tiles(L, band_k, tile_len);
            bprefix.push_back(acc);
        }
        total_n += n;
        eprefix[s + 1] = eprefix[s] + d.P;
        sprefix[s + 1] = sprefix[s] + stat_chunks(d.P);
        tprefix[s + 1] = tprefix[s] + acc;
        cells += 0.5 * static_cast<double>(d.tl) * static_cast<double>(d.tl + 1);
    }
    total_P = eprefix[nseries];
    ntiles = tprefix[nseries];
    if (ntiles >= (int64_t(1) << 31)) return MP_ERR_TOO_LONG;
    return MP_OK;
}

}  // namespace mp

struct mp_batch {
    mp_ctx* ctx = nullptr;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[3] = {};
    mp::BatchPlan plan;
    int refine = 1;
    int accum64 = 1;
    int64_t index_init = -1;
    int excl = -1;
    std::vector<int64_t> lengths;
    std::vector<int32_t> windows;
    // device memory
    char* mem = nullptr;
    size_t bytes = 0;
    mp::F32rBatch dev{};
    // captured graph and the pointers it was captured with
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const double* g_ts = nullptr;
    double* g_prof = nullptr;
    int64_t* g_idx = nullptr;
    int cus = 256;
};

namespace {

int upload_plan(mp_batch* b) {
    const mp::BatchPlan& p = b->plan;
    const int64_t N = p.total_n + 64;
    const size_t need = 2 * al(N, 8) + al(N, mp::term_rec_bytes()) + al(N, 8) +
                        al(p.nseries, sizeof(mp::SeriesDesc)) +
                        3 * al(p.nseries + 1, 8) + al(p.bprefix.size(), 8) + al(1, 4);
    if (need > b->bytes) {
        if (b->mem) (void)hipFree(b->mem);
        b->mem = nullptr;
        b->bytes = 0;
        if (hipMalloc(&b->mem, need) != hipSuccess) return MP_ERR_OOM;
        b->bytes = need;
    }
    char* c = b->mem;
    auto take = [&](size_t bytes) {
        char* r = c;
        c += bytes;
        return r;
    };
    mp::F32rBatch& d = b->dev;
    d.mu = reinterpret_cast<double*>(take(al(N, 8)));
    d.sigma = reinterpret_cast<double*>(take(al(N, 8)));
    d.terms = take(al(N, mp::term_rec_bytes()));
    d.keys = reinterpret_cast<unsigned long long*>(take(al(N, 8)));
    d.accum64 = b->accum64;
    auto* dd = reinterpret_cast<mp::SeriesDesc*>(take(al(p.nseries, sizeof(mp::SeriesDesc))));
    auto* dep = reinterpret_cast<int64_t*>(take(al(p.nseries + 1, 8)));
    auto* dsp = reinterpret_cast<int64_t*>(take(al(p.nseries + 1, 8)));
    auto* dtp = reinterpret_cast<int64_t*>(take(al(p.nseries + 1, 8)));
    auto* dbp = reinterpret_cast<int64_t*>(take(al(p.bprefix.size(), 8)));
    d.counter = reinterpret_cast<unsigned int*>(take(al(1, 4)));
    MP_HIP(hipMemcpyAsync(dd, p.descs.data(), sizeof(mp::SeriesDesc) * p.nseries, hipMemcpyHostToDevice,
                          b->stream));
    MP_HIP(hipMemcpyAsync(dep, p.eprefix.data(), 8 * (p.nseries + 1), hipMemcpyHostToDevice, b->stream));
    MP_HIP(hipMemcpyAsync(dsp, p.sprefix.data(), 8 * (p.nseries + 1), hipMemcpyHostToDevice, b->stream));
    MP_HIP(hipMemcpyAsync(dtp, p.tprefix.data(), 8 * (p.nseries + 1), hipMemcpyHostToDevice, b->stream));
    MP_HIP(hipMemcpyAsync(dbp, p.bprefix.data(), 8 * p.bprefix.size(), hipMemcpyHostToDevice, b->stream));
    MP_HIP(hipStreamSynchronize(b->stream));
    d.descs = dd;
    d.entry_prefix = dep;
    d.stat_prefix = dsp;
    d.nstat = p.sprefix[p.nseries];
    d.tile_prefix = dtp;
    d.band_prefix = dbp;
    d.nseries = p.nseries;
    d.nentries = p.total_P;
    d.ntiles = p.ntiles;
    d.band_k = p.band_k;
    d.tile_len = p.tile_len;
    return MP_OK;
}

// prep + counter reset + tiles + unpack on the batch's stream
int enqueue(mp_batch* b, const double* d_ts, double* d_prof, int64_t* d_idx) {
    mp::F32rBatch d = b->dev;
    d.A = d_ts;
    MP_TRY(mp::launch_f32r_prep(d, b->stream));
    MP_TRY(mp::launch_fill_u32(d.counter, 1, 0u, b->stream));
    if (d.ntiles > 0) {
        const int groups = static_cast<int>(std::min<int64_t>((d.ntiles + 3) / 4, 8 * b->cus));
        MP_TRY(mp::launch_f32r_tiles(d, groups, b->stream));
    }
    MP_TRY(mp::launch_f32r_unpack(d, b->refine, b->index_init, d_prof, d_idx, b->stream));
    return MP_OK;
}

void drop_graph(mp_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// ---- autotuner cache: one line per key "key band_k tile_len cells_per_s" ----
std::mutex g_tune_lock;

std::string tune_cache_path() {
    if (const char* p = std::getenv("SCRIMP_HIP_TUNE_CACHE")) return p;
    const char* home = std::getenv("HOME");
    std::string dir = std::string(home ? home : "/tmp") + "/.cache";
    ::mkdir(dir.c_str(), 0755);
    dir += "/scrimp_hip";
    ::mkdir(dir.c_str(), 0755);
    return dir + "/tune_f32r.txt";
}

std::string tune_key(const mp_batch* b) {
    hipDeviceProp_t prop;
    std::string arch = "unknown";
    if (hipGetDeviceProperties(&prop, b->device) == hipSuccess) arch = prop.gcnArchName;
    std::vector<int32_t> w(b->windows);
    std::sort(w.begin(), w.end());
    const int med_m = w[w.size() / 2];
    const double lg_cells = std::floor(std::log2(std::max(1.0, b->plan.cells)));
    const double lg_n = std::floor(std::log2(std::max<double>(1.0, static_cast<double>(b->plan.total_n) / b->plan.nseries)));
    std::ostringstream s;
    s << arch << (b->accum64 ? "|a64" : "|a32") << "|m" << med_m << "|lc" << lg_cells << "|ln" << lg_n << "|s" << (b->plan.nseries > 1 ? "batch" : "one");
    std::string k = s.str();
    for (auto& ch : k)
        if (ch == ' ') ch = '_';
    return k;
}

bool tune_lookup(const std::string& key, int* bk, int* tl) {
    std::lock_guard<std::mutex> g(g_tune_lock);
    std::ifstream f(tune_cache_path());
    std::string k;
    int a, c;
    double r;
    bool found = false;
    while (f >> k >> a >> c >> r) {
        if (k == key) {
            *bk = a;
            *tl = c;
            found = true;
        }
    }
    return found;
}

void tune_store(const std::string& key, int bk, int tl, double rate) {
    std::lock_guard<std::mutex> g(g_tune_lock);
    std::ofstream f(tune_cache_path(), std::ios::app);
    f << key << " " << bk << " " << tl << " " << rate << "\n";
}

}  // namespace

extern "C" {

int mp_batch_create(mp_ctx* ctx, int32_t nseries, const int64_t* lengths, const int32_t* windows,
                    const mp_opts* o, mp_batch** out) {
    if (!lengths || !windows || !out) return MP_ERR_NULL;
    if (nseries <= 0) return MP_ERR_BAD_ARG;
    mp_batch* b = new (std::nothrow) mp_batch();
    if (!b) return MP_ERR_OOM;
    b->ctx = ctx;
    int dev = 0;
    if (ctx) {
        if (hipSetDevice(ctx->device) != hipSuccess) {
            delete b;
            return MP_ERR_DEVICE;
        }
        dev = ctx->device;
    } else if (hipGetDevice(&dev) != hipSuccess) {
        dev = 0;
    }
    b->device = dev;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
        b->cus = prop.multiProcessorCount;
    auto fail = [&](int rc) {
        mp_batch_destroy(b);
        return rc;
    };
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return fail(MP_ERR_DEVICE);
    for (auto& e : b->ev)
        if (hipEventCreate(&e) != hipSuccess) return fail(MP_ERR_DEVICE);
    b->lengths.assign(lengths, lengths + nseries);
    b->windows.assign(windows, windows + nseries);
    b->refine = o ? (o->refine != 0) : 1;
    b->accum64 = o ? (o->accum64 != 0) : 1;
    b->index_init = (o && o->index_init == 0) ? 0 : -1;
    b->excl = o ? o->exclusion : -1;
    if (o && o->exclusion < -1) return fail(MP_ERR_BAD_WINDOW);
    int bk = (o && o->band_k > 0) ? o->band_k : 8;
    int tl = (o && o->tile_len > 0) ? o->tile_len : 0;
    int rc = MP_OK;
    if (tl > 0) {
        rc = b->plan.build(nseries, lengths, windows, b->excl, bk, tl);
        if (rc != MP_OK) return fail(rc);
    } else {
        // the longest tile (max_auto_tile_len ... 512) that still gives every
        // SIMD four tiles' worth of waves (above kShortTileMax 16 tiles per
        // wave), as for one series
        for (int T = mp::max_auto_tile_len(b->accum64);; T >>= 1) {
            const int64_t target =
                T > mp::kShortTileMax ? mp::long_tile_target(b->cus) : 16LL * std::max(1, b->cus);
            rc = b->plan.build(nseries, lengths, windows, b->excl, bk, T);
            if (rc != MP_OK) return fail(rc);
            tl = T;
            if (b->plan.ntiles >= target || T <= 512) break;
        }
    }
    // a cached tuning result replaces the defaults the caller left open
    if (!(o && o->band_k > 0) || !(o && o->tile_len > 0)) {
        int cbk, ctl;
        if (tune_lookup(tune_key(b), &cbk, &ctl) && ctl <= mp::max_auto_tile_len(b->accum64)) {
            if (!(o && o->band_k > 0)) bk = cbk;
            if (!(o && o->tile_len > 0)) tl = ctl;
            rc = b->plan.build(nseries, lengths, windows, b->excl, bk, tl);
            if (rc != MP_OK) return fail(rc);
        }
    }
    rc = upload_plan(b);
    if (rc != MP_OK) return fail(rc);
    *out = b;
    return MP_OK;
}

int mp_batch_destroy(mp_batch* b) {
    if (!b) return MP_ERR_NULL;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    if (b->mem) (void)hipFree(b->mem);
    for (auto& e : b->ev)
        if (e) (void)hipEventDestroy(e);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return MP_OK;
}

int mp_batch_info(const mp_batch* b, int64_t* total_n, int64_t* total_p, int64_t* ntiles, double* cells,
                  int32_t* band_k, int32_t* tile_len) {
    if (!b) return MP_ERR_NULL;
    if (total_n) *total_n = b->plan.total_n;
    if (total_p) *total_p = b->plan.total_P;
    if (ntiles) *ntiles = b->plan.ntiles;
    if (cells) *cells = b->plan.cells;
    if (band_k) *band_k = b->plan.band_k;
    if (tile_len) *tile_len = b->plan.tile_len;
    return MP_OK;
}

int mp_batch_run(mp_batch* b, const double* d_ts, double* d_profile, int64_t* d_index, int32_t use_graph,
                 mp_timing* tm) {
    if (!b || !d_ts || !d_profile || !d_index) return MP_ERR_NULL;
    MP_HIP(hipEventRecord(b->ev[0], b->stream));
    if (use_graph) {
        if (b->exec && (b->g_ts != d_ts || b->g_prof != d_profile || b->g_idx != d_index)) drop_graph(b);
        if (!b->exec) {
            MP_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            const int rc = enqueue(b, d_ts, d_profile, d_index);
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(b->stream, &g);
            if (rc != MP_OK || ec != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return rc != MP_OK ? rc : MP_ERR_DEVICE;
            }
            b->graph = g;
            if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
                drop_graph(b);
                return MP_ERR_DEVICE;
            }
            b->g_ts = d_ts;
            b->g_prof = d_profile;
            b->g_idx = d_index;
        }
        MP_HIP(hipGraphLaunch(b->exec, b->stream));
    } else {
        MP_TRY(enqueue(b, d_ts, d_profile, d_index));
    }
    MP_HIP(hipEventRecord(b->ev[1], b->stream));
    MP_HIP(hipEventSynchronize(b->ev[1]));
    if (tm) {
        std::memset(tm, 0, sizeof(*tm));
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, b->ev[0], b->ev[1]);
        tm->total_ms = ms;
        tm->kernel_ms = ms;
        tm->cells = b->plan.cells;
        tm->work_items = b->plan.ntiles;
        tm->band_k = b->plan.band_k;
        tm->tile_len = b->plan.tile_len;
    }
    return MP_OK;
}

int mp_batch_tune(mp_batch* b, const double* d_ts, double* d_profile, int64_t* d_index, int32_t max_tile_len,
                  mp_timing* best) {
    if (!b || !d_ts || !d_profile || !d_index) return MP_ERR_NULL;
    // the tuner ranks by speed only among tile lengths whose accuracy the
    // accumulation mode holds (max_auto_tile_len)
    if (max_tile_len <= 0) max_tile_len = mp::max_auto_tile_len(b->accum64);
    max_tile_len = std::min(max_tile_len, mp::max_auto_tile_len(b->accum64));
    const int bks[] = {4, 8};
    const int tls[] = {256, 512, 1024, 2048, 4096, 8192, 16384, 32768};
    double best_rate = -1.0;
    int best_bk = b->plan.band_k, best_tl = b->plan.tile_len;
    for (int bk : bks) {
        for (int tl : tls) {
            if (tl > max_tile_len) continue;
            MP_TRY(b->plan.build(b->plan.nseries, b->lengths.data(), b->windows.data(), b->excl, bk, tl));
            MP_TRY(upload_plan(b));
            drop_graph(b);
            mp_timing t{};
            MP_TRY(mp_batch_run(b, d_ts, d_profile, d_index, 0, &t));  // warm
            MP_TRY(mp_batch_run(b, d_ts, d_profile, d_index, 0, &t));
            const double rate = t.cells / (t.total_ms * 1e-3);
            if (rate > best_rate) {
                best_rate = rate;
                best_bk = bk;
                best_tl = tl;
            }
        }
    }
    MP_TRY(b->plan.build(b->plan.nseries, b->lengths.data(), b->windows.data(), b->excl, best_bk, best_tl));
    MP_TRY(upload_plan(b));
    drop_graph(b);
    tune_store(tune_key(b), best_bk, best_tl, best_rate);
    if (best) {
        std::memset(best, 0, sizeof(*best));
        best->cells = b->plan.cells;
        best->band_k = best_bk;
        best->tile_len = best_tl;
        best->total_ms = b->plan.cells / best_rate * 1e3;
    }
    return MP_OK;
}

}  // extern "C"
