// Serving engine: continuous batching of self-join requests on one GPU.
//
// A request is one FP32-reseeded self-join.  Submissions queue on the host;
// an engine thread admits them in arrival order while their device footprint
// (series, statistics, recurrence terms, partial-profile keys, outputs: 60
// bytes per point) fits the HBM budget, keeps that state resident until the
// request completes, and runs "ticks": each tick is one launch of the
// persistent tile kernel over a slice of the remaining tiles of every active
// request, so requests that arrive while others run share the next launch.
// When a request's last tile has been launched, a second stream waits for that
// tick, unpacks the request's keys into profile/index and copies them out;
// the request then completes and its memory returns to the pool.
// Results do not depend on the mix: every tile is computed the same way
// whichever tick carries it, and tiles merge by an order-independent max.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ctx.hpp"
#include "reseeded_f32.hpp"
#include "scrimp_hip.h"

namespace mp {
int launch_fill_u32(unsigned int* p, int count, unsigned int v, hipStream_t st);
}

namespace {

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point t0, Clock::time_point t) {
    return std::chrono::duration<double, std::milli>(t - t0).count();
}

// First-fit allocator over [0, capacity) with coalescing of freed ranges.
class RangeAlloc {
public:
    void reset(int64_t cap) {
        free_.clear();
        if (cap > 0) free_[0] = cap;
        cap_ = cap;
        used_ = 0;
    }
    int64_t alloc(int64_t len) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= len) {
                const int64_t off = it->first, rem = it->second - len;
                free_.erase(it);
                if (rem > 0) free_[off + len] = rem;
                used_ += len;
                return off;
            }
        }
        return -1;
    }
    void release(int64_t off, int64_t len) {
        if (off < 0 || len <= 0) return;
        used_ -= len;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + len == next->first) {
            len += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += len;
                return;
            }
        }
        free_[off] = len;
    }
    int64_t capacity() const { return cap_; }
    int64_t used() const { return used_; }

private:
    std::map<int64_t, int64_t> free_;
    int64_t cap_ = 0, used_ = 0;
};

enum ReqState { WAITING = 0, ACTIVE = 1, FINALIZING = 2, DONE = 3 };

struct Request {
    uint64_t id = 0;
    int64_t n = 0;
    int m = 0;
    std::vector<double> series;
    double* out_prof = nullptr;
    int64_t* out_idx = nullptr;
    mp::SeriesDesc desc{};
    std::vector<int64_t> bprefix;
    int64_t ntiles = 0, issued = 0;
    int64_t dev_off = -1, dev_len = 0;
    int64_t band_off = -1, band_len = 0;
    int64_t pin_off = -1, pin_len = 0;  // pinned staging (in doubles): series in, then outputs
    Clock::time_point t_submit, t_admit, t_done;
    int state = WAITING;
    int status = MP_OK;
    double cells = 0.0;
};

constexpr int kSlots = 3;
constexpr int kMaxRanges = 8192;

struct Slot {
    hipEvent_t tick_ev = nullptr, fin_ev = nullptr;
    bool tick_busy = false, fin_busy = false;
    // pinned host tables and their device copies
    mp::SeriesDesc *h_tdesc = nullptr, *d_tdesc = nullptr;
    int64_t *h_tpre = nullptr, *d_tpre = nullptr;
    mp::SeriesDesc *h_pdesc = nullptr, *d_pdesc = nullptr;
    int64_t *h_ppre = nullptr, *d_ppre = nullptr;
    mp::SeriesDesc *h_fdesc = nullptr, *d_fdesc = nullptr;
    int64_t *h_fpre = nullptr, *d_fpre = nullptr;
    std::vector<Request*> finishing;  // completes when fin_ev fires
};

}  // namespace

struct mp_server {
    int device = 0;
    int cus = 256;
    hipStream_t cstream = nullptr, fstream = nullptr;
    int band_k = 8, tile_len = 1024, refine = 1, excl = -1, accum64 = 1;
    int64_t index_init = -1;
    int64_t tick_tiles = 1;
    int64_t bytes_per_elem = 60;
    // device pools (element-indexed)
    char* dmem = nullptr;
    double *A = nullptr, *mu = nullptr, *sigma = nullptr, *prof = nullptr;
    float *df = nullptr, *dg = nullptr, *inv = nullptr;
    double *df64 = nullptr, *dg64 = nullptr;
    unsigned long long* keys = nullptr;
    int64_t* idx = nullptr;
    int64_t* bandpool = nullptr;
    unsigned int* counter = nullptr;
    RangeAlloc pool, bpool, ppool;
    double* pinned = nullptr;  // pinned staging pool (doubles)
    Slot slots[kSlots];
    uint64_t tick_no = 0;
    // requests
    std::mutex lock;
    std::condition_variable cv_submit, cv_done;
    std::deque<Request*> waiting;
    std::vector<Request*> active;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> all;
    std::vector<uint64_t> newly_done;  // completed ids not yet drained
    uint64_t next_id = 1;
    bool stop = false;
    int failed = MP_OK;
    std::thread engine;
    // statistics
    std::atomic<int64_t> ticks{0}, completed{0};
    double cells_done = 0.0;
};

namespace {

int64_t footprint_elems(int64_t n) { return n + 64; }

int build_request_plan(mp_server* s, Request* r) {
    const int m = r->m;
    const int64_t n = r->n;
    const int excl = s->excl >= 0 ? s->excl : m / 4;
    mp::SeriesDesc& d = r->desc;
    d = mp::SeriesDesc{};
    d.n = static_cast<int>(n);
    d.m = m;
    d.P = static_cast<int>(n - m + 1);
    d.base = excl + 1;
    d.tl = std::max(0, d.P - d.base);
    const int BW = 64 * s->band_k;
    d.nbands = (d.tl + BW - 1) / BW;
    d.init_c1 = static_cast<float>(1.0 - 1.0 / static_cast<double>(m));
    r->bprefix.assign(1, 0);
    int64_t acc = 0;
    for (int b = 0; b < d.nbands; ++b) {
        const int64_t L = static_cast<int64_t>(d.tl) - static_cast<int64_t>(b) * BW;
        acc += (L + s->tile_len - 1) / s->tile_len;
        r->bprefix.push_back(acc);
    }
    r->ntiles = acc;
    r->cells = 0.5 * static_cast<double>(d.tl) * static_cast<double>(d.tl + 1);
    return MP_OK;
}

void complete(mp_server* s, Request* r, int status) {
    r->status = status;
    r->state = DONE;
    r->t_done = Clock::now();
    s->pool.release(r->dev_off, r->dev_len);
    s->bpool.release(r->band_off, r->band_len);
    s->ppool.release(r->pin_off, r->pin_len);
    r->dev_off = r->band_off = r->pin_off = -1;
    if (status == MP_OK) s->cells_done += r->cells;
    s->completed++;
    s->newly_done.push_back(r->id);
}

// Retire finished finalize batches: copy results from pinned staging to the
// caller's buffers and release the request's memory.  Called with the lock held.
void retire(mp_server* s, bool block) {
    for (auto& sl : s->slots) {
        if (sl.tick_busy && (block ? hipEventSynchronize(sl.tick_ev) == hipSuccess
                                   : hipEventQuery(sl.tick_ev) == hipSuccess)) {
            sl.tick_busy = false;
        }
        if (sl.fin_busy && (block ? hipEventSynchronize(sl.fin_ev) == hipSuccess
                                  : hipEventQuery(sl.fin_ev) == hipSuccess)) {
            for (Request* r : sl.finishing) {
                const int64_t P = r->desc.P;
                const double* ph = s->pinned + r->pin_off;
                std::memcpy(r->out_prof, ph, sizeof(double) * P);
                std::memcpy(r->out_idx, ph + P, sizeof(int64_t) * P);
                complete(s, r, MP_OK);
            }
            sl.finishing.clear();
            sl.fin_busy = false;
            s->cv_done.notify_all();
        }
    }
}

int run_tick(mp_server* s, std::unique_lock<std::mutex>& lk) {
    Slot& sl = s->slots[s->tick_no % kSlots];
    // the slot's tables may still be read by its previous tick / finalize
    if (sl.tick_busy || sl.fin_busy) {
        lk.unlock();
        if (sl.tick_busy && hipEventSynchronize(sl.tick_ev) != hipSuccess) return MP_ERR_DEVICE;
        if (sl.fin_busy && hipEventSynchronize(sl.fin_ev) != hipSuccess) return MP_ERR_DEVICE;
        lk.lock();
        retire(s, false);
    }
    hipStream_t cs = s->cstream;
    // ---- admission, in arrival order, while memory lasts ----
    std::vector<Request*> admitted;
    while (!s->waiting.empty() && static_cast<int>(admitted.size()) < kMaxRanges) {
        Request* r = s->waiting.front();
        const int64_t need = footprint_elems(r->n);
        const int64_t nb = r->desc.nbands + 1;
        const int64_t pin_need = std::max<int64_t>(r->n, 2 * static_cast<int64_t>(r->desc.P));
        const int64_t off = s->pool.alloc(need);
        if (off < 0) break;
        const int64_t boff = s->bpool.alloc(nb);
        if (boff < 0) {
            s->pool.release(off, need);
            break;
        }
        const int64_t poff = s->ppool.alloc(pin_need);
        if (poff < 0) {
            s->pool.release(off, need);
            s->bpool.release(boff, nb);
            break;
        }
        s->waiting.pop_front();
        r->dev_off = off;
        r->dev_len = need;
        r->band_off = boff;
        r->band_len = nb;
        r->pin_off = poff;
        r->pin_len = pin_need;
        r->desc.off = off;
        r->desc.out_off = off;
        r->desc.band_off = static_cast<int>(boff);
        r->state = ACTIVE;
        r->t_admit = Clock::now();
        std::memcpy(s->pinned + poff, r->series.data(), sizeof(double) * r->n);
        std::vector<double>().swap(r->series);
        admitted.push_back(r);
        s->active.push_back(r);
    }
    // ---- tile slices: every active request gets a share of this tick ----
    std::vector<std::pair<Request*, int64_t>> take;
    int64_t budget = s->tick_tiles;
    std::vector<Request*> with_tiles;
    for (Request* r : s->active)
        if (r->state == ACTIVE && r->issued < r->ntiles) with_tiles.push_back(r);
    if (static_cast<int>(with_tiles.size()) > kMaxRanges) with_tiles.resize(kMaxRanges);
    std::vector<int64_t> give(with_tiles.size(), 0);
    for (int pass = 0; pass < 2 && budget > 0; ++pass) {
        const int64_t left = static_cast<int64_t>(with_tiles.size());
        const int64_t share = std::max<int64_t>(1, budget / std::max<int64_t>(1, left));
        for (size_t k = 0; k < with_tiles.size() && budget > 0; ++k) {
            Request* r = with_tiles[k];
            const int64_t rem = r->ntiles - r->issued - give[k];
            const int64_t t = std::min({rem, pass == 0 ? share : budget, budget});
            if (t > 0) {
                give[k] += t;
                budget -= t;
            }
        }
    }
    int nranges = 0;
    int64_t ntiles = 0;
    s->slots[s->tick_no % kSlots].h_tpre[0] = 0;
    for (size_t k = 0; k < with_tiles.size(); ++k) {
        if (give[k] <= 0) continue;
        Request* r = with_tiles[k];
        mp::SeriesDesc d = r->desc;
        d.tile_base = r->issued;
        sl.h_tdesc[nranges] = d;
        ntiles += give[k];
        sl.h_tpre[nranges + 1] = ntiles;
        r->issued += give[k];
        ++nranges;
    }
    // requests whose every tile is now launched (or that have none) finish after this tick
    std::vector<Request*> finishing;
    for (Request* r : s->active)
        if (r->state == ACTIVE && r->issued >= r->ntiles && static_cast<int>(finishing.size()) < kMaxRanges)
            finishing.push_back(r);
    if (admitted.empty() && nranges == 0 && finishing.empty()) return 1;  // idle
    lk.unlock();
    // ---- enqueue on the compute stream ----
    int rc = MP_OK;
    auto H = [&](hipError_t e) {
        if (e != hipSuccess && rc == MP_OK) rc = MP_ERR_DEVICE;
    };
    if (!admitted.empty()) {
        sl.h_ppre[0] = 0;
        for (size_t k = 0; k < admitted.size(); ++k) {
            Request* r = admitted[k];
            sl.h_pdesc[k] = r->desc;
            sl.h_ppre[k + 1] = sl.h_ppre[k] + r->desc.P;
            H(hipMemcpyAsync(s->A + r->dev_off, s->pinned + r->pin_off, sizeof(double) * r->n,
                             hipMemcpyHostToDevice, cs));
            H(hipMemcpyAsync(s->bandpool + r->band_off, r->bprefix.data(), sizeof(int64_t) * r->bprefix.size(),
                             hipMemcpyHostToDevice, cs));
        }
        H(hipMemcpyAsync(sl.d_pdesc, sl.h_pdesc, sizeof(mp::SeriesDesc) * admitted.size(), hipMemcpyHostToDevice, cs));
        H(hipMemcpyAsync(sl.d_ppre, sl.h_ppre, sizeof(int64_t) * (admitted.size() + 1), hipMemcpyHostToDevice, cs));
        mp::F32rBatch pb{};
        pb.descs = sl.d_pdesc;
        pb.entry_prefix = sl.d_ppre;
        pb.nseries = static_cast<int>(admitted.size());
        pb.nentries = sl.h_ppre[admitted.size()];
        pb.A = s->A;
        pb.mu = s->mu;
        pb.sigma = s->sigma;
        pb.df = s->df;
        pb.dg = s->dg;
        pb.inv = s->inv;
        pb.df64 = s->df64;
        pb.dg64 = s->dg64;
        pb.accum64 = s->accum64;
        pb.keys = s->keys;
        if (rc == MP_OK && pb.nentries > 0) rc = mp::launch_f32r_prep(pb, cs);
    }
    if (nranges > 0) {
        H(hipMemcpyAsync(sl.d_tdesc, sl.h_tdesc, sizeof(mp::SeriesDesc) * nranges, hipMemcpyHostToDevice, cs));
        H(hipMemcpyAsync(sl.d_tpre, sl.h_tpre, sizeof(int64_t) * (nranges + 1), hipMemcpyHostToDevice, cs));
        if (rc == MP_OK) rc = mp::launch_fill_u32(s->counter, 1, 0u, cs);
        mp::F32rBatch tb{};
        tb.descs = sl.d_tdesc;
        tb.tile_prefix = sl.d_tpre;
        tb.band_prefix = s->bandpool;
        tb.nseries = nranges;
        tb.ntiles = ntiles;
        tb.band_k = s->band_k;
        tb.tile_len = s->tile_len;
        tb.A = s->A;
        tb.mu = s->mu;
        tb.df = s->df;
        tb.dg = s->dg;
        tb.inv = s->inv;
        tb.df64 = s->df64;
        tb.dg64 = s->dg64;
        tb.accum64 = s->accum64;
        tb.keys = s->keys;
        tb.counter = s->counter;
        const int groups = static_cast<int>(std::min<int64_t>((ntiles + 3) / 4, 8 * s->cus));
        if (rc == MP_OK) rc = mp::launch_f32r_tiles(tb, groups, cs);
    }
    H(hipEventRecord(sl.tick_ev, cs));
    // ---- finalize stream: unpack + copy out after this tick ----
    if (!finishing.empty()) {
        hipStream_t fs = s->fstream;
        sl.h_fpre[0] = 0;
        for (size_t k = 0; k < finishing.size(); ++k) {
            sl.h_fdesc[k] = finishing[k]->desc;
            sl.h_fpre[k + 1] = sl.h_fpre[k] + finishing[k]->desc.P;
        }
        H(hipMemcpyAsync(sl.d_fdesc, sl.h_fdesc, sizeof(mp::SeriesDesc) * finishing.size(), hipMemcpyHostToDevice, fs));
        H(hipMemcpyAsync(sl.d_fpre, sl.h_fpre, sizeof(int64_t) * (finishing.size() + 1), hipMemcpyHostToDevice, fs));
        H(hipStreamWaitEvent(fs, sl.tick_ev, 0));
        mp::F32rBatch ub{};
        ub.descs = sl.d_fdesc;
        ub.entry_prefix = sl.d_fpre;
        ub.nseries = static_cast<int>(finishing.size());
        ub.nentries = sl.h_fpre[finishing.size()];
        ub.A = s->A;
        ub.mu = s->mu;
        ub.sigma = s->sigma;
        ub.keys = s->keys;
        if (rc == MP_OK && ub.nentries > 0) rc = mp::launch_f32r_unpack(ub, s->refine, s->index_init, s->prof, s->idx, fs);
        for (Request* r : finishing) {
            const int64_t P = r->desc.P;
            H(hipMemcpyAsync(s->pinned + r->pin_off, s->prof + r->dev_off, sizeof(double) * P, hipMemcpyDeviceToHost, fs));
            H(hipMemcpyAsync(s->pinned + r->pin_off + P, s->idx + r->dev_off, sizeof(int64_t) * P,
                             hipMemcpyDeviceToHost, fs));
        }
        H(hipEventRecord(sl.fin_ev, fs));
    }
    lk.lock();
    sl.tick_busy = true;
    if (!finishing.empty()) {
        sl.finishing = finishing;
        sl.fin_busy = true;
        for (Request* r : finishing) r->state = FINALIZING;
        s->active.erase(std::remove_if(s->active.begin(), s->active.end(),
                                       [](Request* r) { return r->state != ACTIVE; }),
                        s->active.end());
    }
    s->tick_no++;
    s->ticks++;
    return rc;
}

void engine_main(mp_server* s) {
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->lock);
    for (;;) {
        retire(s, false);
        if (s->failed != MP_OK) break;
        const bool pending = !s->waiting.empty() || !s->active.empty();
        bool in_flight = false;
        for (auto& sl : s->slots) in_flight = in_flight || sl.tick_busy || sl.fin_busy;
        if (s->stop && !pending && !in_flight) break;
        if (!pending) {
            if (in_flight) {
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(50));
                lk.lock();
            } else {
                s->cv_submit.wait_for(lk, std::chrono::milliseconds(2));
            }
            continue;
        }
        const int rc = run_tick(s, lk);
        if (rc == 1) {  // nothing launchable now (waiting for memory or finalizes)
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(50));
            lk.lock();
        } else if (rc != MP_OK) {
            s->failed = rc;
        }
    }
    // fail whatever is left
    for (auto& kv : s->all) {
        Request* r = kv.second.get();
        if (r->state != DONE) {
            r->state = DONE;
            r->status = s->failed != MP_OK ? s->failed : MP_ERR_DEVICE;
            r->t_done = Clock::now();
        }
    }
    s->cv_done.notify_all();
}

}  // namespace

extern "C" {

int mp_server_create(mp_ctx* ctx, uint64_t hbm_budget_bytes, int64_t tick_cells, const mp_opts* o,
                     mp_server** out) {
    if (!out) return MP_ERR_NULL;
    mp_server* s = new (std::nothrow) mp_server();
    if (!s) return MP_ERR_OOM;
    auto fail = [&](int rc) {
        mp_server_destroy(s);
        return rc;
    };
    int dev = ctx ? ctx->device : 0;
    if (hipSetDevice(dev) != hipSuccess) {
        delete s;
        return MP_ERR_NO_DEVICE;
    }
    s->device = dev;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0) s->cus = prop.multiProcessorCount;
    s->band_k = (o && o->band_k > 0) ? o->band_k : 8;
    s->tile_len = (o && o->tile_len > 0) ? o->tile_len : 1024;
    if ((s->band_k != 4 && s->band_k != 8 && s->band_k != 16) || s->tile_len % 16 != 0) return fail(MP_ERR_BAD_ARG);
    s->refine = o ? (o->refine != 0) : 1;
    s->accum64 = o ? (o->accum64 != 0) : 1;
    s->index_init = (o && o->index_init == 0) ? 0 : -1;
    s->excl = o ? o->exclusion : -1;
    const double tile_cells = 64.0 * s->band_k * s->tile_len;
    const double tc = tick_cells > 0 ? static_cast<double>(tick_cells) : 4e9;
    s->tick_tiles = std::max<int64_t>(1, static_cast<int64_t>(tc / tile_cells));
    // device pools: 60 bytes per point
    const int64_t per_elem = 8 * 3 + 4 * 3 + 8 + 8 + 8 + (s->accum64 ? 16 : 0);
    s->bytes_per_elem = per_elem;
    const int64_t cap = static_cast<int64_t>(hbm_budget_bytes / per_elem);
    if (cap < 1024) return fail(MP_ERR_BAD_ARG);
    const int64_t band_cap = cap / 64 + 1024;
    const size_t bytes = mp::aligned_bytes(cap, 8) * (s->accum64 ? 8 : 6) + mp::aligned_bytes(cap, 4) * 3 +
                         mp::aligned_bytes(band_cap, 8) + 256;
    if (hipMalloc(&s->dmem, bytes) != hipSuccess) return fail(MP_ERR_OOM);
    char* p = s->dmem;
    auto take = [&](size_t b) {
        char* r = p;
        p += b;
        return r;
    };
    s->A = reinterpret_cast<double*>(take(mp::aligned_bytes(cap, 8)));
    s->mu = reinterpret_cast<double*>(take(mp::aligned_bytes(cap, 8)));
    s->sigma = reinterpret_cast<double*>(take(mp::aligned_bytes(cap, 8)));
    s->prof = reinterpret_cast<double*>(take(mp::aligned_bytes(cap, 8)));
    s->keys = reinterpret_cast<unsigned long long*>(take(mp::aligned_bytes(cap, 8)));
    s->idx = reinterpret_cast<int64_t*>(take(mp::aligned_bytes(cap, 8)));
    s->df = reinterpret_cast<float*>(take(mp::aligned_bytes(cap, 4)));
    s->dg = reinterpret_cast<float*>(take(mp::aligned_bytes(cap, 4)));
    s->inv = reinterpret_cast<float*>(take(mp::aligned_bytes(cap, 4)));
    if (s->accum64) {
        s->df64 = reinterpret_cast<double*>(take(mp::aligned_bytes(cap, 8)));
        s->dg64 = reinterpret_cast<double*>(take(mp::aligned_bytes(cap, 8)));
    }
    s->bandpool = reinterpret_cast<int64_t*>(take(mp::aligned_bytes(band_cap, 8)));
    s->counter = reinterpret_cast<unsigned int*>(take(256));
    s->pool.reset(cap);
    s->bpool.reset(band_cap);
    // pinned staging: inputs and outputs of the requests in flight
    const int64_t pin_cap = std::min<int64_t>(2 * cap, int64_t(1) << 27);
    if (hipHostMalloc(reinterpret_cast<void**>(&s->pinned), sizeof(double) * pin_cap) != hipSuccess)
        return fail(MP_ERR_OOM);
    s->ppool.reset(pin_cap);
    if (hipStreamCreateWithFlags(&s->cstream, hipStreamNonBlocking) != hipSuccess) return fail(MP_ERR_DEVICE);
    if (hipStreamCreateWithFlags(&s->fstream, hipStreamNonBlocking) != hipSuccess) return fail(MP_ERR_DEVICE);
    for (auto& sl : s->slots) {
        if (hipEventCreateWithFlags(&sl.tick_ev, hipEventDisableTiming) != hipSuccess) return fail(MP_ERR_DEVICE);
        if (hipEventCreateWithFlags(&sl.fin_ev, hipEventDisableTiming) != hipSuccess) return fail(MP_ERR_DEVICE);
        const size_t db = sizeof(mp::SeriesDesc) * kMaxRanges, pb = sizeof(int64_t) * (kMaxRanges + 1);
        void* h = nullptr;
        void* d = nullptr;
        if (hipHostMalloc(&h, 3 * (db + pb)) != hipSuccess) return fail(MP_ERR_OOM);
        if (hipMalloc(&d, 3 * (db + pb)) != hipSuccess) {
            (void)hipHostFree(h);
            return fail(MP_ERR_OOM);
        }
        char* hc = static_cast<char*>(h);
        char* dc = static_cast<char*>(d);
        sl.h_tdesc = reinterpret_cast<mp::SeriesDesc*>(hc);
        sl.h_pdesc = reinterpret_cast<mp::SeriesDesc*>(hc + db);
        sl.h_fdesc = reinterpret_cast<mp::SeriesDesc*>(hc + 2 * db);
        sl.h_tpre = reinterpret_cast<int64_t*>(hc + 3 * db);
        sl.h_ppre = reinterpret_cast<int64_t*>(hc + 3 * db + pb);
        sl.h_fpre = reinterpret_cast<int64_t*>(hc + 3 * db + 2 * pb);
        sl.d_tdesc = reinterpret_cast<mp::SeriesDesc*>(dc);
        sl.d_pdesc = reinterpret_cast<mp::SeriesDesc*>(dc + db);
        sl.d_fdesc = reinterpret_cast<mp::SeriesDesc*>(dc + 2 * db);
        sl.d_tpre = reinterpret_cast<int64_t*>(dc + 3 * db);
        sl.d_ppre = reinterpret_cast<int64_t*>(dc + 3 * db + pb);
        sl.d_fpre = reinterpret_cast<int64_t*>(dc + 3 * db + 2 * pb);
    }
    s->engine = std::thread(engine_main, s);
    *out = s;
    return MP_OK;
}

int mp_server_submit(mp_server* s, const double* ts, int64_t n, int32_t m, double* profile_out,
                     int64_t* index_out, uint64_t* req_id) {
    if (!s || !ts || !profile_out || !index_out || !req_id) return MP_ERR_NULL;
    if (m < 1) return MP_ERR_BAD_WINDOW;
    if (n < m) return MP_ERR_SHORT_SERIES;
    if (n > (int64_t(1) << 30)) return MP_ERR_TOO_LONG;
    auto r = std::make_unique<Request>();
    r->n = n;
    r->m = m;
    r->out_prof = profile_out;
    r->out_idx = index_out;
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> g(s->lock);
    if (s->failed != MP_OK) return s->failed;
    build_request_plan(s, r.get());
    if (footprint_elems(n) > s->pool.capacity() ||
        std::max<int64_t>(n, 2 * static_cast<int64_t>(r->desc.P)) > s->ppool.capacity() ||
        r->desc.nbands + 1 > s->bpool.capacity())
        return MP_ERR_ADMISSION;
    r->series.assign(ts, ts + n);
    r->id = s->next_id++;
    *req_id = r->id;
    s->waiting.push_back(r.get());
    s->all.emplace(r->id, std::move(r));
    s->cv_submit.notify_one();
    return MP_OK;
}

// done = 1 once the request's results are in its output buffers; latency =
// submit -> done in ms; the request is forgotten after a poll that reports done.
int mp_server_poll(mp_server* s, uint64_t req_id, int32_t* done, double* latency_ms, double* queue_ms) {
    if (!s || !done) return MP_ERR_NULL;
    std::lock_guard<std::mutex> g(s->lock);
    auto it = s->all.find(req_id);
    if (it == s->all.end()) return MP_ERR_NOT_FOUND;
    Request* r = it->second.get();
    *done = r->state == DONE;
    if (!*done) return MP_OK;
    const int st = r->status;
    if (latency_ms) *latency_ms = ms_since(r->t_submit, r->t_done);
    if (queue_ms) *queue_ms = r->t_admit.time_since_epoch().count() ? ms_since(r->t_submit, r->t_admit) : 0.0;
    s->all.erase(it);
    return st;
}

int mp_server_wait(mp_server* s, uint64_t req_id, double* latency_ms) {
    if (!s) return MP_ERR_NULL;
    std::unique_lock<std::mutex> lk(s->lock);
    auto it = s->all.find(req_id);
    if (it == s->all.end()) return MP_ERR_NOT_FOUND;
    Request* r = it->second.get();
    s->cv_done.wait(lk, [&] { return r->state == DONE; });
    const int st = r->status;
    if (latency_ms) *latency_ms = ms_since(r->t_submit, r->t_done);
    s->all.erase(it);
    return st;
}

// Drains up to max completed requests (in completion order): their ids,
// statuses, latencies (submit -> done) and queueing delays (submit -> admit).
// Drained requests are forgotten.
int mp_server_completed(mp_server* s, int32_t max, uint64_t* ids, int32_t* statuses, double* latency_ms,
                        double* queue_ms, int32_t* count) {
    if (!s || !ids || !count) return MP_ERR_NULL;
    std::lock_guard<std::mutex> g(s->lock);
    int32_t k = 0;
    size_t used = 0;
    for (; used < s->newly_done.size() && k < max; ++used) {
        auto it = s->all.find(s->newly_done[used]);
        if (it == s->all.end()) continue;  // already taken by poll/wait
        Request* r = it->second.get();
        ids[k] = r->id;
        if (statuses) statuses[k] = r->status;
        if (latency_ms) latency_ms[k] = ms_since(r->t_submit, r->t_done);
        if (queue_ms) queue_ms[k] = ms_since(r->t_submit, r->t_admit);
        s->all.erase(it);
        ++k;
    }
    s->newly_done.erase(s->newly_done.begin(), s->newly_done.begin() + used);
    *count = k;
    return MP_OK;
}

int mp_server_stats(mp_server* s, int64_t* ticks, int64_t* completed, double* cells_done, int64_t* queued,
                    int64_t* active, int64_t* hbm_used_bytes) {
    if (!s) return MP_ERR_NULL;
    std::lock_guard<std::mutex> g(s->lock);
    if (ticks) *ticks = s->ticks.load();
    if (completed) *completed = s->completed.load();
    if (cells_done) *cells_done = s->cells_done;
    if (queued) *queued = static_cast<int64_t>(s->waiting.size());
    if (active) *active = static_cast<int64_t>(s->active.size());
    if (hbm_used_bytes) *hbm_used_bytes = s->pool.used() * s->bytes_per_elem;
    return s->failed;
}

int mp_server_destroy(mp_server* s) {
    if (!s) return MP_ERR_NULL;
    if (s->engine.joinable()) {
        {
            std::lock_guard<std::mutex> g(s->lock);
            s->stop = true;
        }
        s->cv_submit.notify_all();
        s->engine.join();
    }
    (void)hipSetDevice(s->device);
    if (s->cstream) (void)hipStreamSynchronize(s->cstream);
    if (s->fstream) (void)hipStreamSynchronize(s->fstream);
    for (auto& sl : s->slots) {
        if (sl.tick_ev) (void)hipEventDestroy(sl.tick_ev);
        if (sl.fin_ev) (void)hipEventDestroy(sl.fin_ev);
        if (sl.h_tdesc) (void)hipHostFree(sl.h_tdesc);
        if (sl.d_tdesc) (void)hipFree(sl.d_tdesc);
    }
    if (s->pinned) (void)hipHostFree(s->pinned);
    if (s->dmem) (void)hipFree(s->dmem);
    if (s->cstream) (void)hipStreamDestroy(s->cstream);
    if (s->fstream) (void)hipStreamDestroy(s->fstream);
    delete s;
    return MP_OK;
}

}  // extern "C"
