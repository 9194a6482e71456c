// Serving engine: continuous batching of self-join requests on one GPU.
//
// A request is one FP32-reseeded self-join.  Submissions queue on the host;
// an engine thread admits them in arrival order while their device footprint
// (series, statistics, recurrence terms, partial-profile keys, outputs: 64
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

// A tick's descriptor tables, laid out one after another in pinned memory and
// moved to the device with one copy.
struct TableArena {
    char* h = nullptr;  // pinned host side
    char* d = nullptr;  // device side, same layout
    size_t cap = 0, used = 0;
    template <typename T>
    void take(size_t count, T** hp, T** dp) {
        const size_t b = (count * sizeof(T) + 255) & ~size_t(255);
        *hp = reinterpret_cast<T*>(h + used);
        *dp = reinterpret_cast<T*>(d + used);
        used += b;
    }
};

struct Slot {
    hipEvent_t tick_ev = nullptr, fin_ev = nullptr, prep_ev = nullptr;
    bool tick_busy = false, fin_busy = false;
    TableArena cmp;  // admission and tile tables (compute stream)
    TableArena fin;  // finalize tables (finalize stream)
    std::vector<Request*> finishing;  // completes when fin_ev fires
};

}  // namespace

struct mp_server {
    int device = 0;
    int cus = 256;
    // Ticks alternate between two compute streams, so that a tick's tiles start
    // while the previous tick's last tiles drain; each stream has its own work
    // counter.  Set-up (statistics of admitted requests) stays in tick order.
    hipStream_t cstream[2] = {nullptr, nullptr};
    hipStream_t fstream = nullptr;
    hipEvent_t last_prep_ev = nullptr, last_tick_ev = nullptr;
    int band_k = 8, tile_len = 4096, refine = 1, excl = -1, accum64 = 1;
    int64_t index_init = -1;
    int64_t tick_tiles = 1;
    int64_t bytes_per_elem = 60;
    // device pools (element-indexed)
    char* dmem = nullptr;
    double *A = nullptr, *mu = nullptr, *sigma = nullptr, *prof = nullptr;
    void* terms = nullptr;  // one term record per element
    unsigned long long* keys = nullptr;
    int64_t* idx = nullptr;
    int64_t* bandpool = nullptr;
    unsigned int* counter = nullptr;
    RangeAlloc pool, bpool, ppool;
    double* pinned = nullptr;      // pinned staging pool (doubles), host view
    double* pinned_dev = nullptr;  // the same memory as the kernels address it
    Slot slots[kSlots];
    uint64_t tick_no = 0;
    // requests
    std::mutex lock;
    std::condition_variable cv_submit, cv_done;
    std::deque<Request*> waiting;
    std::vector<Request*> active;
    // shared: a waiter keeps its request alive even when poll or completed
    // hands the same request to another thread meanwhile
    std::unordered_map<uint64_t, std::shared_ptr<Request>> all;
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
// pinned staging of a request, in doubles: the series and its band prefix on the
// way in, the profile and index on the way out
int64_t pinned_need(const Request* r) {
    return std::max<int64_t>(r->n + static_cast<int64_t>(r->bprefix.size()), 2 * static_cast<int64_t>(r->desc.P));
}

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
        acc += mp::band_tiles(L, s->band_k, s->tile_len);
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
    hipStream_t cs = s->cstream[s->tick_no & 1];
    unsigned int* counter = s->counter + 64 * (s->tick_no & 1);
    // ---- admission, in arrival order, while memory lasts ----
    std::vector<Request*> admitted;
    while (!s->waiting.empty() && static_cast<int>(admitted.size()) < kMaxRanges) {
        Request* r = s->waiting.front();
        const int64_t need = footprint_elems(r->n);
        const int64_t nb = r->desc.nbands + 1;
        const int64_t pin_need = pinned_need(r);
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
        std::memcpy(s->pinned + poff + r->n, r->bprefix.data(), sizeof(int64_t) * r->bprefix.size());
        std::vector<double>().swap(r->series);
        admitted.push_back(r);
        s->active.push_back(r);
    }
    // ---- tile slices: every active request gets a share of this tick ----
    // The tick budget grows with the backlog (up to 8x the base), so that an
    // overloaded server spends less of its time on per-tick costs; an idle one
    // keeps short ticks for latency.
    std::vector<Request*> with_tiles;
    int64_t backlog = 0;
    for (Request* r : s->active)
        if (r->state == ACTIVE && r->issued < r->ntiles) {
            with_tiles.push_back(r);
            backlog += r->ntiles - r->issued;
        }
    int64_t budget = std::min(8 * s->tick_tiles, std::max(s->tick_tiles, backlog / kSlots));
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
    for (size_t k = 0; k < with_tiles.size(); ++k)
        if (give[k] > 0) ++nranges;
    // ---- this tick's tables, in one arena ----
    sl.cmp.used = 0;
    mp::StageDesc *h_sin = nullptr, *d_sin = nullptr;
    mp::SeriesDesc *h_pdesc = nullptr, *d_pdesc = nullptr, *h_tdesc = nullptr, *d_tdesc = nullptr;
    int64_t *h_ppre = nullptr, *d_ppre = nullptr, *h_tpre = nullptr, *d_tpre = nullptr;
    int64_t *h_spre = nullptr, *d_spre = nullptr;
    sl.cmp.take(admitted.size(), &h_sin, &d_sin);
    sl.cmp.take(admitted.size(), &h_pdesc, &d_pdesc);
    sl.cmp.take(admitted.size() + 1, &h_ppre, &d_ppre);
    sl.cmp.take(admitted.size() + 1, &h_spre, &d_spre);
    sl.cmp.take(static_cast<size_t>(nranges), &h_tdesc, &d_tdesc);
    sl.cmp.take(static_cast<size_t>(nranges) + 1, &h_tpre, &d_tpre);
    int64_t ntiles = 0;
    {
        int q = 0;
        h_tpre[0] = 0;
        for (size_t k = 0; k < with_tiles.size(); ++k) {
            if (give[k] <= 0) continue;
            Request* r = with_tiles[k];
            mp::SeriesDesc d = r->desc;
            d.tile_base = r->issued;
            h_tdesc[q] = d;
            ntiles += give[k];
            h_tpre[q + 1] = ntiles;
            r->issued += give[k];
            ++q;
        }
    }
    int64_t stage_chunks = 0;
    h_ppre[0] = 0;
    h_spre[0] = 0;
    for (size_t k = 0; k < admitted.size(); ++k) {
        Request* r = admitted[k];
        h_pdesc[k] = r->desc;
        h_ppre[k + 1] = h_ppre[k] + r->desc.P;
        h_spre[k + 1] = h_spre[k] + mp::stat_chunks(r->desc.P);
        mp::StageDesc& sd = h_sin[k];
        sd.pin_off = r->pin_off;
        sd.dev_off = r->dev_off;
        sd.band_dev_off = r->band_off;
        sd.chunk_base = stage_chunks;
        sd.n = static_cast<int>(r->n);
        sd.nb = static_cast<int>(r->bprefix.size());
        stage_chunks += (r->n + sd.nb + mp::kStageChunk - 1) / mp::kStageChunk;
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
    // everything the previous tick set up must be in place before this tick's tiles
    if (s->last_prep_ev) H(hipStreamWaitEvent(cs, s->last_prep_ev, 0));
    if (sl.cmp.used > 0) H(hipMemcpyAsync(sl.cmp.d, sl.cmp.h, sl.cmp.used, hipMemcpyHostToDevice, cs));
    if (!admitted.empty()) {
        // one launch copies every admitted series and band prefix from pinned memory
        if (rc == MP_OK)
            rc = mp::launch_stage_in(d_sin, static_cast<int>(admitted.size()), stage_chunks, s->pinned_dev, s->A,
                                     s->bandpool, cs);
        mp::F32rBatch pb{};
        pb.descs = d_pdesc;
        pb.entry_prefix = d_ppre;
        pb.stat_prefix = d_spre;
        pb.nstat = h_spre[admitted.size()];
        pb.nseries = static_cast<int>(admitted.size());
        pb.nentries = h_ppre[admitted.size()];
        pb.A = s->A;
        pb.mu = s->mu;
        pb.sigma = s->sigma;
        pb.terms = s->terms;
        pb.accum64 = s->accum64;
        pb.keys = s->keys;
        if (rc == MP_OK && pb.nentries > 0) rc = mp::launch_f32r_prep(pb, cs);
    }
    H(hipEventRecord(sl.prep_ev, cs));
    if (nranges > 0) {
        if (rc == MP_OK) rc = mp::launch_fill_u32(counter, 1, 0u, cs);
        mp::F32rBatch tb{};
        tb.descs = d_tdesc;
        tb.tile_prefix = d_tpre;
        tb.band_prefix = s->bandpool;
        tb.nseries = nranges;
        tb.ntiles = ntiles;
        tb.band_k = s->band_k;
        tb.tile_len = s->tile_len;
        tb.A = s->A;
        tb.mu = s->mu;
        tb.terms = s->terms;
        tb.accum64 = s->accum64;
        tb.keys = s->keys;
        tb.counter = counter;
        const int groups = static_cast<int>(std::min<int64_t>((ntiles + 3) / 4, 8 * s->cus));
        if (rc == MP_OK) rc = mp::launch_f32r_tiles(tb, groups, cs);
    }
    H(hipEventRecord(sl.tick_ev, cs));
    // ---- finalize stream: unpack + copy out after this tick ----
    if (!finishing.empty()) {
        hipStream_t fs = s->fstream;
        sl.fin.used = 0;
        mp::SeriesDesc *h_fdesc = nullptr, *d_fdesc = nullptr;
        int64_t *h_fpre = nullptr, *d_fpre = nullptr;
        mp::StageDesc *h_sout = nullptr, *d_sout = nullptr;
        sl.fin.take(finishing.size(), &h_fdesc, &d_fdesc);
        sl.fin.take(finishing.size() + 1, &h_fpre, &d_fpre);
        sl.fin.take(finishing.size(), &h_sout, &d_sout);
        h_fpre[0] = 0;
        int64_t out_chunks = 0;
        for (size_t k = 0; k < finishing.size(); ++k) {
            Request* r = finishing[k];
            h_fdesc[k] = r->desc;
            h_fpre[k + 1] = h_fpre[k] + r->desc.P;
            mp::StageDesc& sd = h_sout[k];
            sd.pin_off = r->pin_off;
            sd.dev_off = r->dev_off;
            sd.band_dev_off = 0;
            sd.chunk_base = out_chunks;
            sd.n = r->desc.P;
            sd.nb = 0;
            out_chunks += (r->desc.P + mp::kStageChunk - 1) / mp::kStageChunk;
        }
        H(hipMemcpyAsync(sl.fin.d, sl.fin.h, sl.fin.used, hipMemcpyHostToDevice, fs));
        // a finishing request's tiles ran in this tick and possibly in the previous
        // one, on the other compute stream
        H(hipStreamWaitEvent(fs, sl.tick_ev, 0));
        if (s->last_tick_ev) H(hipStreamWaitEvent(fs, s->last_tick_ev, 0));
        mp::F32rBatch ub{};
        ub.descs = d_fdesc;
        ub.entry_prefix = d_fpre;
        ub.nseries = static_cast<int>(finishing.size());
        ub.nentries = h_fpre[finishing.size()];
        ub.A = s->A;
        ub.mu = s->mu;
        ub.sigma = s->sigma;
        ub.keys = s->keys;
        ub.band_k = s->band_k;
        if (rc == MP_OK && ub.nentries > 0) rc = mp::launch_f32r_unpack(ub, s->refine, s->index_init, s->prof, s->idx, fs);
        // one launch copies every finished profile and index to pinned memory
        if (rc == MP_OK)
            rc = mp::launch_stage_out(d_sout, static_cast<int>(finishing.size()), out_chunks, s->prof, s->idx,
                                      s->pinned_dev, fs);
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
    s->last_prep_ev = sl.prep_ev;
    s->last_tick_ev = sl.tick_ev;
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
    // fail whatever is left; the failed requests are reported by completed() too
    for (auto& kv : s->all) {
        Request* r = kv.second.get();
        if (r->state != DONE) {
            r->state = DONE;
            r->status = s->failed != MP_OK ? s->failed : MP_ERR_DEVICE;
            r->t_done = Clock::now();
            s->newly_done.push_back(r->id);
        }
    }
    s->waiting.clear();
    s->active.clear();
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
    s->refine = o ? (o->refine != 0) : 1;
    s->accum64 = o ? (o->accum64 != 0) : 1;
    s->tile_len = (o && o->tile_len > 0) ? o->tile_len : std::min(4096, mp::max_auto_tile_len(s->accum64));
    if ((s->band_k != 4 && s->band_k != 8 && s->band_k != 16) || s->tile_len % 16 != 0) return fail(MP_ERR_BAD_ARG);
    s->index_init = (o && o->index_init == 0) ? 0 : -1;
    s->excl = o ? o->exclusion : -1;
    const double tile_cells = 64.0 * s->band_k * s->tile_len;
    // 1.5e9 cells (about 0.3 ms of the GPU): measured p99 11 ms at 2,000 requests/s and
    // 36 ms at 80% of the batched throughput, against 19 and 53 ms with 4e9
    const double tc = tick_cells > 0 ? static_cast<double>(tick_cells) : 1.5e9;
    s->tick_tiles = std::max<int64_t>(1, static_cast<int64_t>(tc / tile_cells));
    // device pools: A, mu, sigma, profile, keys, index and a term record per point
    const int64_t per_elem = 8 * 6 + static_cast<int64_t>(mp::term_rec_bytes());
    s->bytes_per_elem = per_elem;
    const int64_t cap = static_cast<int64_t>(hbm_budget_bytes / per_elem);
    if (cap < 1024) return fail(MP_ERR_BAD_ARG);
    const int64_t band_cap = cap / 64 + 1024;
    const size_t bytes = mp::aligned_bytes(cap, 8) * 6 + mp::aligned_bytes(cap, mp::term_rec_bytes()) +
                         mp::aligned_bytes(band_cap, 8) + 512;
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
    s->terms = take(mp::aligned_bytes(cap, mp::term_rec_bytes()));
    s->bandpool = reinterpret_cast<int64_t*>(take(mp::aligned_bytes(band_cap, 8)));
    s->counter = reinterpret_cast<unsigned int*>(take(512));  // two counters, 256 bytes apart
    s->pool.reset(cap);
    s->bpool.reset(band_cap);
    // pinned staging: inputs and outputs of the requests in flight
    const int64_t pin_cap = std::min<int64_t>(2 * cap, int64_t(1) << 27);
    // coherent: the staging kernels read and write it directly over the bus, and a
    // recycled range must never be served from a stale device cache line
    if (hipHostMalloc(reinterpret_cast<void**>(&s->pinned), sizeof(double) * pin_cap,
                      hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
        return fail(MP_ERR_OOM);
    if (hipHostGetDevicePointer(reinterpret_cast<void**>(&s->pinned_dev), s->pinned, 0) != hipSuccess)
        return fail(MP_ERR_DEVICE);
    s->ppool.reset(pin_cap);
    for (auto& st : s->cstream)
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return fail(MP_ERR_DEVICE);
    if (hipStreamCreateWithFlags(&s->fstream, hipStreamNonBlocking) != hipSuccess) return fail(MP_ERR_DEVICE);
    for (auto& sl : s->slots) {
        if (hipEventCreateWithFlags(&sl.tick_ev, hipEventDisableTiming) != hipSuccess) return fail(MP_ERR_DEVICE);
        if (hipEventCreateWithFlags(&sl.prep_ev, hipEventDisableTiming) != hipSuccess) return fail(MP_ERR_DEVICE);
        if (hipEventCreateWithFlags(&sl.fin_ev, hipEventDisableTiming) != hipSuccess) return fail(MP_ERR_DEVICE);
        const size_t db = (sizeof(mp::SeriesDesc) * kMaxRanges + 255) & ~size_t(255);
        const size_t pb = (sizeof(int64_t) * (kMaxRanges + 1) + 255) & ~size_t(255);
        const size_t sb = (sizeof(mp::StageDesc) * kMaxRanges + 255) & ~size_t(255);
        sl.cmp.cap = sb + 2 * db + 3 * pb;
        sl.fin.cap = db + pb + sb;
        void* h = nullptr;
        void* d = nullptr;
        if (hipHostMalloc(&h, sl.cmp.cap + sl.fin.cap) != hipSuccess) return fail(MP_ERR_OOM);
        if (hipMalloc(&d, sl.cmp.cap + sl.fin.cap) != hipSuccess) {
            (void)hipHostFree(h);
            return fail(MP_ERR_OOM);
        }
        sl.cmp.h = static_cast<char*>(h);
        sl.cmp.d = static_cast<char*>(d);
        sl.fin.h = sl.cmp.h + sl.cmp.cap;
        sl.fin.d = sl.cmp.d + sl.cmp.cap;
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
    auto r = std::make_shared<Request>();
    r->n = n;
    r->m = m;
    r->out_prof = profile_out;
    r->out_idx = index_out;
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> g(s->lock);
    if (s->failed != MP_OK) return s->failed;
    build_request_plan(s, r.get());
    if (footprint_elems(n) > s->pool.capacity() ||
        pinned_need(r.get()) > s->ppool.capacity() ||
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
    // hold the request itself: another thread's poll or completed may take the
    // id out of the table while this one sleeps
    const std::shared_ptr<Request> r = it->second;
    s->cv_done.wait(lk, [&] { return r->state == DONE; });
    const int st = r->status;
    if (latency_ms) *latency_ms = ms_since(r->t_submit, r->t_done);
    auto again = s->all.find(req_id);
    if (again != s->all.end() && again->second == r) s->all.erase(again);
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
    for (auto st : s->cstream)
        if (st) (void)hipStreamSynchronize(st);
    if (s->fstream) (void)hipStreamSynchronize(s->fstream);
    for (auto& sl : s->slots) {
        if (sl.tick_ev) (void)hipEventDestroy(sl.tick_ev);
        if (sl.prep_ev) (void)hipEventDestroy(sl.prep_ev);
        if (sl.fin_ev) (void)hipEventDestroy(sl.fin_ev);
        if (sl.cmp.h) (void)hipHostFree(sl.cmp.h);
        if (sl.cmp.d) (void)hipFree(sl.cmp.d);
    }
    if (s->pinned) (void)hipHostFree(s->pinned);
    if (s->dmem) (void)hipFree(s->dmem);
    for (auto st : s->cstream)
        if (st) (void)hipStreamDestroy(st);
    if (s->fstream) (void)hipStreamDestroy(s->fstream);
    delete s;
    return MP_OK;
}

}  // extern "C"
