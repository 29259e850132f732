// hipgrid_server.cpp -- request loop over one GPU (hg_server_* in include/hipgrid.h).
//
// Requests (a graph plus "decompose" or "densest") are copied into the
// server at submit and queue in arrival order.  `lanes` worker threads, each
// with its own HIP stream, take work whenever they are idle.  A request with
// more than `large_entries` directed entries runs alone on the single-graph
// engine (hg_graph_*, the whole GPU per level).  Small requests are batched
// only when the queue holds at least `min_batch` of them at its head: then
// the lane takes them, up to `max_batch` and as many as the device-memory
// budget admits, and runs them as one batch (hg_batch_*: one persistent
// launch, each (graph, lmd) chain on one workgroup -- the throughput form).
// With fewer waiting, the GPU is not saturated and a request finishes
// sooner on the single-graph engine, so the lane takes one and runs it
// alone (the latency form).  Requests that arrive while others run join the
// next batch a lane forms, and the lanes' uploads, launches and copies
// overlap on the device.  A request is admitted only while the bytes of
// everything in flight stay within `hbm_budget`.
//
// Graphs can also be registered once (hg_server_register): uploaded and laid
// out by the single-graph engine at registration and kept resident in
// device memory, their bytes counted against the budget until unregistered.
// Requests on a registered graph (hg_server_submit_resident) copy and upload
// nothing; they run on its resident handle, one at a time per graph, and,
// with reuse_results, a decomposition already on the device answers repeat
// decompositions and every densest search (which then only scores cells).
// A request that has not started can be cancelled.
//
// plain = 1 is the baseline: one lane, one request at a time, each on the
// single-graph engine.
//
// The reference has no serving layer (main.cpp runs one method on one graph
// per process); each answer is the one hg_graph_* gives for that graph alone.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "hipgrid.h"
#include "hipgrid_common.h"

using hgc::fail;
using Clock = std::chrono::steady_clock;

namespace {

struct Resident;

struct Request {
    int64_t id = 0;
    Resident* res = nullptr;  // a registered graph, or nullptr (the graph is copied in)
    int32_t kind = HG_REQ_DECOMPOSE;
    double beta = 1.0;
    int64_t n = 0;
    int32_t L = 0;
    uint64_t D = 0;
    uint64_t maxdeg = 0;
    int64_t est_bytes = 0;
    // the graph, copied at submit (buffers left uninitialised before the copy)
    std::vector<std::unique_ptr<int64_t[]>> rp;
    std::vector<std::unique_ptr<uint32_t[]>> col;
    // result
    int32_t status = 1;  // 1 pending, 0 done, < 0 error
    std::string err;
    std::unique_ptr<int32_t[]> core_index;  // [L][n]
    std::vector<int32_t> kmax;
    hg_wds wds{};
    int32_t batch_size = 0;
    Clock::time_point t_submit, t_start, t_done;
};

// A registered graph: its single-graph engine handle on a stream of its own,
// and the host copy of its last decomposition.
struct Resident {
    int64_t gid = 0;
    hg_graph* g = nullptr;   // single-graph engine handle, or
    hg_batch* rb = nullptr;  // (device_queue, graphs of at most large_entries entries) a one-graph batch
    hipStream_t stream = nullptr;
    int64_t n = 0;
    int32_t L = 0;
    uint64_t D = 0;
    int64_t bytes = 0;
    std::mutex run;           // one request at a time on the handle
    bool decomposed = false;  // the handle holds a decomposition (guarded by run)
    std::unique_ptr<int32_t[]> ci;
    std::vector<int32_t> kmax;
    int64_t pending = 0;  // requests queued or running on it (guarded by the server's mutex)
};

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// Device bytes a request holds while it runs (graph, per-lmd state,
// set-up temporaries) on the single-graph engine and in a batch, array by
// array as the allocation cache hands the blocks out (rounded up, up to 2x
// for blocks of at most 512 MB).  Idle blocks the cache keeps between
// requests (up to HG_POOL_MB) are not counted: hg_server_create empties the
// cache, and a full device releases it before an allocation fails.
int64_t engine_bytes(const Request& r) {
    const uint64_t n = uint64_t(r.n), L = uint64_t(r.L), D = r.D, n1 = n + 1;
    auto R = [](uint64_t b) { return hgc::pool_block_bytes(b); };
    const uint64_t Dl = D / (L ? L : 1) + 1;  // a layer's entries (the upload is per layer)
    const uint64_t setup = L * (R(8 * n1) + R(4 * Dl)) + 4 * R(8 * n1) + 2 * R(4 * n) + R(n) + 2 * R(16 * n1);
    const uint64_t graph = R(8 * n1) + R(4 * n) + R(n) + R(4 * D) + R(D + 16) + 2 * R(4 * n * L) +
                           R(4 * ((n + 31) / 32) * L) + R(8 * (D / 64 + n + 1));
    const uint64_t gcap = n + 8192 * 64;  // per-block alive lists: n plus a run of slack per block
    const uint64_t state = R(4 * n * L * L) + 2 * R(4 * n * L) + 2 * R(4 * gcap * L) + R(16 * (D / 256 + n + 1) * L);
    const uint64_t dens = 4 * R(8 * L * (r.maxdeg + 2));
    // set-up also holds the per-vertex bound bytes (n * L) while it builds adjh
    return int64_t(graph + state + dens + setup + hgc::bin_bytes(D, L) + hgc::hb_bytes(D, L) +
                   (hgc::hb_enabled(D, L) ? R(n * L) : 0));
}
int64_t batch_bytes(const Request& r) {
    const uint64_t n = uint64_t(r.n), L = uint64_t(r.L), D = r.D, n1 = n + 1;
    auto R = [](uint64_t b) { return hgc::pool_block_bytes(b); };
    const uint64_t Dl = D / (L ? L : 1) + 1;
    const uint64_t upload = L * (R(8 * n1) + R(4 * Dl)) + R(8 * n1) + R(5 * n);
    const uint64_t graph = R(4 * n1) + R(4 * D) + R(D + 16) + R(n) + R(4 * n * L) + R(4 * n) + R(4 * n * L);
    const uint64_t state = R(4 * n * L * L) + 4 * R(4 * n * L) + R(4 * (n / 32 + 1) * L);
    const uint64_t dens = r.kind == HG_REQ_DENSEST ? R(4 * L * L * (r.maxdeg + 2)) : 0;
    return int64_t(graph + state + dens + upload);
}
// Large requests always run on the engine; small ones may run either way.
int64_t estimate_bytes(const Request& r, bool large) {
    return large ? engine_bytes(r) : std::max(engine_bytes(r), batch_bytes(r));
}

}  // namespace

struct hg_server {
    hg_server_config cfg{};
    int device = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> pending;
    std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;
    std::unordered_map<int64_t, std::unique_ptr<Resident>> residents;
    int64_t next_id = 1, next_gid = 1;
    int64_t inflight = 0, peak = 0, resident_bytes = 0;
    int64_t submitted = 0, completed = 0, failed = 0, cancelled = 0, batches = 0;
    std::vector<double> latencies;
    bool have_first = false;
    Clock::time_point first_submit, last_done;
    bool stop = false;
    std::vector<std::thread> lanes;
    hg_live* live = nullptr;  // device_queue: the shared device work queue
};

namespace {

// Requests that run alone: large ones, every one in plain mode, and those on
// registered graphs (their resident handle).
bool is_large(const hg_server* s, const Request* r) {
    return s->cfg.plain || r->res || int64_t(r->D) > s->cfg.large_entries;
}

void finish(hg_server* s, std::vector<Request*>& rs, int rc, const std::string& err, int64_t bytes) {
    const auto now = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    s->inflight -= bytes;
    for (Request* r : rs) {
        if (r->res) --r->res->pending;
        r->t_done = now;
        r->status = rc == HG_OK ? 0 : rc;
        r->err = err;
        r->rp.clear();
        r->col.clear();
        if (rc == HG_OK) {
            ++s->completed;
            s->latencies.push_back(ms_between(r->t_submit, now));
        } else {
            ++s->failed;
        }
    }
    s->last_done = now;
    ++s->batches;
    s->cv_done.notify_all();
    s->cv_work.notify_all();
}

int run_alone(Request* r, hipStream_t stream) {
    std::vector<hg_layer_csr> lay(size_t(r->L));
    for (int l = 0; l < r->L; ++l) lay[size_t(l)] = {r->rp[size_t(l)].get(), r->col[size_t(l)].get()};
    hg_graph* g = nullptr;
    int rc = hg_graph_create(lay.data(), r->L, r->n, 64, stream, &g);
    if (rc) return rc;
    rc = hg_decompose(g);
    if (!rc) {
        r->core_index.reset(new int32_t[size_t(r->n) * size_t(r->L)]);
        rc = hg_core_index(g, r->core_index.get());
    }
    if (!rc) {
        r->kmax.resize(size_t(r->L));
        rc = hg_cells(g, nullptr, 0, r->kmax.data());
    }
    if (!rc && r->kind == HG_REQ_DENSEST) rc = hg_densest(g, r->beta, &r->wds);
    hg_graph_destroy(g);
    return rc;
}

// A request on a registered graph: decompose on its resident handle unless
// the handle already holds the decomposition and results may be reused,
// then copy the results (and score the cells for a densest search).
int run_resident(hg_server* s, Request* r) {
    Resident* res = r->res;
    std::lock_guard<std::mutex> lk(res->run);
    int rc = HG_OK;
    if (!res->decomposed || !s->cfg.reuse_results) {
        res->decomposed = false;
        if (!res->ci) res->ci.reset(new int32_t[size_t(res->n) * size_t(res->L)]);
        res->kmax.resize(size_t(res->L));
        if (res->rb) {  // the graph's items join the device work queue
            rc = hg_batch_submit_live(res->rb, s->live);
            if (!rc) rc = hg_batch_wait_live(res->rb, nullptr);
            if (!rc) rc = hg_batch_core_index(res->rb, 0, res->ci.get());
            if (!rc) rc = hg_batch_cells(res->rb, 0, nullptr, 0, res->kmax.data());
        } else {
            rc = hg_decompose(res->g);
            if (!rc) rc = hg_core_index(res->g, res->ci.get());
            if (!rc) rc = hg_cells(res->g, nullptr, 0, res->kmax.data());
        }
        if (rc) return rc;
        res->decomposed = true;
    }
    r->core_index.reset(new int32_t[size_t(res->n) * size_t(res->L)]);
    hgc::par_memcpy(r->core_index.get(), res->ci.get(), 4 * size_t(res->n) * size_t(res->L));
    r->kmax = res->kmax;
    if (r->kind == HG_REQ_DENSEST) {
        if (res->rb) {
            const int32_t g0 = 0;
            rc = hg_batch_densest(res->rb, &g0, 1, r->beta, &r->wds);
        } else {
            rc = hg_densest(res->g, r->beta, &r->wds);
        }
    }
    return rc;
}

int run_batch(hg_server* s, std::vector<Request*>& rs, hipStream_t stream) {
    std::vector<std::vector<hg_layer_csr>> lays(rs.size());
    std::vector<hg_batch_graph> gs(rs.size());
    for (size_t i = 0; i < rs.size(); ++i) {
        Request* r = rs[i];
        lays[i].resize(size_t(r->L));
        for (int l = 0; l < r->L; ++l) lays[i][size_t(l)] = {r->rp[size_t(l)].get(), r->col[size_t(l)].get()};
        gs[i].layers = lays[i].data();
        gs[i].L = r->L;
        gs[i].offset_bits = 64;
        gs[i].n = r->n;
    }
    hg_batch* b = nullptr;
    int rc = hg_batch_create(gs.data(), int32_t(gs.size()), stream, &b);
    if (rc) return rc;
    // a batch runs once: launch it directly rather than capture a graph
    hg_batch_params p{};
    if (s->cfg.params.block) p = s->cfg.params;
    else rc = hg_batch_get_params(b, &p);
    p.use_graph = 0;
    if (!rc) rc = hg_batch_set_params(b, &p);
    if (!rc) {
        if (s->live) {  // joins the workgroups already running
            rc = hg_batch_submit_live(b, s->live);
            if (!rc) rc = hg_batch_wait_live(b, nullptr);
        } else {
            rc = hg_batch_decompose(b);
        }
    }
    // densest requests of one beta are scored in one launch each
    std::vector<int32_t> ids;
    std::vector<double> betas;
    for (size_t i = 0; i < rs.size() && !rc; ++i) {
        Request* r = rs[i];
        r->core_index.reset(new int32_t[size_t(r->n) * size_t(r->L)]);
        rc = hg_batch_core_index(b, int32_t(i), r->core_index.get());
        if (!rc) {
            r->kmax.resize(size_t(r->L));
            rc = hg_batch_cells(b, int32_t(i), nullptr, 0, r->kmax.data());
        }
        if (r->kind == HG_REQ_DENSEST && std::find(betas.begin(), betas.end(), r->beta) == betas.end())
            betas.push_back(r->beta);
    }
    for (double beta : betas) {
        if (rc) break;
        ids.clear();
        for (size_t i = 0; i < rs.size(); ++i)
            if (rs[i]->kind == HG_REQ_DENSEST && rs[i]->beta == beta) ids.push_back(int32_t(i));
        std::vector<hg_wds> w(ids.size());
        rc = hg_batch_densest(b, ids.data(), int32_t(ids.size()), beta, w.data());
        for (size_t j = 0; j < ids.size() && !rc; ++j) rs[size_t(ids[j])]->wds = w[j];
    }
    hg_batch_destroy(b);
    return rc;
}

void lane_main(hg_server* s) {
    if (hipSetDevice(s->device) != hipSuccess) return;
    hipStream_t stream = nullptr;
    if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) return;
    for (;;) {
        std::vector<Request*> take;
        int64_t bytes = 0;
        bool large = false, alone = false;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            for (;;) {
                if (!s->pending.empty()) {
                    Request* r = s->pending.front();
                    if (r->est_bytes + s->resident_bytes > s->cfg.hbm_budget) {  // cannot run: fail it
                        s->pending.pop_front();
                        r->status = HG_ERR_NOMEM;
                        r->err = "request exceeds the device-memory budget (registered graphs included)";
                        r->t_done = Clock::now();
                        ++s->failed;
                        s->cv_done.notify_all();
                        continue;
                    }
                    if (s->inflight + s->resident_bytes + r->est_bytes <= s->cfg.hbm_budget) break;
                }
                if (s->stop && s->pending.empty()) {
                    lk.unlock();
                    (void)hipStreamDestroy(stream);
                    return;
                }
                s->cv_work.wait(lk);
            }
            Request* first = s->pending.front();
            large = is_large(s, first);
            if (!large && !s->live) {  // batch only a queue of at least min_batch small requests
                const int32_t want = std::min(s->cfg.min_batch, std::max(1, s->cfg.max_batch));
                int32_t small = 0;
                for (auto it = s->pending.begin(); it != s->pending.end() && small < want && !is_large(s, *it); ++it)
                    ++small;
                alone = small < want;
            }
            const int32_t cap = large || alone ? 1 : std::max(1, s->cfg.max_batch);
            while (!s->pending.empty() && int32_t(take.size()) < cap) {
                Request* r = s->pending.front();
                if (!take.empty() && is_large(s, r)) break;  // a large one runs alone, in arrival order
                if (s->inflight + s->resident_bytes + bytes + r->est_bytes > s->cfg.hbm_budget) break;
                s->pending.pop_front();
                take.push_back(r);
                bytes += r->est_bytes;
            }
            s->inflight += bytes;
            s->peak = std::max(s->peak, s->inflight);
            const auto now = Clock::now();
            for (Request* r : take) {
                r->t_start = now;
                r->batch_size = int32_t(take.size());
            }
        }
        int rc;
        try {
            rc = take[0]->res ? run_resident(s, take[0])
                 : large || alone ? run_alone(take[0], stream) : run_batch(s, take, stream);
        } catch (const std::bad_alloc&) {
            rc = fail(HG_ERR_NOMEM, "host allocation failed");
        }
        finish(s, take, rc, rc ? hg_last_error() : std::string(), bytes);
    }
}

double percentile(std::vector<double> v, double p) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = p * double(v.size() - 1);
    const size_t lo = size_t(pos);
    const size_t hi = std::min(lo + 1, v.size() - 1);
    return v[lo] + (v[hi] - v[lo]) * (pos - double(lo));
}

}  // namespace

extern "C" {

int hg_server_create(const hg_server_config* cfg, hg_server** out) {
    if (!cfg || !out) return fail(HG_ERR_ARG, "NULL argument");
    *out = nullptr;
    if (cfg->hbm_budget <= 0) return fail(HG_ERR_ARG, "hbm_budget must be > 0");
    if (!cfg->plain && (cfg->lanes < 1 || cfg->lanes > 16)) return fail(HG_ERR_ARG, "lanes must be in [1, 16]");
    if (!cfg->plain && cfg->max_batch < 1) return fail(HG_ERR_ARG, "max_batch must be >= 1");
    std::unique_ptr<hg_server> s(new (std::nothrow) hg_server());
    if (!s) return fail(HG_ERR_NOMEM, "host allocation failed");
    // Start from an empty device-memory cache: blocks of earlier, differently
    // sized work would otherwise only be evicted (hipFree, which stalls every
    // lane) one by one while requests are served.
    hgc::pool_trim();
    s->cfg = *cfg;
    if (s->cfg.min_batch < 0) return fail(HG_ERR_ARG, "min_batch must be >= 0");
    if (s->cfg.min_batch == 0) s->cfg.min_batch = 8;
    if (s->cfg.plain) {
        s->cfg.lanes = 1;
        s->cfg.max_batch = 1;
    }
    HIP_TRY(hipGetDevice(&s->device));
    if (s->cfg.device_queue && !s->cfg.plain) {
        hg_batch_params lp{};
        lp.block = 1024;
        lp.unroll = 4;
        lp.blocks_per_cu = 1;
        if (s->cfg.params.block) lp = s->cfg.params;
        int rc = hg_live_create(&lp, 0, &s->live);
        if (rc) return rc;
    }
    try {
        for (int i = 0; i < s->cfg.lanes; ++i) s->lanes.emplace_back(lane_main, s.get());
    } catch (const std::exception&) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        s->cv_work.notify_all();
        for (auto& t : s->lanes) t.join();
        if (s->live) hg_live_destroy(s->live);
        return fail(HG_ERR_NOMEM, "cannot start lane threads");
    }
    *out = s.release();
    return HG_OK;
}

int hg_server_submit(hg_server* s, const hg_batch_graph* graph, int32_t kind, double beta, int64_t* id) {
    if (!s || !graph || !id) return fail(HG_ERR_ARG, "NULL argument");
    if (kind != HG_REQ_DECOMPOSE && kind != HG_REQ_DENSEST) return fail(HG_ERR_ARG, "unknown request kind");
    if (!graph->layers) return fail(HG_ERR_ARG, "layers is NULL");
    if (graph->L < 1 || graph->L > 64) return fail(HG_ERR_RANGE, "L must be in [1, 64]");
    if (graph->n < 1) return fail(HG_ERR_ARG, "n must be >= 1");
    if (graph->offset_bits != 32 && graph->offset_bits != 64) return fail(HG_ERR_ARG, "offset_bits must be 32 or 64");
    std::unique_ptr<Request> r(new (std::nothrow) Request());
    if (!r) return fail(HG_ERR_NOMEM, "host allocation failed");
    r->t_submit = Clock::now();
    r->kind = kind;
    r->beta = beta;
    r->n = graph->n;
    r->L = graph->L;
    try {
        const size_t n = size_t(graph->n);
        r->rp.resize(size_t(graph->L));
        r->col.resize(size_t(graph->L));
        for (int l = 0; l < graph->L; ++l) {
            const hg_layer_csr& lay = graph->layers[l];
            if (!lay.row_ptr || !lay.col) return fail(HG_ERR_ARG, "layer pointer is NULL");
            int64_t* rp = new int64_t[n + 1];
            r->rp[size_t(l)].reset(rp);
            // copy (widening 32-bit offsets), then check, each in parallel ranges
            hgc::par_ranges(n + 1, size_t(1) << 20, [&](size_t b, size_t e) {
                if (graph->offset_bits == 64) {
                    std::memcpy(rp + b, static_cast<const int64_t*>(lay.row_ptr) + b, 8 * (e - b));
                } else {
                    const int32_t* src = static_cast<const int32_t*>(lay.row_ptr);
                    for (size_t v = b; v < e; ++v) rp[v] = src[v];
                }
            });
            std::atomic<bool> bad{false};
            std::atomic<uint64_t> maxdeg{r->maxdeg};
            hgc::par_ranges(n, size_t(1) << 20, [&](size_t b, size_t e) {
                uint64_t md = 0;
                bool bd = false;
                for (size_t v = b; v < e; ++v) {
                    bd |= rp[v + 1] < rp[v];
                    md = std::max<uint64_t>(md, uint64_t(rp[v + 1] - rp[v]));
                }
                if (bd) bad = true;
                uint64_t cur = maxdeg.load();
                while (md > cur && !maxdeg.compare_exchange_weak(cur, md)) {
                }
            });
            if (rp[0] != 0) return fail(HG_ERR_ARG, "row_ptr[0] must be 0");
            if (bad) return fail(HG_ERR_ARG, "row_ptr must be non-decreasing");
            r->maxdeg = maxdeg.load();
            const size_t m = size_t(rp[n]);
            uint32_t* col = new uint32_t[std::max<size_t>(m, 1)];  // a non-NULL pointer for an empty layer
            r->col[size_t(l)].reset(col);
            col[0] = 0;
            if (m) hgc::par_memcpy(col, lay.col, 4 * m);
            r->D += m;
        }
    } catch (const std::bad_alloc&) {
        return fail(HG_ERR_NOMEM, "host allocation failed");
    }
    std::lock_guard<std::mutex> lk(s->mu);
    r->est_bytes = estimate_bytes(*r, is_large(s, r.get()));
    r->id = s->next_id++;
    *id = r->id;
    if (!s->have_first) {
        s->have_first = true;
        s->first_submit = r->t_submit;
    }
    ++s->submitted;
    s->pending.push_back(r.get());
    s->reqs.emplace(r->id, std::move(r));
    s->cv_work.notify_one();
    return HG_OK;
}

int hg_server_register(hg_server* s, const hg_batch_graph* graph, int64_t* gid) {
    if (!s || !graph || !gid) return fail(HG_ERR_ARG, "NULL argument");
    if (!graph->layers) return fail(HG_ERR_ARG, "layers is NULL");
    std::unique_ptr<Resident> res(new (std::nothrow) Resident());
    if (!res) return fail(HG_ERR_NOMEM, "host allocation failed");
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipStreamCreateWithFlags(&res->stream, hipStreamNonBlocking));
    // entries, to pick the form: with a device queue, graphs the loop would
    // batch are kept as one-graph batches whose items join the queue
    int64_t entries = 0;
    for (int l = 0; l < graph->L; ++l)
        entries += graph->offset_bits == 64 ? static_cast<const int64_t*>(graph->layers[l].row_ptr)[graph->n]
                                            : int64_t(static_cast<const int32_t*>(graph->layers[l].row_ptr)[graph->n]);
    int rc;
    int64_t D = 0, bytes = 0;
    if (s->live && entries <= s->cfg.large_entries) {
        rc = hg_batch_create(graph, 1, res->stream, &res->rb);
        int32_t cnt = 0;
        int64_t items = 0;
        if (!rc) rc = hg_batch_info(res->rb, &cnt, &items, &D, &bytes);
    } else {
        rc = hg_graph_create(graph->layers, graph->L, graph->n, graph->offset_bits, res->stream, &res->g);
        if (!rc) rc = hg_graph_info(res->g, nullptr, nullptr, &D, &bytes);
    }
    if (rc) {
        const std::string err = hg_last_error();
        if (res->g) hg_graph_destroy(res->g);
        if (res->rb) hg_batch_destroy(res->rb);
        (void)hipStreamDestroy(res->stream);
        return fail(rc, err);
    }
    res->n = graph->n;
    res->L = graph->L;
    res->D = uint64_t(D);
    // a densest search adds the edge histograms, L * (largest degree + 2)
    // words, kept on the handle
    uint64_t maxdeg = 0;
    for (int l = 0; l < graph->L; ++l)
        for (int64_t v = 0; v < graph->n; ++v) {
            const int64_t d = graph->offset_bits == 64
                                  ? static_cast<const int64_t*>(graph->layers[l].row_ptr)[v + 1] -
                                        static_cast<const int64_t*>(graph->layers[l].row_ptr)[v]
                                  : int64_t(static_cast<const int32_t*>(graph->layers[l].row_ptr)[v + 1]) -
                                        static_cast<const int32_t*>(graph->layers[l].row_ptr)[v];
            maxdeg = std::max<uint64_t>(maxdeg, uint64_t(std::max<int64_t>(d, 0)));
        }
    res->bytes = bytes + int64_t(8 * uint64_t(graph->L) * (maxdeg + 2));
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->resident_bytes + res->bytes > s->cfg.hbm_budget) {
        if (res->g) hg_graph_destroy(res->g);
        if (res->rb) hg_batch_destroy(res->rb);
        (void)hipStreamDestroy(res->stream);
        return fail(HG_ERR_NOMEM, "registered graphs would exceed the device-memory budget");
    }
    res->gid = s->next_gid++;
    *gid = res->gid;
    s->resident_bytes += res->bytes;
    s->peak = std::max(s->peak, s->inflight + s->resident_bytes);
    s->residents.emplace(res->gid, std::move(res));
    return HG_OK;
}

int hg_server_submit_resident(hg_server* s, int64_t gid, int32_t kind, double beta, int64_t* id) {
    if (!s || !id) return fail(HG_ERR_ARG, "NULL argument");
    if (kind != HG_REQ_DECOMPOSE && kind != HG_REQ_DENSEST) return fail(HG_ERR_ARG, "unknown request kind");
    if (kind == HG_REQ_DENSEST && !std::isfinite(beta)) return fail(HG_ERR_ARG, "beta must be finite");
    std::unique_ptr<Request> r(new (std::nothrow) Request());
    if (!r) return fail(HG_ERR_NOMEM, "host allocation failed");
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->residents.find(gid);
    if (it == s->residents.end()) return fail(HG_ERR_ARG, "unknown graph id");
    Resident* res = it->second.get();
    r->res = res;
    r->kind = kind;
    r->beta = beta;
    r->n = res->n;
    r->L = res->L;
    r->D = res->D;
    r->est_bytes = 0;  // the resident bytes are counted already
    r->id = s->next_id++;
    *id = r->id;
    ++res->pending;
    if (!s->have_first) {
        s->have_first = true;
        s->first_submit = r->t_submit;
    }
    ++s->submitted;
    s->pending.push_back(r.get());
    s->reqs.emplace(r->id, std::move(r));
    s->cv_work.notify_one();
    return HG_OK;
}

int hg_server_unregister(hg_server* s, int64_t gid) {
    if (!s) return fail(HG_ERR_ARG, "server is NULL");
    std::unique_ptr<Resident> res;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->residents.find(gid);
        if (it == s->residents.end()) return fail(HG_ERR_ARG, "unknown graph id");
        if (it->second->pending > 0) return fail(HG_ERR_STATE, "requests on the graph are pending");
        res = std::move(it->second);
        s->residents.erase(it);
        s->resident_bytes -= res->bytes;
        s->cv_work.notify_all();
    }
    if (res->g) hg_graph_destroy(res->g);
    if (res->rb) hg_batch_destroy(res->rb);
    (void)hipStreamDestroy(res->stream);
    return HG_OK;
}

int hg_server_cancel(hg_server* s, int64_t id) {
    if (!s) return fail(HG_ERR_ARG, "server is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(HG_ERR_ARG, "unknown request id");
    Request* r = it->second.get();
    auto q = std::find(s->pending.begin(), s->pending.end(), r);
    if (q == s->pending.end()) {
        if (r->status == HG_ERR_CANCELLED) return HG_OK;
        return fail(HG_ERR_STATE, r->status == 1 ? "request already started" : "request already done");
    }
    s->pending.erase(q);
    if (r->res) --r->res->pending;
    r->status = HG_ERR_CANCELLED;
    r->err = "request cancelled";
    r->t_done = Clock::now();
    r->rp.clear();
    r->col.clear();
    ++s->cancelled;
    s->cv_done.notify_all();
    s->cv_work.notify_all();
    return HG_OK;
}

int hg_debug_fail_alloc(int64_t after) {
    if (after < 0) return fail(HG_ERR_ARG, "after must be >= 0");
    hgc::fail_countdown().store(after);
    return HG_OK;
}

int hg_server_wait(hg_server* s, int64_t id, hg_request_info* info) {
    if (!s) return fail(HG_ERR_ARG, "server is NULL");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(HG_ERR_ARG, "unknown request id");
    Request* r = it->second.get();
    s->cv_done.wait(lk, [&] { return r->status != 1; });
    if (info) {
        std::memset(info, 0, sizeof *info);
        info->status = r->status;
        info->kind = r->kind;
        info->n = r->n;
        info->L = r->L;
        info->batch_size = r->batch_size;
        info->queue_ms = r->status == 0 ? ms_between(r->t_submit, r->t_start) : 0.0;
        info->run_ms = r->status == 0 ? ms_between(r->t_start, r->t_done) : 0.0;
        info->latency_ms = ms_between(r->t_submit, r->t_done);
        info->wds = r->wds;
    }
    if (r->status < 0) return fail(r->status, r->err);
    return HG_OK;
}

int hg_server_result(hg_server* s, int64_t id, int32_t* core_index, int32_t* kmax) {
    if (!s) return fail(HG_ERR_ARG, "server is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(HG_ERR_ARG, "unknown request id");
    Request* r = it->second.get();
    if (r->status == 1) return fail(HG_ERR_STATE, "request still pending");
    if (r->status < 0) return fail(r->status, r->err);
    if (core_index && r->core_index) hgc::par_memcpy(core_index, r->core_index.get(), 4 * size_t(r->n) * size_t(r->L));
    if (kmax) std::memcpy(kmax, r->kmax.data(), 4 * r->kmax.size());
    return HG_OK;
}

int hg_server_release(hg_server* s, int64_t id) {
    if (!s) return fail(HG_ERR_ARG, "server is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(HG_ERR_ARG, "unknown request id");
    if (it->second->status == 1) return fail(HG_ERR_STATE, "request still pending");
    s->reqs.erase(it);
    return HG_OK;
}

int hg_server_drain(hg_server* s) {
    if (!s) return fail(HG_ERR_ARG, "server is NULL");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->completed + s->failed + s->cancelled == s->submitted; });
    return HG_OK;
}

int hg_server_stats(hg_server* s, hg_server_summary* out) {
    if (!s || !out) return fail(HG_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(out, 0, sizeof *out);
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->failed = s->failed;
    out->batches = s->batches;
    out->cancelled = s->cancelled;
    out->resident_graphs = int64_t(s->residents.size());
    out->resident_bytes = s->resident_bytes;
    out->p50_ms = percentile(s->latencies, 0.50);
    out->p99_ms = percentile(s->latencies, 0.99);
    double sum = 0, mx = 0;
    for (double x : s->latencies) {
        sum += x;
        mx = std::max(mx, x);
    }
    out->mean_ms = s->latencies.empty() ? 0.0 : sum / double(s->latencies.size());
    out->max_ms = mx;
    out->span_s = s->have_first && s->completed ? std::chrono::duration<double>(s->last_done - s->first_submit).count() : 0.0;
    out->peak_bytes = s->peak;
    return HG_OK;
}

int hg_server_destroy(hg_server* s) {
    if (!s) return fail(HG_ERR_ARG, "server is NULL");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->lanes) t.join();
    for (auto& kv : s->residents) {
        if (kv.second->g) hg_graph_destroy(kv.second->g);
        if (kv.second->rb) hg_batch_destroy(kv.second->rb);
        (void)hipStreamDestroy(kv.second->stream);
    }
    if (s->live) hg_live_destroy(s->live);
    delete s;
    return HG_OK;
}

}  // extern "C"
