// hipgrid_server.cpp -- request loop over one GPU (hg_server_* in include/hipgrid.h).
//
// Requests (a graph plus "decompose" or "densest") are copied into the
// server at submit and queue in arrival order.  `lanes` worker threads, each
// with its own HIP stream, take work whenever they are idle: a request with
// more than `large_entries` directed entries runs alone on the single-graph
// engine (hg_graph_*, the whole GPU per level); otherwise the lane takes the
// queue's leading small requests, up to `max_batch` and as many as the
// device-memory budget admits, and runs them as one batch (hg_batch_*: one
// persistent launch).  Requests that arrive while batches are on the GPU
// join the next batch a lane forms, and the lanes' uploads, launches and
// copies overlap on the device.  A request is admitted only while the bytes
// of everything in flight stay within `hbm_budget`.
//
// plain = 1 is the baseline: one lane, one request at a time, each on the
// single-graph engine.
//
// The reference has no serving layer (main.cpp runs one method on one graph
// per process); each answer is the one hg_graph_* gives for that graph alone.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
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

struct Request {
    int64_t id = 0;
    int32_t kind = HG_REQ_DECOMPOSE;
    double beta = 1.0;
    int64_t n = 0;
    int32_t L = 0;
    uint64_t D = 0;
    uint64_t maxdeg = 0;
    int64_t est_bytes = 0;
    std::vector<std::vector<int64_t>> rp;
    std::vector<std::vector<uint32_t>> col;
    // result
    int32_t status = 1;  // 1 pending, 0 done, < 0 error
    std::string err;
    std::vector<int32_t> core_index;
    std::vector<int32_t> kmax;
    hg_wds wds{};
    int32_t batch_size = 0;
    Clock::time_point t_submit, t_start, t_done;
};

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// Device bytes a request holds while it runs (graph, per-lmd state,
// upload temporaries), on the engine that will run it.
int64_t estimate_bytes(const Request& r, bool large) {
    const uint64_t n = uint64_t(r.n), L = uint64_t(r.L), D = r.D;
    const uint64_t upload = 8 * L * (n + 1) + 4 * D + 8 * (n + 1) + 5 * n;
    if (large) {
        const uint64_t graph = 8 * (n + 1) + 5 * D + 4 * n * L + 5 * n + 4 * n * L + D / 8 + 8;
        const uint64_t state = L * (4 * n * L + 8 * n + 8 * n + n / 8 + 16 * (D / 4096 + n + 1));
        const uint64_t dens = 8 * L * (r.maxdeg + 2) * 4;
        return int64_t(graph + state + dens + upload);
    }
    const uint64_t graph = 4 * (n + 1) + 5 * D + 5 * n + 4 * n * L + 4 * n * L;
    const uint64_t state = L * (4 * n * L + 16 * n + 4 * (n / 32 + 1));
    const uint64_t dens = r.kind == HG_REQ_DENSEST ? 4 * L * L * (r.maxdeg + 2) : 0;
    return int64_t(graph + state + dens + upload);
}

}  // namespace

struct hg_server {
    hg_server_config cfg{};
    int device = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> pending;
    std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;
    int64_t next_id = 1;
    int64_t inflight = 0, peak = 0;
    int64_t submitted = 0, completed = 0, failed = 0, batches = 0;
    std::vector<double> latencies;
    bool have_first = false;
    Clock::time_point first_submit, last_done;
    bool stop = false;
    std::vector<std::thread> lanes;
};

namespace {

bool is_large(const hg_server* s, const Request* r) { return s->cfg.plain || int64_t(r->D) > s->cfg.large_entries; }

void finish(hg_server* s, std::vector<Request*>& rs, int rc, const std::string& err, int64_t bytes) {
    const auto now = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    s->inflight -= bytes;
    for (Request* r : rs) {
        r->t_done = now;
        r->status = rc == HG_OK ? 0 : rc;
        r->err = err;
        r->rp.clear();
        r->col.clear();
        r->rp.shrink_to_fit();
        r->col.shrink_to_fit();
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

int run_large(Request* r, hipStream_t stream) {
    std::vector<hg_layer_csr> lay(size_t(r->L));
    for (int l = 0; l < r->L; ++l) lay[size_t(l)] = {r->rp[size_t(l)].data(), r->col[size_t(l)].data()};
    hg_graph* g = nullptr;
    int rc = hg_graph_create(lay.data(), r->L, r->n, 64, stream, &g);
    if (rc) return rc;
    rc = hg_decompose(g);
    if (!rc) {
        r->core_index.resize(size_t(r->n) * size_t(r->L));
        rc = hg_core_index(g, r->core_index.data());
    }
    if (!rc) {
        r->kmax.resize(size_t(r->L));
        rc = hg_cells(g, nullptr, 0, r->kmax.data());
    }
    if (!rc && r->kind == HG_REQ_DENSEST) rc = hg_densest(g, r->beta, &r->wds);
    hg_graph_destroy(g);
    return rc;
}

int run_batch(hg_server* s, std::vector<Request*>& rs, hipStream_t stream) {
    std::vector<std::vector<hg_layer_csr>> lays(rs.size());
    std::vector<hg_batch_graph> gs(rs.size());
    for (size_t i = 0; i < rs.size(); ++i) {
        Request* r = rs[i];
        lays[i].resize(size_t(r->L));
        for (int l = 0; l < r->L; ++l) lays[i][size_t(l)] = {r->rp[size_t(l)].data(), r->col[size_t(l)].data()};
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
    if (!rc) rc = hg_batch_decompose(b);
    // densest requests of one beta are scored in one launch each
    std::vector<int32_t> ids;
    std::vector<double> betas;
    for (size_t i = 0; i < rs.size() && !rc; ++i) {
        Request* r = rs[i];
        r->core_index.resize(size_t(r->n) * size_t(r->L));
        rc = hg_batch_core_index(b, int32_t(i), r->core_index.data());
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
        bool large = false;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            for (;;) {
                if (!s->pending.empty()) {
                    Request* r = s->pending.front();
                    if (r->est_bytes > s->cfg.hbm_budget) {  // can never run: fail it
                        s->pending.pop_front();
                        r->status = HG_ERR_NOMEM;
                        r->err = "request exceeds the device-memory budget";
                        r->t_done = Clock::now();
                        ++s->failed;
                        s->cv_done.notify_all();
                        continue;
                    }
                    if (s->inflight + r->est_bytes <= s->cfg.hbm_budget) break;
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
            const int32_t cap = large ? 1 : std::max(1, s->cfg.max_batch);
            while (!s->pending.empty() && int32_t(take.size()) < cap) {
                Request* r = s->pending.front();
                if (!take.empty() && is_large(s, r)) break;  // a large one runs alone, in arrival order
                if (s->inflight + bytes + r->est_bytes > s->cfg.hbm_budget) break;
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
            rc = large ? run_large(take[0], stream) : run_batch(s, take, stream);
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
    s->cfg = *cfg;
    if (s->cfg.plain) {
        s->cfg.lanes = 1;
        s->cfg.max_batch = 1;
    }
    HIP_TRY(hipGetDevice(&s->device));
    try {
        for (int i = 0; i < s->cfg.lanes; ++i) s->lanes.emplace_back(lane_main, s.get());
    } catch (const std::exception&) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        s->cv_work.notify_all();
        for (auto& t : s->lanes) t.join();
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
            std::vector<int64_t>& rp = r->rp[size_t(l)];
            rp.resize(n + 1);
            if (graph->offset_bits == 64) {
                std::memcpy(rp.data(), lay.row_ptr, 8 * (n + 1));
            } else {
                const int32_t* src = static_cast<const int32_t*>(lay.row_ptr);
                for (size_t v = 0; v <= n; ++v) rp[v] = src[v];
            }
            if (rp[0] != 0) return fail(HG_ERR_ARG, "row_ptr[0] must be 0");
            for (size_t v = 0; v < n; ++v) {
                if (rp[v + 1] < rp[v]) return fail(HG_ERR_ARG, "row_ptr must be non-decreasing");
                r->maxdeg = std::max<uint64_t>(r->maxdeg, uint64_t(rp[v + 1] - rp[v]));
            }
            const size_t m = size_t(rp[n]);
            r->col[size_t(l)].assign(lay.col, lay.col + m);
            if (m == 0) r->col[size_t(l)].push_back(0);  // a non-NULL pointer for an empty layer
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
    if (core_index) std::memcpy(core_index, r->core_index.data(), 4 * r->core_index.size());
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
    s->cv_done.wait(lk, [&] { return s->completed + s->failed == s->submitted; });
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
    delete s;
    return HG_OK;
}

}  // extern "C"
