// server.hpp — continuous-batching request loop over resident engines
// (hspz_server_* in include/hspz.h; included by engine.hip).
//
// Requests (one query block each) queue FIFO. Each worker thread owns one
// engine (its own stream and buffers, the same reference resident on the same
// device) and repeatedly admits the longest FIFO prefix of waiting requests
// whose estimated device footprint fits the remaining HBM budget (shared by
// all workers), runs one hspz_filter_batch over them, and publishes every
// request's lists. Requests that arrive while passes are in flight wait for
// the next admission, so with two workers one pass computes while the next is
// admitted and uploaded: the device work overlaps. Results per request equal
// hspz_run_pass on that request alone.
#pragma once

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <unordered_set>

namespace {

struct ServerReq {
    uint64_t id;
    std::vector<uint8_t> fwd, rc;
    bool raw;  // fwd holds raw bytes; the device encodes both strands
    uint64_t footprint;
    double t_submit;
};

struct ServerDone {
    int status = 0;
    std::string msg;  // the worker's error text when status != 0
    std::vector<hspz_segment_pair> hsps;
    std::vector<size_t> counts;
    double t_submit = 0, t_start = 0, t_done = 0;
    uint32_t batch_id = 0, batch_requests = 0;
};

}  // namespace

struct hspz_server {
    std::vector<hspz_engine *> engines;
    hspz_pass_params params;
    std::string pattern;
    uint64_t budget = 0, in_flight = 0;
    uint32_t max_reqs = 256;
    uint64_t max_bases = 64ull << 20;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<ServerReq> queue;
    std::unordered_map<uint64_t, ServerDone> done;
    std::unordered_set<uint64_t> taken;  // ids whose result a wait has already collected
    uint64_t next_id = 1;
    uint32_t next_batch = 0;
    bool stopping = false;
    std::vector<std::thread> workers;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    hspz_server_stats stats = {};

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

// Device bytes one request needs inside a batched pass: per query base both
// strands' positions (valid flag 1 + id 4) and up to 13 seeds each (seed 8 +
// count 4 + prefix 8, and for the k-mer-major order the grouped record 8 +
// k-mer 4 + query record 64), the arena and its planes, plus a fixed share of
// the per-pass survivor lists.
uint64_t request_footprint(size_t len) { return (uint64_t)len * (2 * (5 + 13 * 96) + 3) + (1ull << 20); }

void server_worker(hspz_server *s, size_t w) {
    hspz_engine *h = s->engines[w];
    for (;;) {
        std::vector<ServerReq> batch;
        uint64_t fp = 0;
        uint32_t bid;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            for (;;) {
                if (s->queue.empty()) {
                    if (s->stopping) return;
                    s->cv_work.wait(lk);
                    continue;
                }
                // admit the longest FIFO prefix that fits the free budget
                uint64_t bases = 0;
                size_t take = 0;
                for (auto &r : s->queue) {
                    if (take >= s->max_reqs) break;
                    if (s->in_flight + fp + r.footprint > s->budget) break;
                    if (take && bases + r.fwd.size() > s->max_bases) break;
                    fp += r.footprint;
                    bases += r.fwd.size();
                    take++;
                }
                if (take == 0) {  // the head waits for budget held by other passes
                    fp = 0;
                    s->cv_work.wait(lk);
                    continue;
                }
                for (size_t i = 0; i < take; i++) {
                    batch.push_back(std::move(s->queue.front()));
                    s->queue.pop_front();
                }
                s->in_flight += fp;
                bid = s->next_batch++;
                break;
            }
        }
        double t_start = s->now();
        std::vector<hspz_query_chunk> chunks(batch.size());
        for (size_t i = 0; i < batch.size(); i++)
            chunks[i] = {batch[i].fwd.data(), batch[i].raw ? nullptr : batch[i].rc.data(), batch[i].fwd.size()};
        hspz_batch_output out;
        hspz_pass_stats st;
        int rc = hspz_filter_batch(h, &s->params, chunks.data(), chunks.size(), &out, &st);
        std::string msg = rc ? std::string(hspz_last_error()) : std::string();
        double t_done = s->now();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            size_t at = 0, ci = 0;
            for (size_t i = 0; i < batch.size(); i++) {
                ServerDone d;
                d.status = rc;
                d.msg = msg;
                d.t_submit = batch[i].t_submit;
                d.t_start = t_start;
                d.t_done = t_done;
                d.batch_id = bid;
                d.batch_requests = (uint32_t)batch.size();
                if (rc == 0) {
                    size_t ni = out.chunk_intervals[i];
                    size_t n = 0;
                    for (size_t k = 0; k < 2 * ni; k++) {
                        d.counts.push_back(out.counts[ci + k]);
                        n += out.counts[ci + k];
                    }
                    d.hsps.assign(out.hsps + at, out.hsps + at + n);
                    at += n;
                    ci += 2 * ni;
                }
                s->done[batch[i].id] = std::move(d);
            }
            s->in_flight -= fp;
            s->stats.batches++;
            s->stats.requests += batch.size();
            s->stats.bases += 0;
            for (auto &r : batch) s->stats.bases += r.fwd.size();
            s->stats.hits += rc == 0 ? st.hits : 0;
            s->stats.max_batch_requests = std::max<uint64_t>(s->stats.max_batch_requests, batch.size());
            if (rc) s->stats.failed_batches++;
        }
        if (rc == 0) hspz_batch_output_free(&out);
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
}

}  // namespace

extern "C" int hspz_server_new(hspz_engine **engines, size_t n_engines, const hspz_server_config *cfg,
                               hspz_server **out) {
    clear_error();
    if (!engines || n_engines == 0 || !cfg || !out) return set_error(HSPZ_E_ARG, "null argument");
    *out = nullptr;
    for (size_t i = 0; i < n_engines; i++)
        if (!engines[i]) return set_error(HSPZ_E_ARG, "engine %zu is null", i);
    hspz_server *s = new hspz_server();
    s->engines.assign(engines, engines + n_engines);
    s->params = cfg->params;
    s->pattern = cfg->params.pattern ? cfg->params.pattern : "12of19";
    s->params.pattern = s->pattern.c_str();
    if (s->params.wga_chunk == 0) s->params.wga_chunk = 250000;
    if (s->params.interval_size == 0) s->params.interval_size = 10000000;
    s->budget = cfg->hbm_budget;
    if (s->budget == 0) {
        uint64_t f = 0, t = 0;
        if (hspz_device_memory(engines[0]->device, &f, &t) == 0) s->budget = f / 4;
        if (s->budget == 0) s->budget = 16ull << 30;
    }
    if (cfg->max_batch_requests) s->max_reqs = cfg->max_batch_requests;
    if (cfg->max_batch_bases) s->max_bases = cfg->max_batch_bases;
    for (size_t w = 0; w < n_engines; w++) s->workers.emplace_back(server_worker, s, w);
    *out = s;
    return 0;
}

extern "C" int hspz_server_submit(hspz_server *s, const uint8_t *fwd, const uint8_t *rc, size_t len,
                                  uint64_t *request_id) {
    clear_error();
    if (!s || !request_id || (len && !fwd)) return set_error(HSPZ_E_ARG, "null argument");
    uint64_t fp = request_footprint(len);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stopping) return set_error(HSPZ_E_ARG, "server is shutting down");
    if (fp > s->budget) {
        s->stats.rejected++;
        return set_error(HSPZ_E_OOM, "request of %zu bases needs %llu device bytes, over the %llu-byte budget", len,
                         (unsigned long long)fp, (unsigned long long)s->budget);
    }
    ServerReq r;
    r.id = s->next_id++;
    r.fwd.assign(fwd, fwd + len);
    r.raw = rc == nullptr;
    if (rc) r.rc.assign(rc, rc + len);
    r.footprint = fp;
    r.t_submit = s->now();
    *request_id = r.id;
    s->queue.push_back(std::move(r));
    s->cv_work.notify_one();
    return 0;
}

extern "C" int hspz_server_wait(hspz_server *s, uint64_t id, hspz_request_result *out) {
    clear_error();
    if (!s || !out) return set_error(HSPZ_E_ARG, "null argument");
    memset(out, 0, sizeof *out);
    std::unique_lock<std::mutex> lk(s->mu);
    if (id == 0 || id >= s->next_id) return set_error(HSPZ_E_ARG, "unknown request %llu", (unsigned long long)id);
    // a result is handed out once: a later (or concurrent second) wait on the
    // same id gets an error instead of blocking forever
    s->cv_done.wait(lk, [&] { return s->done.count(id) != 0 || s->taken.count(id) != 0; });
    if (s->taken.count(id))
        return set_error(HSPZ_E_ARG, "request %llu was already collected", (unsigned long long)id);
    ServerDone d = std::move(s->done[id]);
    s->done.erase(id);
    s->taken.insert(id);
    lk.unlock();
    s->cv_done.notify_all();
    out->status = d.status;
    out->t_submit = d.t_submit;
    out->t_start = d.t_start;
    out->t_done = d.t_done;
    out->batch_id = d.batch_id;
    out->batch_requests = d.batch_requests;
    out->n_intervals = d.counts.size() / 2;
    out->n_hsps = d.hsps.size();
    out->hsps = (hspz_segment_pair *)aligned_alloc(16, std::max<size_t>(d.hsps.size(), 1) * 16);
    out->counts = (size_t *)malloc(std::max<size_t>(d.counts.size(), 1) * sizeof(size_t));
    if (!out->hsps || !out->counts) return set_error(HSPZ_E_OOM, "host output allocation");
    memcpy(out->hsps, d.hsps.data(), d.hsps.size() * 16);
    memcpy(out->counts, d.counts.data(), d.counts.size() * sizeof(size_t));
    if (d.status) return set_error(d.status, "request %llu failed in its batch: %s", (unsigned long long)id, d.msg.c_str());
    return 0;
}

extern "C" void hspz_request_result_free(hspz_request_result *r) {
    if (!r) return;
    free(r->hsps);
    free(r->counts);
    r->hsps = nullptr;
    r->counts = nullptr;
}

extern "C" int hspz_server_get_stats(hspz_server *s, hspz_server_stats *out) {
    clear_error();
    if (!s || !out) return set_error(HSPZ_E_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    out->queued = s->queue.size();
    out->in_flight_bytes = s->in_flight;
    return 0;
}

extern "C" void hspz_server_free(hspz_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    delete s;
}
