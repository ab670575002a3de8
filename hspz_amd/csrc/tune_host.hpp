// tune_host.hpp — autotuner of the batched gate (hspz_autotune_batch in
// include/hspz.h; included by engine.hip after batch_host.hpp).
//
// The batched gate has five knobs that change its speed and never its
// output: the hit order (seed-major, or k-mer-major with reference windows
// reused across a k-mer's query seeds), cells per step (16/32/64, the width of
// the score tile a lane evaluates between drop tests), threads per block
// (512/1024, sharing one 64 KiB LDS map table), seeds per wave claim (the
// granule of the grid-wide work counter), and the persistent grid. The tuner
// runs the whole batched pass over the caller's chunks for each cell, takes
// the gate and materializer times from the pass's own stream events, and
// keeps the fastest cell.
#pragma once

#include <cmath>

namespace {

struct TuneKnobs {
    int step, threads, claim, blocks, order;
};

TuneKnobs tune_get(const hspz_engine *h) {
    return {h->bgate_step, h->bgate_threads, h->bgate_claim, h->bgate_blocks, h->batch_order};
}

void tune_set(hspz_engine *h, const TuneKnobs &k) {
    h->bgate_step = k.step;
    h->bgate_threads = k.threads;
    h->bgate_claim = k.claim;
    h->bgate_blocks = k.blocks;
    h->batch_order = k.order;
}

}  // namespace

extern "C" int hspz_autotune_batch(hspz_engine *h, const hspz_pass_params *p, const hspz_query_chunk *chunks,
                                   size_t n_chunks, uint32_t reps, hspz_tune_cell *cells, size_t cap,
                                   size_t *n_cells, hspz_tune_cell *best) {
    clear_error();
    if (!h || !p || !n_cells || (cap && !cells) || (n_chunks && !chunks)) return set_error(HSPZ_E_ARG, "null argument");
    *n_cells = 0;
    if (reps == 0) reps = 1;
    const TuneKnobs saved = tune_get(h);
    std::vector<hspz_tune_cell> all;
    std::vector<TuneKnobs> knobs;
    auto measure = [&](const TuneKnobs &k) {
        tune_set(h, k);
        hspz_tune_cell c;
        memset(&c, 0, sizeof c);
        c.step_cells = (uint32_t)k.step;
        c.block_threads = (uint32_t)k.threads;
        c.claim_seeds = (uint32_t)k.claim;
        c.order = (uint32_t)k.order;
        double best_sum = HUGE_VAL;
        for (uint32_t r = 0; r < reps; r++) {
            hspz_batch_output out;
            hspz_pass_stats st;
            memset(&st, 0, sizeof st);
            int rc = hspz_filter_batch(h, p, chunks, n_chunks, &out, &st);
            if (rc) {
                c.status = rc;
                break;
            }
            hspz_batch_output_free(&out);
            double sum = h->last_gate_ms + h->last_post_ms;
            if (sum < best_sum) {
                best_sum = sum;
                c.gate_ms = h->last_gate_ms;
                c.after_ms = h->last_post_ms;
            }
            c.hits = st.hits;
            c.survivors = h->last_surv;
            c.deferred = h->last_deferred;
        }
        c.grid_blocks = (uint32_t)h->bgate_grid_used;
        all.push_back(c);
        knobs.push_back(k);
    };
    auto best_index = [&]() -> long {
        long bi = -1;
        for (size_t i = 0; i < all.size(); i++) {
            if (all[i].status) continue;
            if (bi < 0 || all[i].gate_ms + all[i].after_ms < all[bi].gate_ms + all[bi].after_ms) bi = (long)i;
        }
        return bi;
    };
    // stage 0: hit order (seed-major, k-mer-major) at the engine's knobs
    for (int order = 0; order < 2; order++) measure({saved.step, saved.threads, saved.claim, 0, order});
    long bi = best_index();
    const int order = bi >= 0 ? knobs[bi].order : saved.order;
    // stage 1: step x block x claim at the resident grid. Behind the coarse
    // bound (k-mer-major order; the stage-0 pass in that order tells) the step
    // does not apply (the candidates' gate has its own) and a wave's claim
    // sets how much work each claim's wait on its loads in flight covers, so
    // only blocks and the larger claims are swept.
    const bool coarse = order == 1 && h->coarse_used != 0;
    const std::vector<int> steps = coarse ? std::vector<int>{saved.step} : std::vector<int>{16, 32, 64, 96, 128};
    const std::vector<int> claims = coarse ? std::vector<int>{64, 256, 1024} : std::vector<int>{16, 64, 256};
    const int threads[] = {512, 1024};
    for (int st : steps)
        for (int th : threads)
            for (int cl : claims) measure({st, th, cl, 0, order});
    bi = best_index();
    // stage 2: grid sizes around the resident grid of the best cell
    if (bi >= 0) {
        TuneKnobs k = knobs[bi];
        const int resident = (int)all[bi].grid_blocks;
        const double factors[] = {0.5, 0.75, 1.5, 2.0};
        for (double f : factors) {
            k.blocks = std::max(1, (int)std::lround(resident * f));
            measure(k);
        }
        bi = best_index();
    }
    *n_cells = all.size();
    for (size_t i = 0; i < all.size() && i < cap; i++) cells[i] = all[i];
    if (bi < 0) {  // every cell failed; the last failure's message stands
        tune_set(h, saved);
        return all.back().status;
    }
    tune_set(h, knobs[bi]);
    if (best) *best = all[bi];
    return 0;
}
