// batch.hip -- many independent tanks stepped together (no reference counterpart: the reference drives one volume
// per proxy, SloshProxy.cpp:339-737).
//
// A batch holds existing contexts (tanks of uneven size). Each tank keeps its own HIP stream; a step runs the tanks
// on `workers` host threads at once, so while one tank's host waits on a readback of its stream, the GPU executes
// the kernels of the others (a tank of 32^3..128^3 fills only part of the 256 CUs on its own). Tanks are assigned to
// workers longest-first by their last step time. The solver's iteration chunks replay as HIP graphs per tank.
//
// The autotuner times candidate (workers, iterations per solver chunk, graphs on/off) settings on real batch steps
// and keeps the fastest. None of these change any arithmetic: every tank's step is the same sequence of kernels on
// the same data as a single-tank step, so batch results equal single-tank results bit for bit.
#include <algorithm>
#include <chrono>
#include <new>
#include <thread>
#include <vector>

#include "slosh_ctx.h"

struct slosh_batch {
    std::vector<slosh_ctx *> tanks;
    std::vector<double> last_ms;
    slosh_batch_config cfg{};
};

namespace {

slosh_status run_batch_step(slosh_batch *b, float dt, const float *gravity)
{
    const size_t n = b->tanks.size();
    for (slosh_ctx *c : b->tanks) {
        c->solver_chunk = (int)b->cfg.solver_chunk;
        c->use_graphs = b->cfg.graphs != 0;
    }
    // longest-processing-time-first assignment of tanks to workers
    std::vector<size_t> order(n);
    for (size_t i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t c) { return b->last_ms[a] > b->last_ms[c]; });
    const uint32_t w = std::max(1u, std::min(b->cfg.workers, (uint32_t)n));
    std::vector<std::vector<size_t>> plan(w);
    std::vector<double> load(w, 0.0);
    for (size_t i : order) {
        const size_t k = (size_t)(std::min_element(load.begin(), load.end()) - load.begin());
        plan[k].push_back(i);
        load[k] += b->last_ms[i] > 0.0 ? b->last_ms[i] : 1.0;
    }
    std::vector<slosh_status> status(n, SLOSH_OK);
    auto work = [&](size_t k) {
        for (size_t i : plan[k]) {
            const auto t0 = std::chrono::steady_clock::now();
            status[i] = slosh_step(b->tanks[i], dt, gravity + 3 * i);
            b->last_ms[i] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
    };
    if (w == 1) {
        work(0);
    } else {
        std::vector<std::thread> threads;
        threads.reserve(w);
        for (size_t k = 0; k < w; ++k) threads.emplace_back(work, k);
        for (auto &t : threads) t.join();
    }
    for (slosh_status s : status)
        if (s != SLOSH_OK) return s;
    return SLOSH_OK;
}

} // namespace

extern "C" {

slosh_status slosh_batch_create(slosh_ctx *const *tanks, uint32_t n, slosh_batch **out)
{
    if (!tanks || !out || n == 0) return SLOSH_EINVAL;
    *out = nullptr;
    for (uint32_t i = 0; i < n; ++i) {
        if (!tanks[i]) return SLOSH_EINVAL;
        for (uint32_t j = 0; j < i; ++j)
            if (tanks[j] == tanks[i]) return SLOSH_EINVAL;
    }
    slosh_batch *b = new (std::nothrow) slosh_batch();
    if (!b) return SLOSH_ENOMEM;
    b->tanks.assign(tanks, tanks + n);
    b->last_ms.assign(n, 0.0);
    b->cfg.workers = std::min(4u, n);
    b->cfg.solver_chunk = 4;
    b->cfg.graphs = 1;
    *out = b;
    return SLOSH_OK;
}

void slosh_batch_destroy(slosh_batch *b) { delete b; }

slosh_status slosh_batch_set_config(slosh_batch *b, const slosh_batch_config *cfg)
{
    if (!b || !cfg || cfg->workers == 0 || cfg->workers > 64 || cfg->solver_chunk < 2 || cfg->solver_chunk > 32 ||
        (cfg->solver_chunk & 1u))
        return SLOSH_EINVAL;
    b->cfg = *cfg;
    return SLOSH_OK;
}

slosh_status slosh_batch_get_config(const slosh_batch *b, slosh_batch_config *cfg)
{
    if (!b || !cfg) return SLOSH_EINVAL;
    *cfg = b->cfg;
    return SLOSH_OK;
}

slosh_status slosh_batch_step(slosh_batch *b, float dt, const float *gravity)
{
    if (!b || !gravity) return SLOSH_EINVAL;
    return run_batch_step(b, dt, gravity);
}

slosh_status slosh_batch_autotune(slosh_batch *b, float dt, const float *gravity, uint32_t steps_per_candidate,
                                  slosh_batch_config *chosen)
{
    if (!b || !gravity || steps_per_candidate == 0) return SLOSH_EINVAL;
    const uint32_t n = (uint32_t)b->tanks.size();
    std::vector<slosh_batch_config> cands;
    for (uint32_t w : {1u, 2u, 4u, 8u})
        for (uint32_t ch : {4u, 8u})
            for (uint32_t gr : {0u, 1u})
                if (w <= n) cands.push_back(slosh_batch_config{w, ch, gr, 0u});
    double best = 1e300;
    slosh_batch_config best_cfg = b->cfg;
    for (const slosh_batch_config &cand : cands) {
        b->cfg = cand;
        const auto t0 = std::chrono::steady_clock::now();
        for (uint32_t s = 0; s < steps_per_candidate; ++s) {
            const slosh_status st = run_batch_step(b, dt, gravity);
            if (st != SLOSH_OK) return st;
        }
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (ms < best) {
            best = ms;
            best_cfg = cand;
        }
    }
    b->cfg = best_cfg;
    if (chosen) *chosen = best_cfg;
    return SLOSH_OK;
}

} // extern "C"
