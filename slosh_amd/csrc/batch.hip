// batch.hip -- many independent tanks stepped together (no reference counterpart: the reference drives one volume
// per proxy, SloshProxy.cpp:339-737).
//
// A batch holds existing contexts (tanks of uneven size). Each tank keeps its own HIP stream. Mode 0 runs every
// tank's whole step on its stream, tanks on `workers` host threads at once, so while one tank's host waits on a
// readback of its stream the GPU executes the kernels of the others (a tank of 32^3..128^3 fills only part of the
// 256 CUs on its own). Mode 1 (lockstep) runs the passes before and after the pressure solve so too, but the MGPCG
// iterations of all tanks together on the batch's own stream: each pass of an iteration (the level-0 column passes,
// the level-1 passes, the fused level-2/3 passes, the coarse tails, the finalizes) is one launch whose workgroups
// find their tank from per-tank prefix sums of workgroup counts, and one read of all tanks' states serves a chunk of
// iterations; a tank whose solve cannot join (another solver, a shallow hierarchy, residual capture) solves on its
// own stream. Tanks are assigned to workers longest-first by their last step time. The solver's iteration chunks
// replay as HIP graphs.
//
// The autotuner times candidate settings on real batch steps and keeps the fastest: mode, workers, iterations per
// solver chunk, graphs on/off, and the workgroup shape of the level-0 passes (bricks per workgroup). None of these
// change any arithmetic: every tank's step is the same kernels' arithmetic on the same data as a single-tank step,
// so batch results equal single-tank results bit for bit.
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
    // lockstep mode: the batch's stream, the tank descriptors and workgroup ranges (device and pinned host copies),
    // the gathered solver states and the chunk graph
    slosh::BatchRun run;
    slosh::BatchTank *h_tanks = nullptr;
    uint32_t *h_range = nullptr;
    slosh::SolverState *d_states = nullptr, *h_states = nullptr;
    hipEvent_t ev_done = nullptr;
    hipGraphExec_t graph = nullptr;
};

namespace {

using slosh::HipError;

// Runs fn(i) for every tank on `workers` host threads, tanks assigned longest-first by their last time; returns
// the first failure.
template <typename F> slosh_status for_tanks(slosh_batch *b, F fn)
{
    const size_t n = b->tanks.size();
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
            try {
                status[i] = fn(i);
            } catch (const HipError &e) {
                fprintf(stderr, "slosh: HIP error %d (%s) at %s:%d: %s\n", (int)e.err, hipGetErrorString(e.err), e.file,
                        e.line, e.what);
                status[i] = SLOSH_EDEVICE;
            } catch (const std::bad_alloc &) {
                status[i] = SLOSH_ENOMEM;
            }
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

slosh_status run_batch_step(slosh_batch *b, float dt, const float *gravity)
{
    for (slosh_ctx *c : b->tanks) {
        c->solver_chunk = (int)b->cfg.solver_chunk;
        c->use_graphs = b->cfg.graphs != 0;
        c->col_bricks = b->cfg.col_bricks;
    }
    if (b->cfg.mode == 0)
        return for_tanks(b, [&](size_t i) {
            const auto t0 = std::chrono::steady_clock::now();
            const slosh_status s = slosh_step(b->tanks[i], dt, gravity + 3 * i);
            b->last_ms[i] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            return s;
        });
    // lockstep: the passes before the solve (and the set-up of each MGPCG solve) per tank, the MGPCG iterations of
    // every tank together on the batch's stream, then the passes after the solve per tank
    const size_t n = b->tanks.size();
    std::vector<char> joins(n, 0), iterates(n, 0), solved(n, 0);
    std::vector<double> ms(n, 0.0);
    slosh_status st = for_tanks(b, [&](size_t i) {
        slosh_ctx *c = b->tanks[i];
        const auto t0 = std::chrono::steady_clock::now();
        slosh::tank_step_front(c, dt, gravity + 3 * i);
        joins[i] = slosh::mgpcg_batchable(c);
        if (joins[i]) iterates[i] = slosh::tank_solve_begin(c);
        else solved[i] = slosh::tank_solve(c);
        ms[i] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return SLOSH_OK;
    });
    if (st != SLOSH_OK) return st;
    std::vector<slosh_ctx *> L;
    for (size_t i = 0; i < n; ++i)
        if (joins[i] && iterates[i]) L.push_back(b->tanks[i]);
    try {
        if (!L.empty()) {
            slosh::BatchRun &B = b->run;
            B.ntanks = (uint32_t)L.size();
            B.col_bricks = b->cfg.col_bricks;
            const uint32_t stride = slosh::kMaxBatchTanks + 1u;
            for (int k = 0; k < slosh::kBatchRanges; ++k) b->h_range[(size_t)k * stride] = 0u;
            for (uint32_t t = 0; t < B.ntanks; ++t) {
                slosh_ctx *c = L[t];
                slosh::batch_tank_of(c, b->h_tanks[t]);
                const uint32_t per[slosh::kBatchRanges] = {(c->nsolve + B.col_bricks - 1u) / B.col_bricks, c->nsolve,
                                                           c->ntiles_f, 1u};
                for (int k = 0; k < slosh::kBatchRanges; ++k)
                    b->h_range[(size_t)k * stride + t + 1u] = b->h_range[(size_t)k * stride + t] + per[k];
                SLOSH_HIP(hipEventRecord(c->ev_batch, c->stream));
                SLOSH_HIP(hipStreamWaitEvent(B.stream, c->ev_batch, 0));
            }
            for (int k = 0; k < slosh::kBatchRanges; ++k) B.total[k] = b->h_range[(size_t)k * stride + B.ntanks];
            SLOSH_HIP(hipMemcpyAsync(B.d_tanks, b->h_tanks, B.ntanks * sizeof(slosh::BatchTank), hipMemcpyHostToDevice,
                                     B.stream));
            SLOSH_HIP(hipMemcpyAsync(B.d_range, b->h_range, slosh::kBatchRanges * stride * sizeof(uint32_t),
                                     hipMemcpyHostToDevice, B.stream));
            const auto t0 = std::chrono::steady_clock::now();
            slosh::batch_iterate(B, L.data(), B.ntanks, (int)b->cfg.solver_chunk, b->cfg.graphs != 0, b->graph,
                                 b->d_states, b->h_states);
            const double share =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / B.ntanks;
            SLOSH_HIP(hipEventRecord(b->ev_done, B.stream));
            for (slosh_ctx *c : L) SLOSH_HIP(hipStreamWaitEvent(c->stream, b->ev_done, 0));
            for (size_t i = 0; i < n; ++i)
                if (joins[i] && iterates[i]) ms[i] += share;
        }
    } catch (const HipError &e) {
        fprintf(stderr, "slosh: HIP error %d (%s) at %s:%d: %s\n", (int)e.err, hipGetErrorString(e.err), e.file, e.line,
                e.what);
        return SLOSH_EDEVICE;
    }
    st = for_tanks(b, [&](size_t i) {
        slosh_ctx *c = b->tanks[i];
        const auto t0 = std::chrono::steady_clock::now();
        bool s = solved[i] != 0;
        if (joins[i]) {
            slosh::tank_solve_end(c, iterates[i] != 0);
            s = true;
        }
        slosh::tank_step_back(c, s);
        b->last_ms[i] = ms[i] + std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return SLOSH_OK;
    });
    return st;
}

void free_lockstep(slosh_batch *b)
{
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    if (b->ev_done) (void)hipEventDestroy(b->ev_done);
    if (b->run.stream) (void)hipStreamDestroy(b->run.stream);
    if (b->run.d_tanks) (void)hipFree(b->run.d_tanks);
    if (b->run.d_range) (void)hipFree(b->run.d_range);
    if (b->d_states) (void)hipFree(b->d_states);
    if (b->h_tanks) (void)hipHostFree(b->h_tanks);
    if (b->h_range) (void)hipHostFree(b->h_range);
    if (b->h_states) (void)hipHostFree(b->h_states);
}

// The lockstep mode's stream and buffers (on the first tank's device); false when they cannot be had.
bool alloc_lockstep(slosh_batch *b)
{
    const uint32_t n = (uint32_t)b->tanks.size();
    if (n > slosh::kMaxBatchTanks) return false;
    for (slosh_ctx *c : b->tanks)
        if (c->device != b->tanks[0]->device) return false;
    if (hipSetDevice(b->tanks[0]->device) != hipSuccess) return false;
    const size_t range_words = slosh::kBatchRanges * (slosh::kMaxBatchTanks + 1u);
    bool ok = hipStreamCreate(&b->run.stream) == hipSuccess &&
              hipEventCreateWithFlags(&b->ev_done, hipEventDisableTiming) == hipSuccess &&
              hipMalloc((void **)&b->run.d_tanks, n * sizeof(slosh::BatchTank)) == hipSuccess &&
              hipMalloc((void **)&b->run.d_range, range_words * sizeof(uint32_t)) == hipSuccess &&
              hipMalloc((void **)&b->d_states, n * sizeof(slosh::SolverState)) == hipSuccess &&
              hipHostMalloc((void **)&b->h_tanks, n * sizeof(slosh::BatchTank), hipHostMallocDefault) == hipSuccess &&
              hipHostMalloc((void **)&b->h_range, range_words * sizeof(uint32_t), hipHostMallocDefault) == hipSuccess &&
              hipHostMalloc((void **)&b->h_states, n * sizeof(slosh::SolverState), hipHostMallocDefault) == hipSuccess;
    if (!ok) (void)hipGetLastError();
    return ok;
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
    b->cfg.mode = 0;
    b->cfg.col_bricks = 4;
    if (n <= slosh::kMaxBatchTanks && !alloc_lockstep(b)) {
        free_lockstep(b);
        delete b;
        return SLOSH_ENOMEM;
    }
    *out = b;
    return SLOSH_OK;
}

void slosh_batch_destroy(slosh_batch *b)
{
    if (!b) return;
    free_lockstep(b);
    delete b;
}

slosh_status slosh_batch_set_config(slosh_batch *b, const slosh_batch_config *cfg)
{
    if (!b || !cfg || cfg->workers == 0 || cfg->workers > 64 || cfg->solver_chunk < 2 || cfg->solver_chunk > 32 ||
        (cfg->solver_chunk & 1u) || cfg->mode > 1u ||
        (cfg->col_bricks != 1u && cfg->col_bricks != 2u && cfg->col_bricks != 4u))
        return SLOSH_EINVAL;
    if (cfg->mode == 1u && !b->run.stream) return SLOSH_ECAPACITY; // more tanks than a lockstep batch takes
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
    if (!b || !gravity || !(dt > 0.0f)) return SLOSH_EINVAL;
    return run_batch_step(b, dt, gravity);
}

slosh_status slosh_batch_autotune(slosh_batch *b, float dt, const float *gravity, uint32_t steps_per_candidate,
                                  slosh_batch_config *chosen)
{
    if (!b || !gravity || steps_per_candidate == 0 || !(dt > 0.0f)) return SLOSH_EINVAL;
    const uint32_t n = (uint32_t)b->tanks.size();
    std::vector<slosh_batch_config> cands;
    for (uint32_t w : {1u, 2u, 4u, 8u})
        for (uint32_t ch : {4u, 8u})
            for (uint32_t gr : {0u, 1u})
                if (w <= n) cands.push_back(slosh_batch_config{w, ch, gr, 0u, 4u});
    if (b->run.stream)
        for (uint32_t w : {2u, 4u, 8u})
            for (uint32_t ch : {4u, 8u})
                for (uint32_t cb : {1u, 2u, 4u})
                    if (w <= n) cands.push_back(slosh_batch_config{w, ch, 1u, 1u, cb});
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
