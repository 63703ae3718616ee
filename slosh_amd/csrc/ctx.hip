// ctx.hip -- context lifecycle, step orchestration and the C ABI (include/slosh.h).
//
// Replaces FSloshProxy (SloshDemo/Plugins/Slosh/Source/SloshCore/Private/SloshProxy.cpp): Allocate (:185-306)
// becomes slosh_create, the emit passes of Step_RenderThread (:502-540) become slosh_emit, and one substep of the
// step loop (:564-713) becomes slosh_step. Device work is ordered on one HIP stream per context; the host reads
// back the allocated-brick count once per step (HIP has no indirect dispatch) and the solver state once per
// chunk of iterations.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <new>
#include <vector>

#include "slosh_ctx.h"

using namespace slosh;

namespace slosh {
const char *const kKernelNames[K_COUNT] = {
    "k_emit_fill", "k_sort_tiles", "k_sort_runs_scatter", "k_sort_key_runs", "k_sort_scatter", "k_scan",
    "k_pool", "k_p2g", "k_normalize", "k_divergence", "k_project", "k_div_stats", "k_extrapolate", "k_g2p_keys",
    "k_mg_setup", "k_coarse_factor", "k_pcg_init", "k_red_dot", "k_restrict0", "k_prolong_black", "k_direction",
    "k_update_black", "k_pdown_lds<4>", "k_pdown_lds<2>", "k_pup_lds<4>", "k_pup_lds<2>", "k_d23down", "k_d23up",
    "k_dsmooth", "k_drestrict", "k_dprolong", "k_coarse_solve", "k_coarse_tail", "k_sor_pair", "k_sor_residual",
    "k_fill", "k_finalize", "k_step_stats", "k_unet_conv_l0", "k_unet_conv_l1", "k_unet_conv_l2", "k_unet_io",
    "k_unet_train", "k_fcg", "k_g2p_move", "k_density", "k_viscosity", "k_coupling"};
} // namespace slosh

static_assert(sizeof(slosh_collider) == 80, "collider layout (SloshTypes.h:32-44)");
static_assert(sizeof(slosh_particle) == 64, "64-byte FP32 particle record");
static_assert(sizeof(slosh_particle_ref48) == 48, "FSloshParticleGPU layout (SloshTypes.h:84-91)");
static_assert(sizeof(slosh_settings) == 128, "settings ABI size");
static_assert(sizeof(slosh_manipulator) == 80, "manipulator layout (SloshTypes.h:66-80)");
static_assert(sizeof(slosh_emit_request) == 100, "emit request ABI size");
static_assert(sizeof(slosh_results) == 96, "results ABI size");
static_assert(sizeof(slosh_collider_feedback) == 64, "collider feedback ABI size");

namespace {

template <typename T> void dalloc(T **p, size_t count)
{
    size_t bytes = (count ? count : 1) * sizeof(T);
    if (hipMalloc((void **)p, bytes) != hipSuccess) throw std::bad_alloc();
    SLOSH_HIP(hipMemset(*p, 0, bytes));
}

template <typename T> void dfree(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

void free_all(slosh_ctx *c)
{
    for (int k = 0; k < 2; ++k) {
        dfree(c->parts[k].pa);
        for (int a = 0; a < 3; ++a) dfree(c->parts[k].v[a]);
        for (int q = 0; q < 9; ++q) dfree(c->parts[k].c[q]);
    }
    dfree(c->packed); dfree(c->tile_nruns); dfree(c->run_key); dfree(c->run_len); dfree(c->run_base);
    dfree(c->run_sorted); dfree(c->key_count); dfree(c->key_off); dfree(c->key_off_next); dfree(c->g2p_keys);
    dfree(c->key_runs); dfree(c->run_off);
    dfree(c->run_cursor); dfree(c->scan_blocks);
    dfree(c->table); dfree(c->slot_brick); dfree(c->active); dfree(c->need); dfree(c->need_rank);
    dfree(c->freeflag); dfree(c->free_rank); dfree(c->free_list); dfree(c->alloc_flag); dfree(c->alloc_rank);
    dfree(c->new_slot); dfree(c->touch);
    dfree(c->tiles); dfree(c->tile_valid);
    for (int k = 0; k < 3; ++k) dfree(c->vel[k]);
    dfree(c->fflags); dfree(c->fweight); dfree(c->ctype); dfree(c->pres); dfree(c->sor_tmp); dfree(c->div); dfree(c->diag); dfree(c->fmask);
    dfree(c->r); dfree(c->r2); dfree(c->z); dfree(c->z1); dfree(c->d); dfree(c->d2); dfree(c->q);
    dfree(c->binfo); dfree(c->binfo_f); dfree(c->fluid_flag); dfree(c->fluid_rank); dfree(c->tile_flag); dfree(c->tile_rank);
    dfree(c->tile_list); dfree(c->fcg_z); dfree(c->fcg_rp); dfree(c->capture);
    unet_free(c);
    for (int l = 1; l < kMaxLevels; ++l) {
        dfree(c->lv[l].type);
        if (c->lv[l].pooled) dfree(c->lv[l].info); // dense levels' words and coordinates live in the two blocks
        dfree(c->lv[l].x); dfree(c->lv[l].x1); dfree(c->lv[l].b);
    }
    dfree(c->dense_info); dfree(c->dense_coord);
    dfree(c->coarse_idx); dfree(c->coarse_n); dfree(c->coarse_inv); dfree(c->coarse_sig); dfree(c->pres_prev); dfree(c->pres_lp); dfree(c->state_save); dfree(c->partials); dfree(c->fold);
    dfree(c->dens_rhs); dfree(c->dens_q); dfree(c->dens_disp);
    dfree(c->bp_fluid); dfree(c->bp_div_max); dfree(c->bp_div_cnt); dfree(c->bp_div_sum); dfree(c->bp_speed);
    dfree(c->bp_clamp); dfree(c->bp_drain); dfree(c->d_man);
    dfree(c->state); dfree(c->stats); dfree(c->d_col); dfree(c->coupling);
    dfree(c->staging);
    for (uint32_t k = 0; k < c->tev_cap; ++k) (void)hipEventDestroy(c->tev[k]);
    delete[] c->tev;
    delete[] c->tev_kernel;
    c->tev = nullptr;
    c->tev_kernel = nullptr;
    c->tev_cap = 0;
    if (c->h_stats) (void)hipHostFree(c->h_stats);
    if (c->h_state) (void)hipHostFree(c->h_state);
    if (c->h_nsolve) (void)hipHostFree(c->h_nsolve);
    c->h_stats = nullptr;
    c->h_state = nullptr;
    if (c->ev_step0) (void)hipEventDestroy(c->ev_step0);
    if (c->ev_step1) (void)hipEventDestroy(c->ev_step1);
    if (c->ev_solve0) (void)hipEventDestroy(c->ev_solve0);
    if (c->ev_solve1) (void)hipEventDestroy(c->ev_solve1);
    if (c->ev_counts) (void)hipEventDestroy(c->ev_counts);
    if (c->ev_batch) (void)hipEventDestroy(c->ev_batch);
    if (c->ev_pool) (void)hipEventDestroy(c->ev_pool);
    if (c->mg_graph) (void)hipGraphExecDestroy(c->mg_graph);
    c->mg_graph = nullptr;
    if (c->fcg_graph) (void)hipGraphExecDestroy(c->fcg_graph);
    c->fcg_graph = nullptr;
    if (c->stream) (void)hipStreamDestroy(c->stream);
    c->stream = nullptr;
}

bool settings_valid(const slosh_settings *s)
{
    for (int a = 0; a < 3; ++a)
        if (s->grid_dim[a] < 4 || s->grid_dim[a] > 4096) return false;
    if (!(s->cell_size > 0.0f) || !(s->sparse_storage_budget > 0.0f) || s->sparse_storage_budget > 1.0f) return false;
    if (s->particle_capacity == 0u) return false;
    if (s->pressure_solver < SLOSH_SOLVER_MGPCG || s->pressure_solver > SLOSH_SOLVER_FCG_UNET) return false;
    if (s->pressure_solver == SLOSH_SOLVER_FCG_UNET)
        for (int a = 0; a < 3; ++a)
            if (s->grid_dim[a] % 8 != 0) return false; // the network's levels need whole bricks
    if (s->flags & ~(uint32_t)(SLOSH_FLAG_CAPTURE_RESIDUALS | SLOSH_FLAG_MG_FP16 | SLOSH_FLAG_MG_FLUID_COARSENING |
                               SLOSH_FLAG_WARM_START_EXTRAPOLATE | SLOSH_FLAG_REPORT_FP64_ERROR |
                               SLOSH_FLAG_CAPTURE_VCYCLE_RESIDUALS | SLOSH_FLAG_DENSITY_PROJECTION |
                               SLOSH_FLAG_BINARY_COUPLING))
        return false;
    if (!(s->tolerance > 0.0f) || s->max_iterations < 0) return false;
    if (!(s->sor_weight >= 0.5f && s->sor_weight <= 1.95f)) return false;
    if (s->extrapolation_layers < 0 || s->extrapolation_layers > 4) return false;
    if (!(s->boundary_friction >= 0.0f && s->boundary_friction <= 1.0f)) return false;
    if (s->mg_coarse_sweeps < 0 || s->mg_coarse_sweeps > 4) return false;
    if (!(s->density_relaxation >= 0.0f && s->density_relaxation <= 1.0f)) return false;
    if (s->density_projection_iterations < 0 || s->density_projection_iterations > 10000) return false;
    if (s->density_projection_interval < 0 || s->density_projection_interval > 8) return false;
    if (!(s->density_correction >= 0.0f) || !(s->density_deadband >= 0.0f)) return false;
    if (!(s->kinematic_viscosity >= 0.0f) || s->viscosity_iterations < 0 || s->viscosity_iterations > 1000) return false;
    if (!(s->fluid_density >= 0.0f && s->fluid_density <= 1e6f)) return false;
    for (int k = 0; k < 7; ++k)
        if (s->reserved[k] != 0) return false;
    return true;
}

void ensure_staging(slosh_ctx *c, size_t bytes)
{
    if (c->staging_bytes >= bytes) return;
    dfree(c->staging);
    c->staging_bytes = 0;
    if (hipMalloc(&c->staging, bytes) != hipSuccess) throw std::bad_alloc();
    c->staging_bytes = bytes;
}

void create_impl(slosh_ctx *c)
{
    const slosh_settings &s = c->s;
    SLOSH_HIP(hipGetDevice(&c->device));
    SLOSH_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    SLOSH_HIP(hipEventCreate(&c->ev_step0));
    SLOSH_HIP(hipEventCreate(&c->ev_step1));
    SLOSH_HIP(hipEventCreate(&c->ev_solve0));
    SLOSH_HIP(hipEventCreate(&c->ev_solve1));
    SLOSH_HIP(hipEventCreateWithFlags(&c->ev_counts, hipEventDisableTiming));
    SLOSH_HIP(hipEventCreateWithFlags(&c->ev_batch, hipEventDisableTiming));
    SLOSH_HIP(hipEventCreateWithFlags(&c->ev_pool, hipEventDisableTiming));
    Geo &g = c->geo;
    for (int a = 0; a < 3; ++a) {
        g.n[a] = s.grid_dim[a];
        g.bd[a] = (s.grid_dim[a] + 7) / 8;
    }
    g.nbricks = (uint32_t)g.bd[0] * (uint32_t)g.bd[1] * (uint32_t)g.bd[2];
    uint32_t slots = (uint32_t)ceil((double)s.sparse_storage_budget * (double)g.nbricks);
    if (slots < 1u) slots = 1u;
    if (slots > g.nbricks) slots = g.nbricks;
    g.npool = slots + 1u;
    const uint64_t pc64 = (uint64_t)g.npool * kBrickCells;
    // 32-bit indices over three face planes, and 32-bit byte offsets into the FP64 pool arrays (the solver's
    // column passes address them so): at most 2^20 - 1 pool slots
    if (pc64 * 3u > 0xFFFFFFFFull || pc64 * sizeof(double) > 0xFFFFFFFFull || g.nbricks >= (1u << 23))
        throw std::bad_alloc();
    g.pool_cells = (uint32_t)pc64;
    g.h = s.cell_size;
    g.inv_h = 1.0f / s.cell_size;
    c->nkeys = g.nbricks + 1u;
    c->cap = s.particle_capacity;
    c->ntiles_max = (c->cap + 255u) / 256u;
    const size_t cap = c->cap, runs = (size_t)c->ntiles_max * 256u, pc = g.pool_cells;
    for (int k = 0; k < 2; ++k) {
        dalloc(&c->parts[k].pa, cap);
        for (int a = 0; a < 3; ++a) dalloc(&c->parts[k].v[a], cap);
        for (int q = 0; q < 9; ++q) dalloc(&c->parts[k].c[q], cap);
    }
    dalloc(&c->packed, cap);
    dalloc(&c->tile_nruns, c->ntiles_max);
    dalloc(&c->run_key, runs);
    dalloc(&c->run_len, runs);
    dalloc(&c->run_base, runs);
    dalloc(&c->run_sorted, runs);
    dalloc(&c->key_count, c->nkeys);
    dalloc(&c->key_runs, c->nkeys);
    dalloc(&c->run_cursor, c->nkeys);
    dalloc(&c->key_off, (size_t)c->nkeys + 1u);
    dalloc(&c->key_off_next, (size_t)c->nkeys + 1u);
    dalloc(&c->g2p_keys, (size_t)c->cap);
    dalloc(&c->run_off, (size_t)c->nkeys + 1u);
    uint32_t maxn = c->nkeys > g.npool ? c->nkeys : g.npool;
    dalloc(&c->scan_blocks, (size_t)(maxn / 1024u + 2u));
    dalloc(&c->table, g.nbricks);
    dalloc(&c->slot_brick, g.npool);
    dalloc(&c->active, g.nbricks);
    dalloc(&c->need, g.nbricks);
    dalloc(&c->need_rank, (size_t)g.nbricks + 1u);
    dalloc(&c->freeflag, g.npool);
    dalloc(&c->free_rank, (size_t)g.npool + 1u);
    dalloc(&c->free_list, g.npool);
    dalloc(&c->alloc_flag, g.nbricks);
    dalloc(&c->alloc_rank, (size_t)g.nbricks + 1u);
    dalloc(&c->new_slot, g.npool);
    dalloc(&c->touch, g.nbricks);
    dalloc(&c->tiles, (size_t)g.npool * kTileInts);
    dalloc(&c->tile_valid, g.npool);
    for (int k = 0; k < 3; ++k) dalloc(&c->vel[k], 3 * pc);
    dalloc(&c->fflags, 3 * pc);
    dalloc(&c->fweight, 3 * pc);
    dalloc(&c->ctype, pc);
    dalloc(&c->pres, pc);
    if (s.pressure_solver == SLOSH_SOLVER_RBSOR) dalloc(&c->sor_tmp, pc);
    if (s.flags & SLOSH_FLAG_WARM_START_EXTRAPOLATE) dalloc(&c->pres_prev, pc);
    if (s.flags & SLOSH_FLAG_REPORT_FP64_ERROR) {
        dalloc(&c->pres_lp, pc);
        dalloc(&c->state_save, 1);
    }
    if (s.flags & SLOSH_FLAG_DENSITY_PROJECTION) {
        dalloc(&c->dens_rhs, pc);
        dalloc(&c->dens_q, pc);
        dalloc(&c->dens_disp, 3 * pc);
    }
    dalloc(&c->div, pc);
    dalloc(&c->diag, pc);
    dalloc(&c->fmask, pc);
    dalloc(&c->r, pc);
    dalloc(&c->r2, pc);
    dalloc(&c->z, pc);
    dalloc(&c->z1, pc);
    dalloc(&c->d, pc);
    dalloc(&c->d2, pc);
    dalloc(&c->q, pc);
    dalloc(&c->binfo, (size_t)g.npool * kBinfoWords);
    dalloc(&c->binfo_f, (size_t)g.npool * kBinfoWords);
    dalloc(&c->fluid_flag, g.npool);
    {
        const uint32_t t3 = (uint32_t)(((g.bd[0] + 3) / 4) * ((g.bd[1] + 3) / 4) * ((g.bd[2] + 3) / 4));
        c->ntiles = t3;
        dalloc(&c->tile_flag, t3);
        dalloc(&c->tile_rank, (size_t)t3 + 1u);
        dalloc(&c->tile_list, t3);
    }
    dalloc(&c->fluid_rank, (size_t)g.npool + 1u);
    if (s.pressure_solver == SLOSH_SOLVER_FCG_UNET) {
        dalloc(&c->fcg_z, pc);
        dalloc(&c->fcg_rp, pc);
    }
    if (s.flags & SLOSH_FLAG_CAPTURE_RESIDUALS) {
        c->capture_cap = 12;
        dalloc(&c->capture, (size_t)c->capture_cap * pc);
    }
    // MG hierarchy: halve (rounding up) until a level holds <= 64 cells or 8 levels exist. Levels 1..3 live in the
    // brick pool ((8 >> l)^3 cells per slot), coarser ones are dense. dalloc zeroes them: the null slot of a
    // pooled level must read as AIR with x = b = 0.
    c->nlevels = 1;
    c->lv[0].dim[0] = g.n[0];
    c->lv[0].dim[1] = g.n[1];
    c->lv[0].dim[2] = g.n[2];
    c->lv[0].cells = (uint32_t)g.n[0] * (uint32_t)g.n[1] * (uint32_t)g.n[2];
    while (c->lv[c->nlevels - 1].cells > kCoarsestCells && c->nlevels < kMaxLevels) {
        const int li = c->nlevels;
        Level &p = c->lv[li - 1];
        Level &l = c->lv[li];
        for (int a = 0; a < 3; ++a) l.dim[a] = (p.dim[a] + 1) / 2;
        l.cells = (uint32_t)l.dim[0] * (uint32_t)l.dim[1] * (uint32_t)l.dim[2];
        l.pooled = li <= 2;
        l.half = (s.flags & SLOSH_FLAG_MG_FP16) != 0;
        const size_t n = l.pooled ? (size_t)g.npool * (size_t)(512u >> (3 * li)) : (size_t)l.cells;
        const size_t esz = l.half ? 2u : 4u;
        dalloc(&l.type, n);
        if (l.pooled) dalloc(&l.info, n);
        dalloc((uint8_t **)&l.x, n * esz);
        dalloc((uint8_t **)&l.b, n * esz);
        dalloc((uint8_t **)&l.x1, n * esz);
        if (!l.pooled && (l.dim[0] > 1024 || l.dim[1] > 1024 || l.dim[2] > 1024)) throw std::bad_alloc();
        c->nlevels++;
    }
    if (c->lv[c->nlevels - 1].cells > kCoarsestCells) throw std::bad_alloc();
    // The dense levels' stencil words and cell coordinates are two contiguous blocks, level by level from the
    // first dense level, so the coarse tail reads all of its levels with one flat index.
    {
        size_t total = 0;
        for (int l = 1; l < c->nlevels; ++l)
            if (!c->lv[l].pooled) {
                c->lv[l].dense_off = (uint32_t)total;
                total += c->lv[l].cells;
            }
        dalloc(&c->dense_info, total);
        dalloc(&c->dense_coord, total);
        std::vector<uint32_t> coord(total ? total : 1);
        for (int li = 1; li < c->nlevels; ++li) {
            Level &l = c->lv[li];
            if (l.pooled) continue;
            l.info = c->dense_info + l.dense_off;
            l.coord = c->dense_coord + l.dense_off;
            for (uint32_t i = 0; i < l.cells; ++i) {
                const uint32_t x = i % (uint32_t)l.dim[0], y = (i / (uint32_t)l.dim[0]) % (uint32_t)l.dim[1],
                               z = i / (uint32_t)(l.dim[0] * l.dim[1]);
                coord[l.dense_off + i] = x | y << 10 | z << 20;
            }
        }
        if (total) SLOSH_HIP(hipMemcpy(c->dense_coord, coord.data(), total * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    dalloc(&c->coarse_idx, kCoarsestCells);
    dalloc(&c->coarse_n, 1);
    dalloc(&c->coarse_inv, (size_t)kCoarsestCells * kCoarsestCells);
    dalloc(&c->coarse_sig, (size_t)kCoarsestCells + 1u);
    dalloc(&c->partials, (size_t)kMaxBlockPartials * ((size_t)g.npool + 1u));
    dalloc(&c->fold, (size_t)kMaxBlockPartials * kFoldGroups);
    dalloc(&c->bp_fluid, g.npool);
    dalloc(&c->bp_div_max, g.npool);
    dalloc(&c->bp_div_cnt, g.npool);
    dalloc(&c->bp_div_sum, g.npool);
    dalloc(&c->bp_speed, g.nbricks);
    dalloc(&c->bp_clamp, g.nbricks);
    dalloc(&c->bp_drain, g.nbricks);
    dalloc(&c->state, 1);
    dalloc(&c->stats, 1);
    dalloc(&c->d_col, SLOSH_MAX_COLLIDERS);
    dalloc(&c->coupling, (size_t)SLOSH_MAX_COLLIDERS * 16u);
    dalloc(&c->d_man, SLOSH_MAX_MANIPULATORS);
    if (hipHostMalloc((void **)&c->h_stats, sizeof(StepStats), hipHostMallocDefault) != hipSuccess) throw std::bad_alloc();
    if (hipHostMalloc((void **)&c->h_nsolve, 2 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) throw std::bad_alloc();
    if (hipHostMalloc((void **)&c->h_state, sizeof(SolverState), hipHostMallocDefault) != hipSuccess) throw std::bad_alloc();
    memset(c->h_stats, 0, sizeof(StepStats));
    memset(c->h_state, 0, sizeof(SolverState));
    g.table = c->table;
    g.active = c->active;
    launch_fill_u32(c->stream, c->slot_brick, kNoBrick, g.npool);
    SLOSH_HIP(hipStreamSynchronize(c->stream));
}

// The submerged-cell pass of the last step, when it has not run yet (see step_back).
void flush_submerged(slosh_ctx *c)
{
    if (!c->submerged_pending) return;
    launch_submerged(c);
    c->submerged_pending = false;
}

void reset_impl(slosh_ctx *c)
{
    c->submerged_pending = false;
    const Geo &g = c->geo;
    c->np = 0;
    c->ndead = 0;
    c->sorted = false;
    c->cur = 0;
    c->vcur = 0;
    c->nactive = 0;
    c->frame = 0;
    c->prev_dt = 0.0f;
    c->drained_total = 0;
    c->coupling_ncol = 0;
    memset(&c->res, 0, sizeof(c->res));
    launch_fill_u32(c->stream, c->table, 0u, g.nbricks);
    launch_fill_u32(c->stream, c->slot_brick, kNoBrick, g.npool);
    launch_fill_u8(c->stream, c->touch, 0u, g.nbricks);
    launch_fill_u8(c->stream, c->new_slot, 0u, g.npool);
    launch_fill_f64(c->stream, c->pres, 0.0, g.pool_cells);
    if (c->sor_tmp) launch_fill_f64(c->stream, c->sor_tmp, 0.0, g.pool_cells);
    SLOSH_HIP(hipStreamSynchronize(c->stream));
}

// Folds the event pairs recorded since the last harvest into the per-kernel totals (stream must be idle).
void harvest_timing(slosh_ctx *c)
{
    for (uint32_t k = 0; k + 1 < c->tev_used; k += 2) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, c->tev[k], c->tev[k + 1]) == hipSuccess) {
            int id = c->tev_kernel[k / 2];
            c->ktime_ms[id] += (double)ms;
            c->klaunches[id] += 1u;
        }
    }
    c->tev_used = 0;
}

void read_stats(slosh_ctx *c)
{
    SLOSH_HIP(hipMemcpyAsync(c->h_stats, c->stats, sizeof(StepStats), hipMemcpyDeviceToHost, c->stream));
    SLOSH_HIP(hipStreamSynchronize(c->stream));
    harvest_timing(c);
}

template <typename T> std::vector<T> fetch(const slosh_ctx *c, const T *d, size_t n)
{
    std::vector<T> h(n);
    SLOSH_HIP(hipMemcpyAsync(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, c->stream));
    SLOSH_HIP(hipStreamSynchronize(c->stream));
    return h;
}

// The pressure solve of the current system with the configured solver (SloshProxy.cpp:618-641); the solver's final
// state is copied back asynchronously. Returns whether a solver ran.
bool run_solver(slosh_ctx *c)
{
    SLOSH_HIP(hipEventRecord(c->ev_solve0, c->stream));
    c->fcg_iters = 0;
    c->fcg_fallback = 0;
    if (c->s.pressure_solver == SLOSH_SOLVER_MGPCG) run_mgpcg(c);
    else if (c->s.pressure_solver == SLOSH_SOLVER_RBSOR) run_rbsor(c);
    else if (c->s.pressure_solver == SLOSH_SOLVER_FCG_UNET) run_fcg(c);
    SLOSH_HIP(hipEventRecord(c->ev_solve1, c->stream));
    c->res.fallback_used = c->fcg_fallback;
    // the solver's final state comes back with the step's statistics (no host wait between solve and G2P)
    const bool solved = c->s.pressure_solver != SLOSH_SOLVER_NONE;
    if (solved) SLOSH_HIP(hipMemcpyAsync(c->h_state, c->state, sizeof(SolverState), hipMemcpyDeviceToHost, c->stream));
    return solved;
}

// A lockstep batch's solve of this tank, before the batched iterations: the solver's set-up (run_solver's first
// part). Returns whether the solve needs iterations (false: no fluid).
bool batch_solve_begin(slosh_ctx *c)
{
    SLOSH_HIP(hipEventRecord(c->ev_solve0, c->stream));
    c->fcg_iters = 0;
    c->fcg_fallback = 0;
    return mgpcg_begin(c);
}

// After the batched iterations (the tank's stream waits for them): a restart when the FP64 true residual missed the
// tolerance, then run_solver's last part.
void batch_solve_end(slosh_ctx *c, bool iterated)
{
    if (iterated) mgpcg_resume(c);
    SLOSH_HIP(hipEventRecord(c->ev_solve1, c->stream));
    c->res.fallback_used = c->fcg_fallback;
    SLOSH_HIP(hipMemcpyAsync(c->h_state, c->state, sizeof(SolverState), hipMemcpyDeviceToHost, c->stream));
}

// The solver's figures in the results, once the stream has passed run_solver's copy of the state.
void solver_results(slosh_ctx *c)
{
    if (c->nactive) {
        // flexible CG: its own iterations, plus those of the MGPCG fallback when it ran
        const bool fcg = c->s.pressure_solver == SLOSH_SOLVER_FCG_UNET && c->fcg_fallback != 2;
        c->res.solver_iterations = fcg ? (c->fcg_fallback ? c->fcg_iters + c->h_state->iters : c->fcg_iters)
                                       : c->h_state->iters;
        c->res.residual_rel = c->h_state->rel;
        c->res.true_residual_rel = c->s.pressure_solver == SLOSH_SOLVER_RBSOR ? (float)c->h_state->rel
                                                                              : (float)c->h_state->rel_true;
        c->res.rhs_norm = sqrt(c->h_state->bb);
        c->res.converged = c->h_state->converged;
        if ((c->s.flags & SLOSH_FLAG_REPORT_FP64_ERROR) && c->s.pressure_solver != SLOSH_SOLVER_RBSOR)
            c->res.fp64_error_rel = c->h_state->err_rel;
    }
    float ms = 0.0f;
    SLOSH_HIP(hipEventElapsedTime(&ms, c->ev_solve0, c->ev_solve1));
    c->res.solve_ms = ms;
}

// Puts a pressure system given per cell in dense order into the context as P2G and the divergence pass would leave
// it: the bricks holding fluid are touched and the pool allocates them with their neighbourhood, then cell types,
// the operator diagonal, the divergence, the fluid-neighbour masks and the per-brick fluid counts are written.
// Returns false, with the context reset, when the pool cannot hold the bricks.
bool load_system_impl(slosh_ctx *c, const uint8_t *types, const float *diag, const float *div)
{
    reset_impl(c);
    const Geo &g = c->geo;
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    auto dense = [&](int x, int y, int z) { return (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * (size_t)z); };
    auto fluid_at = [&](int x, int y, int z) {
        return x >= 0 && y >= 0 && z >= 0 && x < nx && y < ny && z < nz && types[dense(x, y, z)] == kCellFluid;
    };
    std::vector<uint8_t> touch(g.nbricks, 0u);
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x)
                if (types[dense(x, y, z)] == kCellFluid)
                    touch[(size_t)((x >> 3) + g.bd[0] * ((y >> 3) + g.bd[1] * (z >> 3)))] = 1u;
    SLOSH_HIP(hipMemcpyAsync(c->touch, touch.data(), touch.size(), hipMemcpyHostToDevice, c->stream));
    run_pool(c);
    read_stats(c);
    if (c->h_stats->overflow) {
        reset_impl(c);
        return false;
    }
    c->nactive = c->h_stats->nactive;
    c->res.active_bricks = c->nactive;
    c->res.pool_overflow_bricks = 0;
    build_brick_rows(c, 0u, c->nactive);
    const std::vector<uint32_t> table = fetch(c, c->table, g.nbricks);
    const std::vector<uint32_t> active = fetch(c, c->active, c->nactive);
    std::vector<uint32_t> ctype(g.pool_cells, 0u), bp_fluid(c->nactive, 0u);
    std::vector<float> pdiag(g.pool_cells, 0.0f), pdiv(g.pool_cells, 0.0f);
    std::vector<uint8_t> fmask(g.pool_cells, 0u);
    for (uint32_t e = 0; e < c->nactive; ++e) {
        const uint32_t b = active[e], slot = table[b];
        const int bx = (int)(b % (uint32_t)g.bd[0]), by = (int)(b / (uint32_t)g.bd[0] % (uint32_t)g.bd[1]),
                  bz = (int)(b / (uint32_t)(g.bd[0] * g.bd[1]));
        for (uint32_t l = 0; l < kBrickCells; ++l) {
            const int x = bx * 8 + (int)(l & 7u), y = by * 8 + (int)((l >> 3) & 7u), z = bz * 8 + (int)(l >> 6);
            if (x >= nx || y >= ny || z >= nz) continue;
            const size_t i = (size_t)slot * kBrickCells + l, d = dense(x, y, z);
            ctype[i] = types[d];
            if (types[d] != kCellFluid) continue;
            // fluid-neighbour bits in the order -x +x -y +y -z +z, bit 6 the cell itself (the divergence pass's mask)
            uint32_t m = 1u << 6;
            const int off[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
            for (int n = 0; n < 6; ++n)
                if (fluid_at(x + off[n][0], y + off[n][1], z + off[n][2])) m |= 1u << n;
            fmask[i] = (uint8_t)m;
            pdiag[i] = diag[d];
            pdiv[i] = div[d];
            bp_fluid[e]++;
        }
    }
    SLOSH_HIP(hipMemcpyAsync(c->ctype, ctype.data(), ctype.size() * 4u, hipMemcpyHostToDevice, c->stream));
    SLOSH_HIP(hipMemcpyAsync(c->diag, pdiag.data(), pdiag.size() * 4u, hipMemcpyHostToDevice, c->stream));
    SLOSH_HIP(hipMemcpyAsync(c->div, pdiv.data(), pdiv.size() * 4u, hipMemcpyHostToDevice, c->stream));
    SLOSH_HIP(hipMemcpyAsync(c->fmask, fmask.data(), fmask.size(), hipMemcpyHostToDevice, c->stream));
    if (c->nactive)
        SLOSH_HIP(hipMemcpyAsync(c->bp_fluid, bp_fluid.data(), bp_fluid.size() * 4u, hipMemcpyHostToDevice, c->stream));
    SLOSH_HIP(hipStreamSynchronize(c->stream));
    uint32_t nf = 0;
    for (uint32_t v : bp_fluid) nf += v;
    c->res.fluid_cells = nf;
    return true;
}

// One substep in the order of SloshProxy.cpp:544-713, stopping after `stop` (SLOSH_STAGE_*): the passes before the
// pressure solve (step_front), the solve, the passes after it (step_back). A batch in lockstep mode runs the solves
// of its tanks between the two halves.
void step_front(slosh_ctx *c, float dt, const float gv[3], int stop)
{
    c->st_dt = dt;
    c->st_stop = stop;
    for (int a = 0; a < 3; ++a) c->st_g[a] = gv[a];
    c->frame++;
    const float warm = c->prev_dt > 0.0f ? dt / c->prev_dt : 1.0f;
    const float dt_prev = c->prev_dt;
    c->prev_dt = dt;
    c->res.sim_frame = c->frame;
    c->res.solver_iterations = 0;
    c->res.residual_rel = 0.0;
    c->res.true_residual_rel = 0.0f;
    c->res.fp64_error_rel = -1.0;
    c->res.converged = 1;
    c->res.solve_ms = 0.0f;
    launch_fill_u32(c->stream, (uint32_t *)c->stats, 0u, sizeof(StepStats) / 4u);
    c->n_p2g_blocks = 0;
    c->n_div_blocks = 0;
    c->n_g2p_blocks = 0;
    SLOSH_HIP(hipEventRecord(c->ev_step0, c->stream));
    if (!c->sorted) run_sort(c);
    c->np -= c->ndead;
    c->ndead = 0;
    c->res.live_particles = c->np;
    if (stop > SLOSH_STAGE_SORT) {
        run_pool(c);
        // The brick rows and P2G are launched over a bound of the new active-brick count (the last count plus an
        // eighth) and clip to the count on the device, so the host reads the count back while P2G runs; bricks
        // past the bound follow in a second launch. With kernel timing on, the host waits first (event pairs are
        // harvested on an idle stream).
        const uint32_t npool_max = c->geo.npool > 0u ? c->geo.npool - 1u : 0u;
        uint32_t bound = 0u;
        if (c->timing_mask == 0u) {
            const uint64_t b64 = (uint64_t)c->nactive + c->nactive / 8u + 64u;
            bound = (uint32_t)(b64 < npool_max ? b64 : npool_max);
            SLOSH_HIP(hipMemcpyAsync(c->h_stats, c->stats, sizeof(StepStats), hipMemcpyDeviceToHost, c->stream));
            SLOSH_HIP(hipEventRecord(c->ev_pool, c->stream));
            build_brick_rows(c, 0u, bound);
            if (stop > SLOSH_STAGE_POOL) launch_p2g(c, warm, dt, dt_prev, 0u, bound);
            SLOSH_HIP(hipEventSynchronize(c->ev_pool));
        } else {
            read_stats(c);
        }
        c->nactive = c->h_stats->nactive;
        c->res.active_bricks = c->nactive;
        c->res.pool_overflow_bricks = c->h_stats->overflow;
        if (c->nactive > bound) {
            build_brick_rows(c, bound, c->nactive - bound);
            if (stop > SLOSH_STAGE_POOL) launch_p2g(c, warm, dt, dt_prev, bound, c->nactive - bound);
        }
        c->n_p2g_blocks = stop > SLOSH_STAGE_POOL ? c->nactive : 0u;
    }
    if (stop > SLOSH_STAGE_P2G) {
        launch_normalize(c, dt, gv);
        // SloshProxy.cpp:603-617: implicit viscosity, alpha = nu dt / h^2
        const float alpha = c->s.kinematic_viscosity * dt / (c->geo.h * c->geo.h);
        if (alpha > 0.0f && c->s.viscosity_iterations > 0) run_viscosity(c, alpha, c->s.viscosity_iterations);
    }
    if (stop > SLOSH_STAGE_NORMALIZE) launch_divergence(c, dt);
}

void step_back(slosh_ctx *c, bool solved)
{
    const float dt = c->st_dt;
    const float *gv = c->st_g;
    const int stop = c->st_stop;
    if (stop > SLOSH_STAGE_SOLVE) launch_project(c);
    // collider coupling of this step (SloshProxy.cpp:397 clears it per frame; one step is one substep here): the
    // contact and wetted-face terms of the projection and the smoothed force (Slosh.usf:1138-1216) on this step's
    // cell types and solved pressure. The submerged cells (SloshProxy.cpp:715-718) are a once-per-frame pass over
    // the last substep's cell types in the reference: here it runs when the feedback is read (or before the
    // colliders change), on the same cell types, so steps whose feedback nobody reads do not pay for it.
    c->coupling_ncol = 0;
    c->submerged_pending = false;
    if (stop > SLOSH_STAGE_SOLVE && c->ncol) {
        launch_fill_u32(c->stream, (uint32_t *)c->coupling, 0u, (size_t)c->ncol * 16u);
        launch_coupling(c, dt);
        c->submerged_pending = true;
        c->coupling_ncol = c->ncol;
    }
    // SloshProxy.cpp:656-702: the two-sided density projection after the pressure projection, every
    // density_projection_interval steps with the relaxation compounded to 1 - (1 - r)^N
    c->dens_applied = false;
    if (stop > SLOSH_STAGE_SOLVE && (c->s.flags & SLOSH_FLAG_DENSITY_PROJECTION) && c->nactive) {
        const int interval = c->s.density_projection_interval > 0 ? c->s.density_projection_interval : 1;
        const float relax = 1.0f - powf(1.0f - c->s.density_relaxation, (float)interval);
        if (relax > 0.0f && c->frame % (uint64_t)interval == 0u) {
            const int keep = c->s.pressure_solver == SLOSH_SOLVER_NONE; // q as written by the caller
            launch_density_rhs(c, relax, keep);
            if (!keep) {
                // the solve reads its state back: keep the pressure solve's final state for the step's results
                SLOSH_HIP(hipStreamSynchronize(c->stream));
                const SolverState saved = *c->h_state;
                run_density_solve(c);
                *c->h_state = saved;
            }
            launch_density_disp(c);
            c->dens_applied = true;
        }
    }
    if (stop > SLOSH_STAGE_PROJECT)
        for (int k = 0; k < c->s.extrapolation_layers; ++k) launch_extrapolate(c, (uint32_t)k);
    if (stop > SLOSH_STAGE_EXTRAPOLATE) launch_g2p(c, dt, gv);
    launch_step_stats(c);
    SLOSH_HIP(hipEventRecord(c->ev_step1, c->stream));
    read_stats(c);
    if (solved) solver_results(c);
    const StepStats &st = *c->h_stats;
    if (stop > SLOSH_STAGE_EXTRAPOLATE && st.drained) {
        // drained records took the dead key: the sorted array's live records are the first np - drained
        c->np -= st.drained;
        c->drained_total += st.drained;
    }
    c->res.drained_particles = c->drained_total;
    c->res.live_particles = c->np;
    c->res.fluid_cells = st.fluid_cells;
    float f;
    memcpy(&f, &st.max_speed_bits, 4);
    c->res.max_particle_speed = f;
    memcpy(&f, &st.max_div_bits, 4);
    c->res.max_divergence = f;
    c->res.mean_divergence = st.div_count ? (float)(st.div_sum / (double)st.div_count) : 0.0f;
    c->res.clamped_particles = st.clamped;
    if (st.misplaced)
        throw HipError(hipErrorUnknown, "G2P passes disagree on particles' bricks (record order inconsistent)",
                       __FILE__, __LINE__);
    float ms = 0.0f;
    SLOSH_HIP(hipEventElapsedTime(&ms, c->ev_step0, c->ev_step1));
    c->res.step_ms = ms;
}

void step_impl(slosh_ctx *c, float dt, const float gv[3], int stop)
{
    step_front(c, dt, gv, stop);
    bool solved = false;
    if (stop > SLOSH_STAGE_DIVERGENCE) solved = run_solver(c);
    step_back(c, solved);
}

// Pool -> dense conversion on the host (x fastest; unallocated cells read 0).
template <typename Src, typename Dst>
void pool_to_dense(const slosh_ctx *c, const std::vector<uint32_t> &table, const Src *pool, Dst *dense, int planes)
{
    const Geo &g = c->geo;
    const size_t cells = (size_t)g.n[0] * g.n[1] * g.n[2];
    for (int pl = 0; pl < planes; ++pl)
        for (int z = 0; z < g.n[2]; ++z)
            for (int y = 0; y < g.n[1]; ++y)
                for (int x = 0; x < g.n[0]; ++x) {
                    uint32_t b = (uint32_t)((x >> 3) + g.bd[0] * ((y >> 3) + g.bd[1] * (z >> 3)));
                    uint32_t slot = table[b];
                    size_t di = (size_t)pl * cells + (size_t)x + (size_t)g.n[0] * ((size_t)y + (size_t)g.n[1] * z);
                    if (slot == 0u) {
                        dense[di] = (Dst)0;
                        continue;
                    }
                    uint32_t l = (uint32_t)((x & 7) + 8 * ((y & 7) + 8 * (z & 7)));
                    dense[di] = (Dst)pool[(size_t)pl * g.pool_cells + (size_t)slot * kBrickCells + l];
                }
}

slosh_status read_grid_impl(slosh_ctx *c, int32_t field, void *dst, uint64_t bytes)
{
    const Geo &g = c->geo;
    const size_t cells = (size_t)g.n[0] * g.n[1] * g.n[2];
    std::vector<uint32_t> table = fetch(c, c->table, g.nbricks);
    if (field == SLOSH_FIELD_BRICK_TABLE) {
        if (bytes < (uint64_t)g.nbricks * 4u) return SLOSH_ECAPACITY;
        memcpy(dst, table.data(), (size_t)g.nbricks * 4u);
        return SLOSH_OK;
    }
    int planes = 1;
    if (field == SLOSH_FIELD_VEL || field == SLOSH_FIELD_FACE_FLAGS || field == SLOSH_FIELD_FACE_WEIGHT) planes = 3;
    if (field == SLOSH_FIELD_ACCUM) planes = 6;
    if (field == SLOSH_FIELD_PRESSURE_F64) {
        if (bytes < (uint64_t)cells * 8u) return SLOSH_ECAPACITY;
        auto h = fetch(c, c->pres, g.pool_cells);
        pool_to_dense(c, table, h.data(), (double *)dst, 1);
        return SLOSH_OK;
    }
    if (field == SLOSH_FIELD_DENSITY_Q || field == SLOSH_FIELD_DENSITY_RHS || field == SLOSH_FIELD_DENSITY_DISP) {
        if (!c->dens_q) return SLOSH_ESTATE; // the settings do not enable the density projection
        const int pl = field == SLOSH_FIELD_DENSITY_DISP ? 3 : 1;
        if (bytes < (uint64_t)cells * 4u * (uint64_t)pl) return SLOSH_ECAPACITY;
        if (field == SLOSH_FIELD_DENSITY_Q) {
            auto h = fetch(c, c->dens_q, g.pool_cells);
            pool_to_dense(c, table, h.data(), (float *)dst, 1);
        } else if (field == SLOSH_FIELD_DENSITY_RHS) {
            auto h = fetch(c, c->dens_rhs, g.pool_cells);
            pool_to_dense(c, table, h.data(), (float *)dst, 1);
        } else {
            auto h = fetch(c, c->dens_disp, 3 * (size_t)g.pool_cells);
            pool_to_dense(c, table, h.data(), (float *)dst, 3);
        }
        return SLOSH_OK;
    }
    if (field < 0 || field > SLOSH_FIELD_ACCUM) return SLOSH_EINVAL;
    if (bytes < (uint64_t)cells * 4u * (uint64_t)planes) return SLOSH_ECAPACITY;
    const size_t pc = g.pool_cells;
    switch (field) {
    case SLOSH_FIELD_CELL_FLAGS: {
        auto h = fetch(c, c->ctype, pc);
        pool_to_dense(c, table, h.data(), (uint32_t *)dst, 1);
        break;
    }
    case SLOSH_FIELD_PRESSURE: {
        auto h = fetch(c, c->pres, pc);
        pool_to_dense(c, table, h.data(), (float *)dst, 1);
        break;
    }
    case SLOSH_FIELD_DIVERGENCE: {
        auto h = fetch(c, c->div, pc);
        pool_to_dense(c, table, h.data(), (float *)dst, 1);
        break;
    }
    case SLOSH_FIELD_CELL_DIAG: {
        auto h = fetch(c, c->diag, pc);
        pool_to_dense(c, table, h.data(), (float *)dst, 1);
        break;
    }
    case SLOSH_FIELD_VEL: {
        auto h = fetch(c, c->vel[c->vcur], 3 * pc);
        pool_to_dense(c, table, h.data(), (float *)dst, 3);
        break;
    }
    case SLOSH_FIELD_FACE_FLAGS: {
        auto h = fetch(c, c->fflags, 3 * pc);
        pool_to_dense(c, table, h.data(), (uint32_t *)dst, 3);
        break;
    }
    case SLOSH_FIELD_FACE_WEIGHT: {
        auto h = fetch(c, c->fweight, 3 * pc);
        pool_to_dense(c, table, h.data(), (float *)dst, 3);
        break;
    }
    case SLOSH_FIELD_ACCUM: {
        // P2G sums: the overlapping brick tiles of each face, summed as the normalise pass does.
        auto tiles = fetch(c, c->tiles, (size_t)g.npool * kTileInts);
        auto valid = fetch(c, c->tile_valid, g.npool);
        int32_t *out = (int32_t *)dst;
        memset(out, 0, cells * 4u * 6u);
        for (int z = 0; z < g.n[2]; ++z)
            for (int y = 0; y < g.n[1]; ++y)
                for (int x = 0; x < g.n[0]; ++x) {
                    int p[3] = {x, y, z};
                    uint32_t b = (uint32_t)((x >> 3) + g.bd[0] * ((y >> 3) + g.bd[1] * (z >> 3)));
                    if (table[b] == 0u) continue;
                    size_t di = (size_t)x + (size_t)g.n[0] * ((size_t)y + (size_t)g.n[1] * z);
                    for (int dz = -1; dz <= 1; ++dz)
                        for (int dy = -1; dy <= 1; ++dy)
                            for (int dx = -1; dx <= 1; ++dx) {
                                int nb[3] = {(x >> 3) + dx, (y >> 3) + dy, (z >> 3) + dz};
                                bool ok = true;
                                int tc[3];
                                for (int a = 0; a < 3; ++a) {
                                    if (nb[a] < 0 || nb[a] >= g.bd[a]) ok = false;
                                    tc[a] = p[a] - (nb[a] * 8 - 1);
                                    if (tc[a] < 0 || tc[a] >= kTileDim) ok = false;
                                }
                                if (!ok) continue;
                                uint32_t ns = table[(uint32_t)(nb[0] + g.bd[0] * (nb[1] + g.bd[1] * nb[2]))];
                                if (ns == 0u || !valid[ns]) continue;
                                const int32_t *t = tiles.data() + (size_t)ns * kTileInts +
                                                   (tc[0] + kTileDim * (tc[1] + kTileDim * tc[2]));
                                for (int q = 0; q < 6; ++q) out[(size_t)q * cells + di] += t[q * kTileNodes];
                            }
                }
        break;
    }
    default:
        return SLOSH_EINVAL;
    }
    return SLOSH_OK;
}

thread_local char g_info[256];

// ---- staging of the allocation-free particle readback (slosh_read_particles_noalloc) ----
// One library-wide device buffer per GPU, SLOSH_NOALLOC_STAGING_BYTES long, held from the module's load on that GPU
// until the process ends. Host code never passes `scratch` itself (in host code the name is only a stand-in for
// the symbol): each device's address comes from hipGetSymbolAddress, taken once per device in slosh_create with
// that device current, and only that address reaches kernels and copies.
static __device__ __align__(256) uint8_t scratch[SLOSH_NOALLOC_STAGING_BYTES];
constexpr int kMaxDevices = 64;
void *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mutex;

// Looks up (once) this device's address of the staging buffer; the caller has made `device` current.
hipError_t scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device]) return hipSuccess;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = ptr;
    return hipSuccess;
}

void *scratch_of(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    return g_scratch_ptr[device];
}

} // namespace

namespace slosh {
void tank_step_front(slosh_ctx *c, float dt, const float gv[3]) { ::step_front(c, dt, gv, SLOSH_STAGE_ALL); }
void tank_step_back(slosh_ctx *c, bool solved) { ::step_back(c, solved); }
bool tank_solve(slosh_ctx *c) { return ::run_solver(c); }
bool tank_solve_begin(slosh_ctx *c) { return ::batch_solve_begin(c); }
void tank_solve_end(slosh_ctx *c, bool iterated) { ::batch_solve_end(c, iterated); }
} // namespace slosh

#define ABI_GUARD(...)                                                                                           \
    try {                                                                                                        \
        __VA_ARGS__                                                                                              \
    } catch (const slosh::HipError &e) {                                                                        \
        fprintf(stderr, "slosh: HIP error %d (%s) at %s:%d: %s\n", (int)e.err, hipGetErrorString(e.err), e.file, \
                e.line, e.what);                                                                                 \
        return SLOSH_EDEVICE;                                                                                    \
    } catch (const std::bad_alloc &) {                                                                          \
        return SLOSH_ENOMEM;                                                                                     \
    }

extern "C" {

uint32_t slosh_abi_version(void) { return SLOSH_ABI_VERSION; }

const char *slosh_build_info(void)
{
    snprintf(g_info, sizeof(g_info), "slosh_amd abi %u; offload gfx950; HIP %d.%d; -ffp-contract=off", SLOSH_ABI_VERSION,
             HIP_VERSION_MAJOR, HIP_VERSION_MINOR);
    return g_info;
}

const char *slosh_status_string(slosh_status s)
{
    switch (s) {
    case SLOSH_OK: return "ok";
    case SLOSH_EINVAL: return "invalid argument";
    case SLOSH_ENOMEM: return "out of memory";
    case SLOSH_EDEVICE: return "device error";
    case SLOSH_ECAPACITY: return "capacity exceeded";
    case SLOSH_ESTATE: return "invalid state";
    }
    return "unknown";
}

void slosh_default_settings(slosh_settings *s)
{
    if (!s) return;
    memset(s, 0, sizeof(*s));
    s->grid_dim[0] = 64;
    s->grid_dim[1] = 64;
    s->grid_dim[2] = 64;
    s->cell_size = 1.0f / 64.0f;
    s->particle_capacity = 1u << 20;
    s->sparse_storage_budget = 1.0f;
    s->pressure_solver = SLOSH_SOLVER_MGPCG;
    s->tolerance = 1e-6f;
    s->max_iterations = 200;
    s->sor_weight = 1.8f;
    s->ghost_fluid = 1;
    s->extrapolation_layers = 2;
    s->boundary_friction = 0.0f;
    s->max_speed = 20.0f;
    s->particle_radius_cells = 0.25f;
    s->density_correction = 0.0f;
    s->density_deadband = 0.05f;
    s->density_relaxation = 0.5f;
    s->density_projection_iterations = 30;
    s->density_projection_interval = 1;
    s->kinematic_viscosity = 0.0f;
    s->viscosity_iterations = 12;
    s->fluid_density = 1000.0f;
}

slosh_status slosh_create(const slosh_settings *s, slosh_ctx **out)
{
    if (!s || !out) return SLOSH_EINVAL;
    *out = nullptr;
    if (!settings_valid(s)) return SLOSH_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return SLOSH_EDEVICE;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return SLOSH_EDEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return SLOSH_EDEVICE;
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        fprintf(stderr, "slosh: device %d is %s; the kernels are built for gfx950\n", dev, prop.gcnArchName);
        return SLOSH_EDEVICE;
    }
    {
        const hipError_t e = scratch_init(dev);
        if (e != hipSuccess) {
            fprintf(stderr, "slosh: staging buffer address on device %d: %s\n", dev, hipGetErrorString(e));
            return SLOSH_EDEVICE;
        }
    }
    slosh_ctx *c = new (std::nothrow) slosh_ctx();
    if (!c) return SLOSH_ENOMEM;
    c->s = *s;
    try {
        create_impl(c);
    } catch (const slosh::HipError &e) {
        fprintf(stderr, "slosh: HIP error %d (%s) at %s:%d\n", (int)e.err, hipGetErrorString(e.err), e.file, e.line);
        free_all(c);
        delete c;
        return SLOSH_EDEVICE;
    } catch (const std::bad_alloc &) {
        free_all(c);
        delete c;
        return SLOSH_ENOMEM;
    }
    *out = c;
    return SLOSH_OK;
}

void slosh_destroy(slosh_ctx *c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_all(c);
    delete c;
}

slosh_status slosh_reset(slosh_ctx *c)
{
    if (!c) return SLOSH_EINVAL;
    ABI_GUARD(reset_impl(c); return SLOSH_OK;)
}

slosh_status slosh_emit(slosh_ctx *c, const slosh_emit_request *r)
{
    if (!c || !r) return SLOSH_EINVAL;
    if (r->shape > SLOSH_EMIT_BOX || r->reserved != 0u) return SLOSH_EINVAL;
    if (!(r->velocity_spread >= 0.0f) || !(r->speed_jitter >= 0.0f)) return SLOSH_EINVAL;
    uint64_t count = 0;
    if (r->shape == SLOSH_EMIT_FILL_CELLS) {
        for (int a = 0; a < 3; ++a)
            if (r->cell_dims[a] < 0) return SLOSH_EINVAL;
        count = (uint64_t)r->cell_dims[0] * (uint64_t)r->cell_dims[1] * (uint64_t)r->cell_dims[2] * 8u;
    } else {
        for (int a = 0; a < 3; ++a)
            if (!(r->extent[a] >= 0.0f) || !std::isfinite(r->center[a])) return SLOSH_EINVAL;
        count = r->count;
    }
    if (count == 0) return SLOSH_OK;
    if ((uint64_t)c->np + count > (uint64_t)c->cap) return SLOSH_ECAPACITY;
    ABI_GUARD(launch_emit(c, r, (uint32_t)count); c->np += (uint32_t)count; c->sorted = false;
              SLOSH_HIP(hipStreamSynchronize(c->stream)); return SLOSH_OK;)
}

slosh_status slosh_set_colliders(slosh_ctx *c, const slosh_collider *col, uint32_t n)
{
    if (!c || (n && !col)) return SLOSH_EINVAL;
    if (n > SLOSH_MAX_COLLIDERS) return SLOSH_ECAPACITY;
    ABI_GUARD(flush_submerged(c); if (n) {
        memcpy(c->h_col, col, n * sizeof(slosh_collider));
        SLOSH_HIP(hipMemcpyAsync(c->d_col, c->h_col, n * sizeof(slosh_collider), hipMemcpyHostToDevice, c->stream));
        SLOSH_HIP(hipStreamSynchronize(c->stream));
    } c->ncol = n;
              return SLOSH_OK;)
}

slosh_status slosh_set_manipulators(slosh_ctx *c, const slosh_manipulator *m, uint32_t n)
{
    if (!c || (n && !m)) return SLOSH_EINVAL;
    if (n > SLOSH_MAX_MANIPULATORS) return SLOSH_ECAPACITY;
    for (uint32_t k = 0; k < n; ++k)
        if (m[k].kind > 4u || m[k].shape > 3u || !(m[k].falloff >= 0.0f) || !(m[k].param0 >= 0.0f || m[k].kind == 3u))
            return SLOSH_EINVAL;
    ABI_GUARD(if (n) {
        SLOSH_HIP(hipMemcpyAsync(c->d_man, m, n * sizeof(slosh_manipulator), hipMemcpyHostToDevice, c->stream));
        SLOSH_HIP(hipStreamSynchronize(c->stream));
    } c->nman = n;
              return SLOSH_OK;)
}

slosh_status slosh_step_until(slosh_ctx *c, float dt, const float gravity[3], int32_t stop_after)
{
    if (!c || !gravity || !(dt > 0.0f) || stop_after < SLOSH_STAGE_SORT || stop_after > SLOSH_STAGE_ALL)
        return SLOSH_EINVAL;
    ABI_GUARD(step_impl(c, dt, gravity, stop_after); return SLOSH_OK;)
}

slosh_status slosh_step(slosh_ctx *c, float dt, const float gravity[3])
{
    return slosh_step_until(c, dt, gravity, SLOSH_STAGE_ALL);
}

slosh_status slosh_read_particles(slosh_ctx *c, void *dst, uint64_t dst_bytes, uint32_t *count, int32_t layout)
{
    if (!c || !count || (layout != SLOSH_LAYOUT_F32_64 && layout != SLOSH_LAYOUT_REF_48)) return SLOSH_EINVAL;
    const size_t rec = layout == SLOSH_LAYOUT_F32_64 ? 64u : 48u;
    *count = 0;
    if ((uint64_t)c->np * rec > dst_bytes) return SLOSH_ECAPACITY;
    if (c->np == 0) return SLOSH_OK;
    if (!dst) return SLOSH_EINVAL;
    ABI_GUARD(ensure_staging(c, (size_t)c->np * rec); launch_pack_particles(c, c->staging, layout);
              SLOSH_HIP(hipMemcpyAsync(dst, c->staging, (size_t)c->np * rec, hipMemcpyDeviceToHost, c->stream));
              SLOSH_HIP(hipStreamSynchronize(c->stream)); *count = c->np; return SLOSH_OK;)
}

slosh_status slosh_read_particles_noalloc(slosh_ctx *c, void *dst, uint64_t dst_bytes, uint32_t *count,
                                          int32_t layout)
{
    if (!c || !count || (layout != SLOSH_LAYOUT_F32_64 && layout != SLOSH_LAYOUT_REF_48)) return SLOSH_EINVAL;
    const size_t rec = layout == SLOSH_LAYOUT_F32_64 ? 64u : 48u;
    *count = 0;
    const uint64_t bytes = (uint64_t)c->np * rec;
    if (bytes > (uint64_t)SLOSH_NOALLOC_STAGING_BYTES) {
        fprintf(stderr, "slosh: slosh_read_particles_noalloc: %u particles of %zu bytes exceed the %u-byte staging "
                        "buffer\n", c->np, rec, (unsigned)SLOSH_NOALLOC_STAGING_BYTES);
        return SLOSH_ECAPACITY;
    }
    if (bytes > dst_bytes) return SLOSH_ECAPACITY;
    if (c->np == 0) return SLOSH_OK;
    if (!dst) return SLOSH_EINVAL;
    void *stage = scratch_of(c->device);
    if (!stage) return SLOSH_ESTATE; // slosh_create looked the address up on this context's device
    ABI_GUARD(launch_pack_particles(c, stage, layout);
              SLOSH_HIP(hipMemcpyAsync(dst, stage, (size_t)bytes, hipMemcpyDeviceToHost, c->stream));
              SLOSH_HIP(hipStreamSynchronize(c->stream)); *count = c->np; return SLOSH_OK;)
}

slosh_status slosh_device_particles(slosh_ctx *c, int32_t layout, void **dev_ptr, uint32_t *count)
{
    if (!c || !dev_ptr || !count || (layout != SLOSH_LAYOUT_F32_64 && layout != SLOSH_LAYOUT_REF_48)) return SLOSH_EINVAL;
    const size_t rec = layout == SLOSH_LAYOUT_F32_64 ? 64u : 48u;
    ABI_GUARD(ensure_staging(c, (size_t)(c->np ? c->np : 1u) * rec); launch_pack_particles(c, c->staging, layout);
              SLOSH_HIP(hipStreamSynchronize(c->stream)); *dev_ptr = c->staging; *count = c->np; return SLOSH_OK;)
}

slosh_status slosh_write_particles(slosh_ctx *c, const void *src, uint32_t n)
{
    if (!c || (n && !src)) return SLOSH_EINVAL;
    if (n > c->cap) return SLOSH_ECAPACITY;
    ABI_GUARD(launch_fill_u8(c->stream, c->touch, 0u, c->geo.nbricks); c->np = 0; c->ndead = 0; c->sorted = false;
              if (n) {
        ensure_staging(c, (size_t)n * sizeof(slosh_particle));
        SLOSH_HIP(hipMemcpyAsync(c->staging, src, (size_t)n * sizeof(slosh_particle), hipMemcpyHostToDevice, c->stream));
        launch_unpack_particles(c, (const slosh_particle *)c->staging, n);
        c->np = n;
        const slosh_particle *sp = (const slosh_particle *)src;
        for (uint32_t i = 0; i < n; ++i) c->ndead += (sp[i].age >= 0.0f) ? 0u : 1u;
    } SLOSH_HIP(hipStreamSynchronize(c->stream));
              return SLOSH_OK;)
}

slosh_status slosh_read_grid(slosh_ctx *c, int32_t field, void *dst, uint64_t dst_bytes, int32_t order)
{
    if (!c || !dst || order != SLOSH_ORDER_DENSE) return SLOSH_EINVAL;
    ABI_GUARD(return read_grid_impl(c, field, dst, dst_bytes);)
}

slosh_status slosh_write_grid(slosh_ctx *c, int32_t field, const void *src, uint64_t src_bytes, int32_t order)
{
    if (!c || !src || order != SLOSH_ORDER_DENSE) return SLOSH_EINVAL;
    if (field != SLOSH_FIELD_PRESSURE && field != SLOSH_FIELD_DENSITY_Q) return SLOSH_EINVAL;
    if (field == SLOSH_FIELD_DENSITY_Q && !c->dens_q) return SLOSH_ESTATE;
    double *const target = field == SLOSH_FIELD_PRESSURE ? c->pres : c->dens_q;
    const Geo &g = c->geo;
    const size_t cells = (size_t)g.n[0] * g.n[1] * g.n[2];
    if (src_bytes < cells * 4u) return SLOSH_ECAPACITY;
    ABI_GUARD({
        std::vector<uint32_t> table = fetch(c, c->table, g.nbricks);
        std::vector<double> pool = fetch(c, target, g.pool_cells);
        const float *in = (const float *)src;
        for (int z = 0; z < g.n[2]; ++z)
            for (int y = 0; y < g.n[1]; ++y)
                for (int x = 0; x < g.n[0]; ++x) {
                    uint32_t slot = table[(uint32_t)((x >> 3) + g.bd[0] * ((y >> 3) + g.bd[1] * (z >> 3)))];
                    if (!slot) continue;
                    pool[(size_t)slot * kBrickCells + (size_t)((x & 7) + 8 * ((y & 7) + 8 * (z & 7)))] =
                        (double)in[(size_t)x + (size_t)g.n[0] * ((size_t)y + (size_t)g.n[1] * z)];
                }
        SLOSH_HIP(hipMemcpyAsync(target, pool.data(), pool.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
        SLOSH_HIP(hipStreamSynchronize(c->stream));
        return SLOSH_OK;
    })
}

slosh_status slosh_load_system(slosh_ctx *c, const uint8_t *types, const float *diag, const float *divergence)
{
    if (!c || !types || !diag || !divergence) return SLOSH_EINVAL;
    const Geo &g = c->geo;
    const size_t cells = (size_t)g.n[0] * g.n[1] * g.n[2];
    for (size_t i = 0; i < cells; ++i)
        if (types[i] > kCellSolid || (types[i] == kCellFluid && !(diag[i] >= 0.0f && diag[i] <= 1e6f)) ||
            (types[i] == kCellFluid && !(fabsf(divergence[i]) <= 3.0e38f)))
            return SLOSH_EINVAL;
    ABI_GUARD(return load_system_impl(c, types, diag, divergence) ? SLOSH_OK : SLOSH_ECAPACITY;)
}

slosh_status slosh_solve(slosh_ctx *c)
{
    if (!c) return SLOSH_EINVAL;
    ABI_GUARD({
        c->res.solver_iterations = 0;
        c->res.residual_rel = 0.0;
        c->res.true_residual_rel = 0.0f;
        c->res.fp64_error_rel = -1.0;
        c->res.converged = 1;
        if (run_solver(c)) {
            SLOSH_HIP(hipStreamSynchronize(c->stream));
            solver_results(c);
        }
        SLOSH_HIP(hipStreamSynchronize(c->stream));
        harvest_timing(c);
        return SLOSH_OK;
    })
}

slosh_status slosh_set_kernel_timing(slosh_ctx *c, uint64_t mask)
{
    if (!c) return SLOSH_EINVAL;
    ABI_GUARD({
        if (mask && !c->tev) {
            const uint32_t cap = 8192;
            c->tev = new hipEvent_t[cap];
            c->tev_kernel = new int[cap / 2];
            for (uint32_t k = 0; k < cap; ++k) {
                SLOSH_HIP(hipEventCreate(&c->tev[k]));
                c->tev_cap = k + 1;
            }
        }
        SLOSH_HIP(hipStreamSynchronize(c->stream));
        harvest_timing(c);
        c->timing_mask = mask;
        return SLOSH_OK;
    })
}

slosh_status slosh_kernel_times(slosh_ctx *c, slosh_kernel_time *out, uint32_t cap, uint32_t *n)
{
    if (!c || !n) return SLOSH_EINVAL;
    *n = K_COUNT;
    if (!out) {
        for (int k = 0; k < K_COUNT; ++k) {
            c->ktime_ms[k] = 0.0;
            c->klaunches[k] = 0u;
        }
        return SLOSH_OK;
    }
    for (uint32_t k = 0; k < cap && k < (uint32_t)K_COUNT; ++k) {
        memset(&out[k], 0, sizeof(out[k]));
        strncpy(out[k].name, kKernelNames[k], sizeof(out[k].name) - 1);
        out[k].launches = c->klaunches[k];
        out[k].total_ms = c->ktime_ms[k];
    }
    return SLOSH_OK;
}

uint64_t slosh_unet_param_count(void) { return (uint64_t)slosh::unet_param_count(); }

slosh_status slosh_unet_set_weights(slosh_ctx *c, const float *w, uint64_t n)
{
    if (!c || !w || n != (uint64_t)unet_param_count()) return SLOSH_EINVAL;
    ABI_GUARD({
        unet_init(c);
        SLOSH_HIP(hipMemcpyAsync(c->unet.w32, w, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
        unet_upload_bf16(c);
        if (c->unet.m) {
            SLOSH_HIP(hipMemsetAsync(c->unet.m, 0, n * sizeof(float), c->stream));
            SLOSH_HIP(hipMemsetAsync(c->unet.v, 0, n * sizeof(float), c->stream));
            SLOSH_HIP(hipMemsetAsync(c->unet.grad, 0, n * sizeof(float), c->stream));
        }
        c->unet.adam_t = 0;
        SLOSH_HIP(hipStreamSynchronize(c->stream));
        c->unet.loaded = true;
        return SLOSH_OK;
    })
}

slosh_status slosh_unet_get_weights(slosh_ctx *c, float *w, uint64_t n)
{
    if (!c || !w || n != (uint64_t)unet_param_count()) return SLOSH_EINVAL;
    if (!c->unet.loaded) return SLOSH_ESTATE;
    ABI_GUARD({
        SLOSH_HIP(hipMemcpyAsync(w, c->unet.w32, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        SLOSH_HIP(hipStreamSynchronize(c->stream));
        return SLOSH_OK;
    })
}

namespace {
// Dense (x fastest) float field -> pool vector of the current allocation (unallocated cells dropped), on device.
void dense_to_pool_device(slosh_ctx *c, const float *dense, float *pool_dev)
{
    const Geo &g = c->geo;
    std::vector<uint32_t> table = fetch(c, c->table, g.nbricks);
    std::vector<float> pool((size_t)g.pool_cells, 0.0f);
    for (int z = 0; z < g.n[2]; ++z)
        for (int y = 0; y < g.n[1]; ++y)
            for (int x = 0; x < g.n[0]; ++x) {
                uint32_t slot = table[(uint32_t)((x >> 3) + g.bd[0] * ((y >> 3) + g.bd[1] * (z >> 3)))];
                if (!slot) continue;
                pool[(size_t)slot * kBrickCells + (size_t)((x & 7) + 8 * ((y & 7) + 8 * (z & 7)))] =
                    dense[(size_t)x + (size_t)g.n[0] * ((size_t)y + (size_t)g.n[1] * z)];
            }
    SLOSH_HIP(hipMemcpyAsync(pool_dev, pool.data(), pool.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    SLOSH_HIP(hipStreamSynchronize(c->stream));
}

bool unet_ready(const slosh_ctx *c)
{
    const Geo &g = c->geo;
    return c->unet.loaded && c->nactive > 0 && g.n[0] % 8 == 0 && g.n[1] % 8 == 0 && g.n[2] % 8 == 0;
}
} // namespace

slosh_status slosh_unet_apply(slosh_ctx *c, const float *r_dense, float *z_dense)
{
    if (!c || !r_dense || !z_dense) return SLOSH_EINVAL;
    if (!unet_ready(c)) return SLOSH_ESTATE;
    ABI_GUARD({
        const Geo &g = c->geo;
        float *rv = nullptr, *zv = nullptr;
        SLOSH_HIP(hipMalloc((void **)&rv, (size_t)g.pool_cells * 4u));
        SLOSH_HIP(hipMalloc((void **)&zv, (size_t)g.pool_cells * 4u));
        SLOSH_HIP(hipMemsetAsync(zv, 0, (size_t)g.pool_cells * 4u, c->stream));
        dense_to_pool_device(c, r_dense, rv);
        l0_brick_info(c);
        unet_apply(c, rv, zv);
        std::vector<uint32_t> table = fetch(c, c->table, g.nbricks);
        std::vector<float> h = fetch(c, zv, g.pool_cells);
        pool_to_dense(c, table, h.data(), z_dense, 1);
        (void)hipFree(rv);
        (void)hipFree(zv);
        return SLOSH_OK;
    })
}

slosh_status slosh_unet_loss_grad(slosh_ctx *c, const float *r_dense, double *loss, float *grad)
{
    if (!c || !r_dense || !loss || !grad) return SLOSH_EINVAL;
    if (!unet_ready(c)) return SLOSH_ESTATE;
    ABI_GUARD({
        const Geo &g = c->geo;
        UNet &u = c->unet;
        float *rv = nullptr;
        SLOSH_HIP(hipMalloc((void **)&rv, (size_t)g.pool_cells * 4u));
        dense_to_pool_device(c, r_dense, rv);
        l0_brick_info(c);
        unet_backward(c, rv);
        UNetScalars sc;
        SLOSH_HIP(hipMemcpyAsync(&sc, u.sc, sizeof(sc), hipMemcpyDeviceToHost, c->stream));
        SLOSH_HIP(hipMemcpyAsync(grad, u.grad, u.nparams * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        SLOSH_HIP(hipMemsetAsync(u.grad, 0, u.nparams * sizeof(float), c->stream));
        SLOSH_HIP(hipStreamSynchronize(c->stream));
        *loss = sc.rhat2 > 0.0 ? sc.err2 / sc.rhat2 : 0.0;
        (void)hipFree(rv);
        return SLOSH_OK;
    })
}

slosh_status slosh_unet_read_capture(slosh_ctx *c, uint32_t k, float *dst, uint64_t dst_bytes, uint32_t *n)
{
    if (!c || !dst || !n) return SLOSH_EINVAL;
    *n = c->capture_n;
    if (!c->capture) return SLOSH_ESTATE;
    if (k >= c->capture_n) return SLOSH_ECAPACITY;
    const Geo &g = c->geo;
    const size_t cells = (size_t)g.n[0] * g.n[1] * g.n[2];
    if (dst_bytes < cells * 4u) return SLOSH_ECAPACITY;
    ABI_GUARD({
        std::vector<uint32_t> table = fetch(c, c->table, g.nbricks);
        std::vector<float> h = fetch(c, c->capture + (size_t)k * g.pool_cells, g.pool_cells);
        pool_to_dense(c, table, h.data(), dst, 1);
        return SLOSH_OK;
    })
}

slosh_status slosh_unet_train(slosh_ctx *c, const slosh_unet_train_config *cfg, slosh_unet_train_report *out)
{
    if (!c || !cfg || cfg->iterations == 0 || cfg->batch == 0 || !(cfg->learning_rate > 0.0f)) return SLOSH_EINVAL;
    if (!unet_ready(c) || !c->capture || c->capture_n == 0) return SLOSH_ESTATE;
    ABI_GUARD({
        UNet &u = c->unet;
        l0_brick_info(c);
        double first = 0.0, last = 0.0;
        uint32_t k = cfg->offset;
        for (uint32_t it = 0; it < cfg->iterations; ++it) {
            SLOSH_HIP(hipMemsetAsync(&u.sc->loss_sum, 0, 2 * sizeof(double), c->stream));
            for (uint32_t b = 0; b < cfg->batch; ++b, ++k)
                unet_backward(c, c->capture + (size_t)(k % c->capture_n) * c->geo.pool_cells);
            unet_adam(c, cfg->learning_rate, 1.0f / (float)cfg->batch);
            if (it == 0 || it + 1 == cfg->iterations) {
                UNetScalars sc;
                SLOSH_HIP(hipMemcpyAsync(&sc, u.sc, sizeof(sc), hipMemcpyDeviceToHost, c->stream));
                SLOSH_HIP(hipStreamSynchronize(c->stream));
                const double mean = sc.loss_count > 0.0 ? sc.loss_sum / sc.loss_count : 0.0;
                if (it == 0) first = mean;
                last = mean;
            }
        }
        SLOSH_HIP(hipStreamSynchronize(c->stream));
        if (out) {
            out->first_loss = first;
            out->last_loss = last;
            out->systems = c->capture_n;
            out->adam_steps = (uint32_t)u.adam_t;
        }
        return SLOSH_OK;
    })
}

// SloshProxy.cpp:755-768 for one substep: sums in 1/1024 units, means over the wetted faces / submerged cells.
slosh_status slosh_get_collider_feedback(slosh_ctx *c, slosh_collider_feedback *out, uint32_t cap, uint32_t *n)
{
    if (!c || !n || (cap && !out)) return SLOSH_EINVAL;
    ABI_GUARD({
        flush_submerged(c);
        const uint32_t nc = c->coupling_ncol;
        *n = nc;
        const uint32_t m = nc < cap ? nc : cap;
        if (m) {
            int32_t raw[SLOSH_MAX_COLLIDERS * 16];
            SLOSH_HIP(hipMemcpyAsync(raw, c->coupling, (size_t)m * 16u * sizeof(int32_t), hipMemcpyDeviceToHost,
                                     c->stream));
            SLOSH_HIP(hipStreamSynchronize(c->stream));
            const float inv = 1.0f / 1024.0f;
            for (uint32_t k = 0; k < m; ++k) {
                const int32_t *C = raw + k * 16u;
                slosh_collider_feedback &o = out[k];
                memset(&o, 0, sizeof(o));
                for (int a = 0; a < 3; ++a) {
                    o.pressure_force[a] = (float)C[a] * inv;
                    o.pressure_torque[a] = (float)C[3 + a] * inv;
                    o.mean_fluid_velocity[a] = C[6] > 0 ? (float)C[8 + a] * (inv / (float)C[6]) : 0.0f;
                    o.submerged_centroid_offset[a] = C[7] > 0 ? (float)C[12 + a] * (inv / (float)C[7]) : 0.0f;
                }
                o.wetted_faces = C[6];
                o.submerged_cells = C[7];
                o.solid_cells = C[11];
            }
        }
        return SLOSH_OK;
    })
}

slosh_status slosh_get_results(slosh_ctx *c, slosh_results *out)
{
    if (!c || !out) return SLOSH_EINVAL;
    *out = c->res;
    out->live_particles = c->np;
    return SLOSH_OK;
}

} // extern "C"
