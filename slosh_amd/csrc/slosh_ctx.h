// slosh_ctx.h -- host-side context of one simulation domain and the launchers of the gfx950 kernels.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "slosh_device.h"

namespace slosh {

// Particles in structure-of-arrays form: position + age as one float4 (the G2P and sort-key passes read 16 B),
// velocity and the APIC matrix as scalar planes.
struct Parts {
    float4 *pa;
    float *v[3];
    float *c[9];
};

// PCG / SOR scalars, kept on the device so no iteration needs a host round trip.
struct SolverState {
    double rho;    // r.z of the current preconditioned residual
    double dq;     // d.Ad
    double alpha;
    double beta;
    double rr;     // ||r||^2
    double bb;     // ||b||^2
    double rel;    // ||r|| / ||b||
    double tol;
    double rel_true; // ||b - A p|| / ||b|| of the last true-residual check
    double nf;       // fluid cells
    double mg_scale; // FP16 levels: the V-cycle input scale sqrt(nf / ||r||^2), so the levels hold O(1) values
    double err_rel;  // SLOSH_FLAG_REPORT_FP64_ERROR: relative L2 error of the solution against the FP64-accurate one
    double hist[4];  // flexible CG: relative residuals of the last four iterations (stall test)
    int32_t done;  // 1 once converged or out of iterations: every solver kernel then returns at once; 2 when the
                   // MGPCG update being made is the last (only the update pass runs; the next finalize sets 1)
    int32_t iters;
    int32_t maxit;
    int32_t converged;
    int32_t stalled; // flexible CG: stopped without converging (no halving over four iterations, the iteration
                     // limit, or a non-finite residual): the MGPCG fallback takes over
    int32_t pad;
};

// Multigrid level >= 1. Levels 1 and 2 are pooled (stored per allocated brick, (8 >> l)^3 cells per slot); level 3
// (one cell per brick) and coarser levels are dense (x fastest; level 3 is indexed like the brick grid). Every level keeps a per-cell stencil word (see kInfo*), the down-sweep iterate x,
// the up-sweep result x1 and the right-hand side b.
struct Level {
    int dim[3];
    uint32_t cells;
    bool pooled;
    bool half;       // x, x1, b stored as FP16 (SLOSH_FLAG_MG_FP16), else FP32
    uint8_t *type;
    uint16_t *info;
    uint32_t *coord; // dense levels: packed cell coordinates x | y << 10 | z << 20
    void *x, *x1, *b;
    uint32_t dense_off; // dense levels: offset of this level's words and coordinates in the contiguous blocks
};

// A coarse-level vector in FP32 or FP16 storage; arithmetic is FP32 either way (FP16 values are rounded to
// nearest even on store). v[i] reads and writes like a float.
struct LVec {
    void *p;
    int half;
    struct Ref {
        void *p;
        int half;
        uint32_t i;
        __device__ __forceinline__ operator float() const
        {
            return half ? __half2float(((const __half *)p)[i]) : ((const float *)p)[i];
        }
        __device__ __forceinline__ Ref &operator=(float v)
        {
            if (half) ((__half *)p)[i] = __float2half(v);
            else ((float *)p)[i] = v;
            return *this;
        }
        __device__ __forceinline__ Ref &operator=(const Ref &o) { return *this = (float)o; }
        __device__ __forceinline__ Ref &operator+=(float v) { return *this = (float)*this + v; }
    };
    __device__ __forceinline__ Ref operator[](uint32_t i) const { return Ref{p, half, i}; }
};

// Words per allocated brick in the per-step brick table: slots of the 3x3x3 neighbourhood (entry (dx+1) +
// 3(dy+1) + 9(dz+1), 0 = none), then the brick index.
constexpr uint32_t kBinfoWords = 32u;
// Most values a level-0 reduction writes per workgroup (k_direction: d.q, r.q, q.q, r.r; k_pcg_init three); the
// partials buffer holds this many per pool slot, one workgroup per allocated brick.
constexpr uint32_t kMaxBlockPartials = 4u;
// Reductions over more partial sets than kFinOneMax (one set per brick: grids of about 400^3 and more) are first
// folded by up to kFoldGroups workgroups into the context's `fold` array, so that no single workgroup reads them
// all. A tank above the bound solves on its own stream rather than in a lockstep batch, whose finalize reads each
// tank's partials in one workgroup.
constexpr uint32_t kFinOneMax = 16384u;
constexpr uint32_t kFoldGroups = 512u;

// Wavefront sum of an FP64 value in a fixed order, without LDS round trips: inclusive sums within each row of 16
// lanes by DPP row shifts of 1, 2, 4 and 8 (a lane shifted in from outside its row adds +0), then the four row
// totals (lanes 15, 31, 47, 63) added in row order. Every lane returns the same value.
template <int kCtrl> __device__ __forceinline__ double dpp_row_shr(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), kCtrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), kCtrl, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_value(double v, int lane)
{
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                            __builtin_amdgcn_readlane(__double2loint(v), lane));
}
__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_row_shr<0x111>(v); // row_shr:1
    v += dpp_row_shr<0x112>(v); // row_shr:2
    v += dpp_row_shr<0x114>(v); // row_shr:4
    v += dpp_row_shr<0x118>(v); // row_shr:8
    return ((lane_value(v, 15) + lane_value(v, 31)) + lane_value(v, 47)) + lane_value(v, 63);
}

// Brick handled by workgroup w of a one-brick-per-workgroup pass over n bricks. Workgroups are dispatched
// round-robin over the MI355X's 8 XCDs, each with its own L2; giving XCD k a contiguous range of the ascending
// brick list keeps most of a brick's neighbours (and their halo reads) in the same L2. A bijection on [0, n).
__device__ __forceinline__ uint32_t xcd_remap(uint32_t w, uint32_t n)
{
    const uint32_t q = n >> 3, r = n & 7u, k = w & 7u, j = w >> 3;
    return k * q + (k < r ? k : r) + j;
}

// A workgroup's index among the workgroups of its pass and their count: blockIdx.x and gridDim.x of a single-tank
// launch; a launch over a batch of tanks gives each tank's workgroups their own range.
struct Blk {
    uint32_t b, nb;
};
__device__ __forceinline__ Blk this_blk() { return Blk{blockIdx.x, gridDim.x}; }

// Stencil word of a coarse cell: bits 0-5 fluid neighbours (-x +x -y +y -z +z), bit 6 the cell is fluid, bit 7 red
// ((x + y + z) even), bits 8-11 the number of non-solid neighbours (the coarse diagonal without the 1e-4).
constexpr uint32_t kInfoSelf = 1u << 6;
constexpr uint32_t kInfoRed = 1u << 7;

// Diagonal of a coarse cell, 1e-4 plus one per non-solid neighbour, summed in the same order as the reference.
__device__ __forceinline__ float info_diag(uint32_t info)
{
    float d = kDiagEps;
    for (uint32_t k = 0; k < (info >> 8); ++k) d += 1.0f;
    return d;
}

// Kernel-side views of a pooled and of a dense level.
struct PLevel {
    int dim[3];
    int ls;        // log2 of cells per brick axis
    int s;         // cells per brick axis (8 >> level)
    uint32_t cpb;  // cells per brick
    uint8_t *type;
    uint16_t *info;
    LVec x, x1, b;
};
struct DenseLevel {
    int dim[3];
    uint32_t cells;
    uint8_t *type;
    uint16_t *info;
    const uint32_t *coord;
    LVec x, x1, b;
};

// The dense levels the coarse tail runs in LDS (one workgroup).
struct TailArgs {
    int n;                   // dense levels in the tail, the last one the coarsest
    int sweeps;              // red-black smoothing pairs per visit of a level
    uint32_t total;          // cells of all tail levels
    const uint16_t *info;    // stencil words of the tail's levels, contiguous from the first one
    const uint32_t *coord;   // their packed cell coordinates
    uint32_t off[kMaxLevels];  // first cell of each tail level in the two arrays above and in LDS
    DenseLevel L[kMaxLevels];  // per level: dimensions and cell count; b and the red-initialised x of the first
                               // level are in memory
};

// One tank of a batched MGPCG solve: what its solver passes read, in device memory. The batched passes launch the
// workgroups of every tank at once, each workgroup finding its tank from per-tank prefix sums of workgroup counts.
struct BatchTank {
    const uint32_t *binfo_f;
    const uint8_t *fmask;
    const float *diag, *div;
    double *pres;
    float *r[2], *z, *z1, *d[2], *q;
    double *partials;
    SolverState *state;
    uint32_t nsolve, ntiles_f;
    Geo geo;
    const uint32_t *tile_list;
    PLevel L1, L2;
    DenseLevel L3, L4;
    TailArgs tail;
    const uint32_t *coarse_idx, *coarse_n;
    const double *coarse_inv;
};
constexpr uint32_t kMaxBatchTanks = 256;
// Workgroup ranges of the batched passes: column passes (bricks per workgroup as configured), one brick per
// workgroup, one level-3 tile per workgroup, one workgroup per tank.
enum BatchRange { kRangeCol = 0, kRangeBrick = 1, kRangeTile = 2, kRangeTank = 3, kBatchRanges = 4 };

// The tank of workgroup b of a batched pass (range: the pass's prefix sums over the tanks, range[0] = 0) and the
// workgroup's place among that tank's workgroups.
__device__ __forceinline__ uint32_t batch_tank(const uint32_t *range, uint32_t ntanks, uint32_t b, Blk &G)
{
    uint32_t lo = 0, hi = ntanks; // range[lo] <= b < range[hi]
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (range[mid] <= b) lo = mid;
        else hi = mid;
    }
    G.b = b - range[lo];
    G.nb = range[lo + 1] - range[lo];
    return lo;
}

__device__ __forceinline__ bool pinside(const PLevel &L, int x, int y, int z)
{
    return x >= 0 && y >= 0 && z >= 0 && x < L.dim[0] && y < L.dim[1] && z < L.dim[2];
}

// Pooled index of level coordinates inside the level's extent (through the brick table).
__device__ __forceinline__ uint32_t pindex(const Geo &g, const PLevel &L, int x, int y, int z)
{
    const int m = L.s - 1;
    uint32_t b = (uint32_t)((x >> L.ls) + g.bd[0] * ((y >> L.ls) + g.bd[1] * (z >> L.ls)));
    uint32_t local = (uint32_t)((x & m) + L.s * ((y & m) + L.s * (z & m)));
    return g.table[b] * L.cpb + local;
}

// Kernels that can be timed with HIP events on the context's stream (slosh_set_kernel_timing).
enum KernelId {
    K_EMIT = 0, K_SORT_TILES, K_SORT_RUNS, K_SORT_KEY_RUNS, K_SORT_SCATTER, K_SCAN, K_POOL, K_P2G, K_NORMALIZE,
    K_DIVERGENCE, K_PROJECT, K_DIV_STATS, K_EXTRAPOLATE, K_G2P, K_MG_SETUP, K_COARSE_FACTOR, K_PCG_INIT, K_RED_DOT,
    K_RESTRICT0, K_PROLONG_BLACK, K_DIRECTION, K_UPDATE_BLACK, K_PDOWN1, K_PDOWN2, K_PUP1, K_PUP2, K_D3DOWN, K_D3UP,
    K_DSMOOTH, K_DRESTRICT, K_DPROLONG, K_COARSE_SOLVE, K_COARSE_TAIL, K_SOR_PAIR, K_SOR_RESIDUAL, K_FILL,
    K_FINALIZE, K_STEP_STATS, K_UNET_CONV0, K_UNET_CONV1, K_UNET_CONV2, K_UNET_IO, K_UNET_TRAIN, K_FCG, K_G2P_MOVE,
    K_DENSITY, K_VISCOSITY, K_COUPLING, K_COUNT
};
// Names as the compiled kernels are called (what a kernel trace of the library shows).
extern const char *const kKernelNames[K_COUNT];

// ---- learned preconditioner (unet.hip) ----
struct UNetScalars {
    float s, s_inv;      // RMS of the input residual over fluid cells, and its inverse
    double err2, rhat2;  // loss terms of the last backward pass: sum (A zhat - rhat)^2, sum rhat^2
    double loss_sum;     // sum of err2 / rhat2 over backward passes since the last reset
    double loss_count;
};
struct UNetLayerDesc {
    int cin, cout, level, taps;
};
constexpr int kUNetLayerCount = 11;
extern const UNetLayerDesc kUNetLayers[kUNetLayerCount];
struct UNet {
    size_t nparams = 0;
    size_t w_off[kUNetLayerCount]{}, b_off[kUNetLayerCount]{}, wf_off[kUNetLayerCount]{}, wt_off[kUNetLayerCount]{};
    float *w32 = nullptr;   // FP32 master parameters (flat)
    uint16_t *wbf = nullptr; // BF16 weights per layer: forward [tap][cout][cin (padded to 16)], mirrored-transposed
    UNetScalars *sc = nullptr;
    // activations, channel-last BF16 in the pool layout of their level
    uint16_t *x0 = nullptr, *a0 = nullptr, *s0 = nullptr, *t0 = nullptr, *d0 = nullptr;                  // level 0
    uint16_t *p0 = nullptr, *a1 = nullptr, *s1 = nullptr, *t1 = nullptr, *d1 = nullptr, *c1 = nullptr;   // level 1
    uint16_t *p1 = nullptr, *a2 = nullptr, *b2 = nullptr, *c2 = nullptr;                                 // level 2
    float *zhat = nullptr;
    double *partials = nullptr;
    int32_t *zero_flag = nullptr;     // a device 0: the forward kernels' skip flag outside a flexible-CG solve
    const int32_t *skip = nullptr;    // forward kernels return at once while *skip != 0
    // training state (allocated by the first backward pass)
    float *grad = nullptr, *m = nullptr, *v = nullptr, *e = nullptr;
    uint16_t *g16a = nullptr, *g16b = nullptr, *g16c = nullptr, *g16_1a = nullptr, *g16_1b = nullptr;
    uint16_t *g32a = nullptr, *g32b = nullptr, *g32c = nullptr, *g32_2a = nullptr, *g32_2b = nullptr;
    uint16_t *g64a = nullptr, *g64b = nullptr;
    int adam_t = 0;
    bool loaded = false; // weights were set
};

struct StepStats {
    uint32_t fluid_cells;
    uint32_t clamped;
    uint32_t max_speed_bits;
    uint32_t max_div_bits;
    uint32_t overflow;
    uint32_t nactive;
    uint32_t nfree;
    uint32_t misplaced; // particles whose two G2P passes disagree on the new brick (never expected; fails the step)
    uint32_t drained;   // particles removed by drain manipulators this step
    double div_sum;
    uint32_t div_count;
    uint32_t pad2;
};

} // namespace slosh

struct slosh_ctx {
    slosh_settings s;
    int device = 0;
    hipStream_t stream = nullptr;
    slosh::Geo geo{};      // host copy; table/active point to device memory
    uint32_t nkeys = 0;    // nbricks + 1 (dead key)
    uint32_t cap = 0;      // particle capacity
    uint32_t np = 0;       // particle slots in use (host-exact)
    uint32_t ndead = 0;    // dead records among them, dropped by the next sort
    bool sorted = false;   // the records are in brick-key order with key_off current (the last step's G2P sorted
                           // them); emitting or writing particles clears it and the next step sorts
    uint32_t ntiles_max = 0;
    uint32_t nactive = 0;  // allocated bricks after the last pool pass (read back once per step)

    slosh::Parts parts[2]{};
    int cur = 0; // which particle buffer holds the live, sorted particles

    // sort scratch
    uint32_t *packed = nullptr, *tile_nruns = nullptr, *run_key = nullptr, *run_len = nullptr, *run_base = nullptr,
             *run_sorted = nullptr, *key_count = nullptr, *key_off = nullptr, *key_runs = nullptr,
             *key_off_next = nullptr, *g2p_keys = nullptr, *run_off = nullptr,
             *run_cursor = nullptr;
    // scan scratch
    uint32_t *scan_blocks = nullptr;
    // brick pool
    uint32_t *table = nullptr, *slot_brick = nullptr, *active = nullptr, *need = nullptr, *need_rank = nullptr,
             *freeflag = nullptr, *free_rank = nullptr, *free_list = nullptr, *alloc_flag = nullptr,
             *alloc_rank = nullptr;
    uint8_t *new_slot = nullptr; // slots assigned this step: their warm-start pressure reads as 0
    uint8_t *touch = nullptr;
    // grid (pool layout)
    int32_t *tiles = nullptr;
    uint8_t *tile_valid = nullptr;
    float *vel[3] = {nullptr, nullptr, nullptr};
    int vcur = 0;
    uint8_t *fflags = nullptr;
    float *fweight = nullptr;
    uint32_t *ctype = nullptr;
    double *pres = nullptr;
    double *sor_tmp = nullptr; // RB-SOR: the second pressure array of the two-colour passes
    float *div = nullptr, *diag = nullptr;
    uint8_t *fmask = nullptr; // solver stencil mask per cell (fluid neighbours, self fluid)
    // solver: residual and search direction ping-pong between iterations, z is the down-sweep iterate and z1 the
    // V-cycle result; binfo holds kBinfoWords per allocated brick (its 3x3x3 neighbourhood slots, brick index)
    float *r = nullptr, *r2 = nullptr, *z = nullptr, *z1 = nullptr, *d = nullptr, *d2 = nullptr, *q = nullptr;
    uint32_t *binfo = nullptr;
    // rows of binfo for the bricks holding fluid cells, in active order: the solver's level-0 and pooled passes
    // visit these only (a brick without fluid has no fluid cell on levels 0-2 and nothing reads its values)
    uint32_t *binfo_f = nullptr;
    uint32_t *fluid_flag = nullptr, *fluid_rank = nullptr;
    uint32_t *h_nsolve = nullptr; // pinned: [0] the count of such bricks, [1] of level-3 tiles with fluid bricks
    uint32_t nsolve = 0;
    // 4^3 tiles of level 3 whose bricks hold fluid: the fused level-2/3 passes visit these only
    uint32_t *tile_flag = nullptr, *tile_rank = nullptr, *tile_list = nullptr;
    uint32_t ntiles = 0, ntiles_f = 0;
    int nlevels = 1;
    slosh::Level lv[slosh::kMaxLevels]{};
    uint32_t *coarse_idx = nullptr; // [kCoarsestCells]
    uint32_t *coarse_n = nullptr;
    double *coarse_inv = nullptr;   // [kCoarsestCells^2]
    double *pres_prev = nullptr;    // SLOSH_FLAG_WARM_START_EXTRAPOLATE: last solution per unit dt (NaN: none)
    double *pres_lp = nullptr;      // SLOSH_FLAG_REPORT_FP64_ERROR: the solution kept while the FP64-accurate one is formed
    // SLOSH_FLAG_DENSITY_PROJECTION: the projection's right-hand side, potential and face displacements
    float *dens_rhs = nullptr;
    double *dens_q = nullptr;
    float *dens_disp = nullptr;
    bool dens_applied = false;      // this step's G2P applies the displacement
    slosh::SolverState *state_save = nullptr;
    uint32_t *coarse_sig = nullptr; // the coarsest level's cell types and indices at its last factorisation
    double *partials = nullptr;
    double *fold = nullptr; // [kFoldGroups][kMaxBlockPartials]: the partials of a large reduction, folded
    uint32_t nparts = 0;              // partial sets the last reduction pass wrote (what the next finalize sums)
    uint32_t col_bricks = 4;          // bricks (wavefronts) per workgroup of the level-0 column passes: 1, 2 or 4
    // the step between its two halves (step_front / step_back): dt, gravity and the last stage
    float st_dt = 0.0f;
    float st_g[3]{};
    int st_stop = 0;
    hipEvent_t ev_batch = nullptr;    // a lockstep batch waits on the tank's stream at this event
    bool submerged_pending = false;   // the last step's submerged-cell pass has not run yet (it runs on demand)
    uint16_t *dense_info = nullptr;   // stencil words of all dense levels, level by level
    uint32_t *dense_coord = nullptr;  // their packed cell coordinates
    slosh::SolverState *state = nullptr;
    slosh::StepStats *stats = nullptr;
    // per-workgroup statistics partials, reduced once at the end of a step (no cross-workgroup atomics)
    uint32_t *bp_fluid = nullptr;   // [npool]   P2G: fluid cells per brick
    uint32_t *bp_div_max = nullptr; // [npool]   |div|/h max bits per brick
    uint32_t *bp_div_cnt = nullptr; // [npool]
    double *bp_div_sum = nullptr;   // [npool]
    uint32_t *bp_speed = nullptr;   // [nbricks] G2P: max speed bits per brick of the sorted particles
    uint32_t *bp_clamp = nullptr;   // [nbricks]
    uint32_t *bp_drain = nullptr;   // [nbricks] G2P: particles drained per brick
    uint32_t n_p2g_blocks = 0, n_div_blocks = 0, n_g2p_blocks = 0;
    // MGPCG iterations per host check of the solver state, and their replay as a captured HIP graph
    int solver_chunk = 4;
    int last_mg_iters = 0; // iterations of the previous MGPCG solve (chunks enqueued before the first state read)
    bool use_graphs = true;
    hipGraphExec_t mg_graph = nullptr;
    // learned preconditioner and the residuals captured for its training
    slosh::UNet unet;
    float *fcg_z = nullptr;           // [pool_cells] preconditioned residual of the flexible CG
    float *fcg_rp = nullptr;          // [pool_cells] the residual the V-cycle leaves (the network's input)
    hipGraphExec_t fcg_graph = nullptr; // a flexible-CG iteration chunk, captured once per solve
    int last_fcg_iters = 0;
    float *capture = nullptr;         // [capture_cap][pool_cells]
    uint32_t capture_cap = 0, capture_n = 0;
    int capture_iter[32] = {};        // MGPCG iteration of each captured residual
    int fcg_fallback = 0, fcg_iters = 0;
    // colliders and manipulators
    slosh_collider *d_col = nullptr;
    slosh_collider h_col[SLOSH_MAX_COLLIDERS];
    uint32_t ncol = 0;
    slosh_manipulator *d_man = nullptr;
    uint32_t nman = 0;
    int32_t *coupling = nullptr;      // [SLOSH_MAX_COLLIDERS][16] fixed-point coupling sums of the last step
    uint32_t coupling_ncol = 0;       // colliders the last step's sums cover
    uint64_t drained_total = 0;
    // host pinned mirrors
    slosh::StepStats *h_stats = nullptr;
    slosh::SolverState *h_state = nullptr;
    // readback staging
    void *staging = nullptr;
    size_t staging_bytes = 0;
    // results and timing
    slosh_results res{};
    uint64_t frame = 0;
    float prev_dt = 0.0f;
    hipEvent_t ev_step0 = nullptr, ev_step1 = nullptr, ev_solve0 = nullptr, ev_solve1 = nullptr;
    hipEvent_t ev_pool = nullptr;   // after the pool pass's readback of the step statistics
    hipEvent_t ev_counts = nullptr; // after the multigrid set-up's readback of the fluid brick and tile counts
    // per-kernel timing (event pairs around selected launches, accumulated after each step's final sync)
    uint64_t timing_mask = 0;
    hipEvent_t *tev = nullptr;
    uint32_t tev_cap = 0, tev_used = 0;
    int *tev_kernel = nullptr;
    double ktime_ms[slosh::K_COUNT] = {};
    uint32_t klaunches[slosh::K_COUNT] = {};
};

namespace slosh {

// ---- particles.hip ----
void launch_emit(slosh_ctx *c, const slosh_emit_request *r, uint32_t count);
void launch_unpack_particles(slosh_ctx *c, const slosh_particle *d_src, uint32_t n);
void launch_pack_particles(slosh_ctx *c, void *d_dst, int layout);
void run_sort(slosh_ctx *c);
// Active bricks [first, first + count) of the pool, clipped on the device to the count the pool pass wrote (so that
// they can be launched before the host has read that count back).
void launch_p2g(slosh_ctx *c, float warm_scale, float dt, float dt_prev, uint32_t first, uint32_t count);
void launch_g2p(slosh_ctx *c, float dt, const float g[3]);

// ---- grid.hip ----
void run_pool(slosh_ctx *c);
void build_brick_rows(slosh_ctx *c, uint32_t first, uint32_t count);
// Device address of the active-brick count in the step statistics.
const uint32_t *device_nactive(const slosh_ctx *c);
void launch_normalize(slosh_ctx *c, float dt, const float g[3]);
void launch_divergence(slosh_ctx *c, float dt);
void launch_project(slosh_ctx *c);
void launch_extrapolate(slosh_ctx *c, uint32_t layer);
void launch_step_stats(slosh_ctx *c);
void launch_density_rhs(slosh_ctx *c, float relax, int keep_q);
// Slosh.usf:657-686 / SloshProxy.cpp:603-617: `iterations` Jacobi sweeps of the implicit viscosity on the
// normalised velocities (alpha = nu dt / h^2); the result becomes the current velocity buffer.
void launch_coupling(slosh_ctx *c, float dt);
void launch_submerged(slosh_ctx *c);
void run_viscosity(slosh_ctx *c, float alpha, int iterations);
void launch_density_disp(slosh_ctx *c);
void scan_exclusive(slosh_ctx *c, const uint32_t *in, uint32_t *out, uint32_t n);
// Two independent exclusive scans (one launch when both fit one workgroup).
void scan_exclusive2(slosh_ctx *c, const uint32_t *in0, uint32_t *out0, uint32_t n0, const uint32_t *in1, uint32_t *out1,
                     uint32_t n1);
void launch_fill_u32(hipStream_t st, uint32_t *p, uint32_t v, size_t n);
void launch_zero3_u32(hipStream_t st, uint32_t *a, uint32_t *b, uint32_t *c, size_t n); // three arrays of n, one launch
void launch_fill_u8(hipStream_t st, uint8_t *p, uint8_t v, size_t n);
void launch_fill_f64(hipStream_t st, double *p, double v, size_t n);

// ---- solver.hip ----
void run_mgpcg(slosh_ctx *c);
void run_rbsor(slosh_ctx *c);
// `iters` red-black SOR iterations on c->pres (two-colour passes alternating between c->pres and c->sor_tmp).
void sor_iterate(slosh_ctx *c, int iters, double w);
// The density projection's solve A q = b, b = -dens_rhs, from q = 0, with the pressure solver's operator: two MGPCG
// iterations (RB-SOR: density_projection_iterations sweeps at omega 1.7), as the reference's projection pass.
void run_density_solve(slosh_ctx *c);

// ---- pooled.hip: MG levels 1 and 2 in the brick pool ----
PLevel plevel_of(const slosh_ctx *c, int l);
// 1 when the settings ask for fluid-first coarse cell types (SLOSH_FLAG_MG_FLUID_COARSENING).
inline int fluid_coarsening(const slosh_ctx *c) { return (c->s.flags & SLOSH_FLAG_MG_FLUID_COARSENING) ? 1 : 0; }
// Red-black smoothing pairs per visit of a dense coarse level (settings.mg_coarse_sweeps; 0 means 1).
inline int coarse_sweeps(const slosh_ctx *c) { return c->s.mg_coarse_sweeps > 0 ? c->s.mg_coarse_sweeps : 1; }
DenseLevel dlevel_of(const slosh_ctx *c, int l);
void pooled_build_type(slosh_ctx *c, int l);
void pooled_build_info(slosh_ctx *c, int l);
void dense_build_type_from_pooled(slosh_ctx *c, int l); // dense level l from pooled level l - 1
// Black sweep of pooled level l (1 or 2), then its residual restricted into level l + 1 (b, and x = b / diag on
// red cells).
void pooled_down(slosh_ctx *c, int l);
// Prolongation of the parent level's result into pooled level l, then a black and a red sweep: x1 of level l.
void pooled_up(slosh_ctx *c, int l);

// ---- level0.hip: finest-level solver passes (one brick per workgroup; one reduction partial per brick) ----
void l0_brick_info(slosh_ctx *c, bool readback = true);
void l0_pcg_init(slosh_ctx *c, float *r_out);
// p += alpha d, r_out = r_in - alpha q (first: r_out = r_in), z = r_out / diag on red cells, then the black sweep;
// partials ||r||^2 unless first.
void l0_update_black(slosh_ctx *c, const float *r_in, float *r_out, const float *d, int first);
void l0_restrict(slosh_ctx *c, const float *r);
// z1 = z + prolongation of level 1's x1, then the black sweep into z1.
void l0_prolong_black(slosh_ctx *c, const float *r);
// Red sweep on z1 and the partials r.z1.
void l0_red_dot(slosh_ctx *c, const float *r);
// d_out = z1 + beta d_in (first: d_out = z1), q = A d_out; partials d.q.
void l0_direction(slosh_ctx *c, const float *d_in, float *d_out, const float *r, int first);
void l0_sor_sweep(slosh_ctx *c, uint32_t colour, double w);
// A red and then a black sweep in one pass, from pin into pout (distinct arrays whose non-fluid cells agree).
void l0_sor_pair(slosh_ctx *c, const double *pin, double *pout, double w);
// Partials ||b||^2 and ||b - A p||^2 of the FP64 pressure (the true residual).
void l0_sor_residual(slosh_ctx *c);
// The same partials whatever the solver state (the true residual of a finished solve).
void l0_true_residual_if_done(slosh_ctx *c);
void l0_true_residual(slosh_ctx *c);
void l0_fcg_zq(slosh_ctx *c, const float *z);
void l0_fcg_direction(slosh_ctx *c, const float *z, const float *d_in, float *d_out, const float *r, int first);
void l0_fcg_update(slosh_ctx *c, float *r, const float *d);
// Partials sum (p_ref - pres)^2 and sum pres^2 over fluid cells, pres the current FP64 pressure.
void l0_error_partials(slosh_ctx *c, const double *p_ref);
// out = r - A z over the fluid cells (the residual a preconditioner's result z leaves).
void l0_fcg_resid(slosh_ctx *c, const float *r, const float *z, float *out);

// ---- batched MGPCG iterations over many tanks (batch.hip drives them; each file launches its own passes) ----
struct BatchRun {
    hipStream_t stream = nullptr;
    uint32_t ntanks = 0;
    BatchTank *d_tanks = nullptr;
    uint32_t *d_range = nullptr;     // [kBatchRanges][kMaxBatchTanks + 1] prefix sums of workgroups per tank
    uint32_t total[kBatchRanges]{};  // workgroups of each range over the whole batch
    uint32_t col_bricks = 4;         // bricks per workgroup of the column passes
    const uint32_t *range(int k) const { return d_range + (size_t)k * (kMaxBatchTanks + 1u); }
};
// level0.hip: the column passes over every tank (buffer selectors: r[ri] in, r[ro] out, d[di], d[dout])
void b_update_black(const BatchRun &B, int ri, int ro, int di, bool first);
void b_restrict(const BatchRun &B, int ri);
void b_prolong_black(const BatchRun &B, int ri);
void b_red_dot(const BatchRun &B, int ri);
void b_direction(const BatchRun &B, int din, int dout, int ri, bool first);
void b_true_residual_if_done(const BatchRun &B);
// pooled.hip: the level-1 down and up passes
void b_pooled_down1(const BatchRun &B);
void b_pooled_up1(const BatchRun &B);
// solver.hip: the fused level-2/3 passes, the coarse tail, the finalize (one workgroup per tank)
void b_d23down(const BatchRun &B);
void b_d23up(const BatchRun &B);
void b_coarse_tail(const BatchRun &B);
void b_finalize(const BatchRun &B, int kind);

// solver.hip: the MGPCG set-up of a solve (returns whether it needs iterations), whether a tank's solve can join a
// lockstep batch, its descriptor, the batched iterations, and the single-tank continuation after them (restarts on
// a missed true residual)
bool mgpcg_begin(slosh_ctx *c);
bool mgpcg_batchable(const slosh_ctx *c);
void batch_tank_of(const slosh_ctx *c, BatchTank &T);
void batch_iterate(BatchRun &B, slosh_ctx *const *tanks, uint32_t n, int chunk, bool graphs, hipGraphExec_t &graph,
                   SolverState *d_states, SolverState *h_states);
void mgpcg_resume(slosh_ctx *c);
// ctx.hip: the halves of a full step around its pressure solve, the single-tank solve, and a batched solve's
// per-tank ends
void tank_step_front(slosh_ctx *c, float dt, const float gv[3]);
void tank_step_back(slosh_ctx *c, bool solved);
bool tank_solve(slosh_ctx *c);
bool tank_solve_begin(slosh_ctx *c);
void tank_solve_end(slosh_ctx *c, bool iterated);

// ---- unet.hip ----
size_t unet_param_count();
void unet_init(slosh_ctx *c);
void unet_free(slosh_ctx *c);
void unet_upload_bf16(slosh_ctx *c);
void unet_apply(slosh_ctx *c, const float *r, float *z, const float *zadd = nullptr);
void unet_backward(slosh_ctx *c, const float *r);
void unet_adam(slosh_ctx *c, float lr, float scale);

// ---- fcg.hip: flexible CG preconditioned by the network, MGPCG as fallback ----
void run_fcg(slosh_ctx *c);

} // namespace slosh

#define SLOSH_HIP(call)                                                                                          \
    do {                                                                                                         \
        hipError_t e_ = (call);                                                                                  \
        if (e_ != hipSuccess) throw slosh::HipError(e_, #call, __FILE__, __LINE__);                            \
    } while (0)

namespace slosh {
struct HipError {
    hipError_t err;
    const char *what;
    const char *file;
    int line;
    HipError(hipError_t e, const char *w, const char *f, int l) : err(e), what(w), file(f), line(l) {}
};

// Records an event pair around the launches in its scope when kernel `id` is selected for timing.
struct KTimer {
    slosh_ctx *c;
    int slot;
    KTimer(slosh_ctx *ctx, int id) : c(ctx), slot(-1)
    {
        if (!(c->timing_mask >> id & 1ull) || c->tev_used + 2 > c->tev_cap) return;
        slot = (int)c->tev_used;
        c->tev_used += 2;
        c->tev_kernel[slot / 2] = id;
        (void)hipEventRecord(c->tev[slot], c->stream);
    }
    ~KTimer()
    {
        if (slot >= 0) (void)hipEventRecord(c->tev[slot + 1], c->stream);
    }
};

// Launches one kernel on the context's stream. When kernel `id` is selected for timing, the event pair rides on
// the dispatch itself (hipExtLaunchKernelGGL), so the elapsed time is the kernel's own begin-to-end time, without
// the gaps an event recorded before and after the launch would add.
template <typename F, typename... Args>
inline void launch_timed(slosh_ctx *c, int id, F kernel, dim3 grid, dim3 block, uint32_t shm, Args... args)
{
    if ((c->timing_mask >> id & 1ull) && c->tev_used + 2 <= c->tev_cap) {
        const uint32_t slot = c->tev_used;
        c->tev_used += 2;
        c->tev_kernel[slot / 2] = id;
        hipExtLaunchKernelGGL(kernel, grid, block, shm, c->stream, c->tev[slot], c->tev[slot + 1], 0u, args...);
    } else {
        hipLaunchKernelGGL(kernel, grid, block, shm, c->stream, args...);
    }
}
} // namespace slosh

#define SLOSH_TIMED(c, id) slosh::KTimer slosh_ktimer_(c, slosh::id)
