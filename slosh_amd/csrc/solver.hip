// solver.hip -- pressure solvers on gfx950: MGPCG (SloshPressureMG.usf) and red-black SOR (Slosh.usf:856-892).
//
// Both solve (CellDiag_c + 1e-4) p_c - sum_{fluid n} p_n = -Divergence_c on the fluid cells of the brick pool.
// MGPCG keeps FP32 vectors, an FP64 pressure (the alpha-update accumulator), FP64 dot products and alpha/beta, and
// solves the coarsest level (<= 64 cells) exactly in FP64. All scalars live in a device SolverState; every solver
// kernel returns at once when `done` is set, so the host enqueues iterations in chunks and reads the state back
// once per chunk.
//
// Level layout: level 0 is the brick pool (level0.hip), levels 1 and 2 are pooled per brick (pooled.hip), level 3
// (one cell per brick) is dense by brick index, and levels 4 and coarser are dense and run inside one workgroup, in
// LDS (k_coarse_tail below).
//
// Reductions are deterministic and free of cross-workgroup synchronisation: each workgroup writes its FP64
// partial sum, and a one-workgroup finalize kernel sums the partials in a fixed order and updates the state. (A
// grid-wide "last workgroup" ticket needs a device-scope fence per workgroup, which on gfx950's separate per-XCD
// L2 caches costs more than the whole pass.)
#include <algorithm>
#include <cmath>

#include "slosh_ctx.h"

namespace slosh {

namespace {

inline uint32_t blocks_for(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

constexpr int kFirstDense = 4;        // first dense MG level
constexpr uint32_t kTailCells = 4800; // LDS capacity of the tail (16^3 + 8^3 + 4^3 + ...)

__device__ __forceinline__ void dcoord(const DenseLevel &L, uint32_t i, int &x, int &y, int &z)
{
    const uint32_t pc = L.coord[i];
    x = (int)(pc & 1023u);
    y = (int)((pc >> 10) & 1023u);
    z = (int)(pc >> 20);
}

// Index offset of neighbour n (order -x +x -y +y -z +z) in a dense level.
__device__ __forceinline__ int doff(const DenseLevel &L, int n)
{
    const int s = n < 2 ? 1 : (n < 4 ? L.dim[0] : L.dim[0] * L.dim[1]);
    return (n & 1) ? s : -s;
}

template <typename V> __device__ __forceinline__ float dsum(const DenseLevel &L, const V &v, uint32_t i, uint32_t info)
{
    float s = 0.0f;
    for (int n = 0; n < 6; ++n)
        if (info >> n & 1u) s += v[(int)i + doff(L, n)];
    return s;
}

enum FinKind { kFinInit = 0, kFinRho = 1, kFinAlphaStop = 2, kFinRr = 3, kFinSor = 4, kFinBetaF = 5,
               kFinAlphaF = 6, kFinCheck = 7, kFinRestart = 8, kFinErr = 9, kFinCheckDone = 10 };

__host__ __device__ constexpr int fin_values(int kind)
{
    if (kind == kFinAlphaStop) return 4;
    if (kind == kFinInit || kind == kFinRestart) return 3;
    if (kind == kFinSor || kind == kFinAlphaF || kind == kFinCheck || kind == kFinErr || kind == kFinCheckDone) return 2;
    return 1;
}

// Sums the partials of the preceding reduction in a fixed order (per thread: its strided run in increasing order;
// then wave_sum and the wavefront sums in order) and applies them to the solver state.
constexpr uint32_t kFinThreads = 1024;
__device__ __forceinline__ void finalize_body(int kind, const double *partials, uint32_t nblocks, int iters,
                                              SolverState *st)
{
    constexpr int kMaxV = 4;
    constexpr uint32_t kWaves = kFinThreads / 64u;
    const int nv = fin_values(kind);
    __shared__ double sh[kMaxV][kWaves];
    const uint32_t t = threadIdx.x;
    // the first 8 partials of this thread (b = t + kFinThreads k) are read together with the done flag; a thread
    // sums its partials in increasing b (an absent one adds +0, which leaves the sum unchanged: it starts at +0 and
    // never becomes -0)
    constexpr int kFirst = 8;
    const SolverState s0 = *st; // read with the partials; thread 0 updates the state from it
    const uint32_t done = (uint32_t)s0.done;
    double v[kFirst][kMaxV];
    for (int k = 0; k < kFirst; ++k) {
        const uint32_t b = t + kFinThreads * (uint32_t)k;
        for (int j = 0; j < kMaxV; ++j) v[k][j] = (b < nblocks && j < nv) ? partials[(size_t)b * nv + j] : 0.0;
    }
    issue_point();
    // kFinCheckDone: the true-residual check enqueued with every read of the state; it applies once the solve is done
    if (kind == kFinCheckDone && !done) return;
    if (kind != kFinInit && kind != kFinCheck && kind != kFinCheckDone && kind != kFinRestart && kind != kFinErr &&
        done) {
        // the update that met the stopping test has run: from here on every pass returns at once
        if (kind == kFinAlphaStop && done == 2u && t == 0) st->done = 1;
        return;
    }
    double a[kMaxV] = {0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < kFirst; ++k)
        for (int j = 0; j < kMaxV; ++j) a[j] += v[k][j];
    for (uint32_t b = t + kFinThreads * kFirst; b < nblocks; b += kFinThreads)
        for (int j = 0; j < nv; ++j) a[j] += partials[(size_t)b * nv + j];
    for (int j = 0; j < kMaxV; ++j) a[j] = wave_sum(a[j]);
    if ((t & 63u) == 0u)
        for (int j = 0; j < kMaxV; ++j) sh[j][t >> 6] = a[j];
    __syncthreads();
    if (t != 0) return;
    double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
    for (uint32_t w = 0; w < kWaves; ++w) {
        t0 += sh[0][w];
        t1 += sh[1][w];
        t2 += sh[2][w];
        t3 += sh[3][w];
    }
    if (kind == kFinInit || kind == kFinRestart) {
        st->nf = t2;
        st->mg_scale = t1 > 0.0 ? sqrt(t2 / t1) : 1.0;
    }
    if (kind == kFinRestart) {
        // MGPCG restarted from the current pressure with r = b - A p (its true residual exceeded the tolerance
        // although the recurrence residual met it): fresh directions, the iteration count carried on, at least one
        // more iteration
        st->bb = t0;
        st->rr = t1;
        st->rel = t0 > 0.0 ? sqrt(t1) / sqrt(t0) : 0.0;
        st->rho = 0.0;
        st->beta = 0.0;
        st->alpha = 0.0;
        st->converged = 0;
        st->stalled = 0;
        for (int j = 0; j < 4; ++j) st->hist[j] = 0.0;
        st->done = 0;
    } else if (kind == kFinInit || kind == kFinSor) {
        // t0 = ||b||^2, t1 = ||r||^2
        st->bb = t0;
        st->rr = t1;
        const double bn = sqrt(t0), rel = bn > 0.0 ? sqrt(t1) / bn : 0.0;
        st->rel = rel;
        if (kind == kFinInit) {
            st->iters = 0;
            st->rho = 0.0;
            st->beta = 0.0;
            st->alpha = 0.0;
            const int conv = (bn == 0.0 || rel <= s0.tol) ? 1 : 0;
            st->converged = conv;
            st->done = conv;
        } else {
            st->iters = iters;
            if (bn == 0.0 || rel <= s0.tol) {
                st->converged = 1;
                st->done = 1;
            } else if (iters >= s0.maxit) {
                st->done = 1;
            }
        }
    } else if (kind == kFinRho) {
        st->beta = fabs(s0.rho) > 1e-30 ? t0 / s0.rho : 0.0;
        st->rho = t0;
    } else if (kind == kFinAlphaStop) {
        // t0 = d.q, t1 = r.q, t2 = q.q, t3 = r.r of the current residual r: alpha, and the stopping test on
        // ||r - alpha q||^2 = r.r - 2 alpha r.q + alpha^2 q.q (each term summed afresh from this iteration's vectors)
        st->dq = t0;
        const double alpha = fabs(t0) > 1e-30 ? s0.rho / t0 : 0.0;
        st->alpha = alpha;
        const double rr = fmax(t3 - 2.0 * alpha * t1 + alpha * alpha * t2, 0.0);
        st->rr = rr;
        st->mg_scale = rr > 0.0 ? sqrt(s0.nf / rr) : 1.0;
        const double rel = sqrt(rr) / sqrt(s0.bb);
        st->rel = rel;
        const int it = s0.iters + 1;
        st->iters = it;
        if (rel <= s0.tol) {
            st->converged = 1;
            st->done = 2;
        } else if (it >= s0.maxit) {
            st->done = 2;
        }
    } else if (kind == kFinBetaF) {
        // flexible CG: beta = -(z . q_prev) / (d_prev . q_prev)
        st->beta = fabs(s0.dq) > 1e-30 ? -t0 / s0.dq : 0.0;
    } else if (kind == kFinAlphaF) {
        // flexible CG: alpha = (d . r) / (d . q)
        st->dq = t0;
        st->alpha = fabs(t0) > 1e-30 ? t1 / t0 : 0.0;
    } else if (kind == kFinErr) {
        // t0 = sum (p_ref - p)^2, t1 = sum p^2 with p the FP64-accurate solution
        st->err_rel = t1 > 0.0 ? sqrt(t0 / t1) : 0.0;
    } else if (kind == kFinCheck || kind == kFinCheckDone) {
        // the FP64 true residual ||b - A p|| / ||b|| of the solution: the solve counts as converged only when it
        // meets the tolerance (the recurrence residual that stopped the iteration can differ from it)
        const double rel_true = t0 > 0.0 ? sqrt(t1) / sqrt(t0) : 0.0;
        st->rel_true = rel_true;
        st->converged = rel_true <= s0.tol ? 1 : 0;
    } else {
        // kFinRr, the flexible CG's update: its stopping test, and the stall test that hands the solve to MGPCG
        // (a residual that is not finite, the iteration limit, or no halving over the last four iterations)
        st->rr = t0;
        st->mg_scale = t0 > 0.0 ? sqrt(s0.nf / t0) : 1.0;
        const double rel = sqrt(t0) / sqrt(s0.bb);
        st->rel = rel;
        const int it = s0.iters + 1;
        st->iters = it;
        // hist is read through the pointer: an indexed copy in the thread's own state would put the whole state
        // copy of every thread in scratch memory
        const int k = (it - 1) & 3;
        if (!(rel <= 1e300)) {
            st->stalled = 1;
            st->done = 1;
        } else if (rel <= s0.tol) {
            st->converged = 1;
            st->done = 1;
        } else if (it >= s0.maxit || (it - 1 >= 4 && rel > 0.5 * st->hist[k])) {
            st->stalled = 1;
            st->done = 1;
        }
        st->hist[k] = rel;
    }
}

__global__ __launch_bounds__(kFinThreads) void k_finalize(int kind, const double *partials, uint32_t nblocks, int iters,
                                                          SolverState *st)
{
    finalize_body(kind, partials, nblocks, iters, st);
}

// First stage of a large reduction: workgroup g sums partial sets [g chunk, (g + 1) chunk) in a fixed order (per
// thread its strided run in increasing order, then wave_sum and the four wavefront sums in order) into fold[g].
// It returns at once where the finalize it feeds would (a solve already done), like every pass of an iteration.
__global__ __launch_bounds__(256) void k_fold_partials(int kind, const double *partials, uint32_t nblocks,
                                                       uint32_t chunk, double *fold, const SolverState *st)
{
    constexpr int kMaxV = 4;
    __shared__ double sh[kMaxV][4];
    const uint32_t done = st->done;
    issue_point();
    if (kind == kFinCheckDone && !done) return;
    if (kind != kFinInit && kind != kFinCheck && kind != kFinCheckDone && kind != kFinRestart && kind != kFinErr &&
        done)
        return;
    const int nv = fin_values(kind);
    const uint32_t b0 = blockIdx.x * chunk, b1 = min(b0 + chunk, nblocks);
    double a[kMaxV] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t b = b0 + threadIdx.x; b < b1; b += 256u)
        for (int j = 0; j < kMaxV; ++j) a[j] += j < nv ? partials[(size_t)b * nv + j] : 0.0;
    for (int j = 0; j < kMaxV; ++j) a[j] = wave_sum(a[j]);
    if ((threadIdx.x & 63u) == 0u)
        for (int j = 0; j < kMaxV; ++j) sh[j][threadIdx.x >> 6] = a[j];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int j = 0; j < nv; ++j) fold[(size_t)blockIdx.x * nv + j] = ((sh[j][0] + sh[j][1]) + sh[j][2]) + sh[j][3];
}

// The finalize of every tank of a batch (workgroup t: tank t, whose partials are per brick with fluid).
__global__ __launch_bounds__(kFinThreads) void kb_finalize(int kind, const BatchTank *T)
{
    const BatchTank &B = T[blockIdx.x];
    finalize_body(kind, B.partials, B.nsolve, 0, B.state);
}

// ---- dense levels (>= 4) ----------------------------------------------------------------------------------
// SloshPressureMG.usf:256-278 (MGBuildTypeCS): AIR if any child is air, else FLUID if any child is fluid.
__global__ void k_build_type(DenseLevel F, DenseLevel C, int fluid_first)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C.cells) return;
    int x, y, z;
    dcoord(C, i, x, y, z);
    bool air = false, fluid = false;
    for (uint32_t oi = 0; oi < 8u; ++oi) {
        int fx = 2 * x + (int)(oi & 1u), fy = 2 * y + (int)((oi >> 1) & 1u), fz = 2 * z + (int)(oi >> 2);
        if (fx >= F.dim[0] || fy >= F.dim[1] || fz >= F.dim[2]) continue;
        uint32_t t = F.type[(uint32_t)(fx + F.dim[0] * (fy + F.dim[1] * fz))];
        air = air || t == kCellAir;
        fluid = fluid || t == kCellFluid;
    }
    C.type[i] = coarse_type(air, fluid, fluid_first);
}

// Stencil word of a dense level's cells (after its types).
__global__ void k_dinfo(DenseLevel L)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L.cells) return;
    uint32_t info = 0u;
    if (L.type[i] == kCellFluid) {
        int x, y, z;
        dcoord(L, i, x, y, z);
        info = kInfoSelf | (is_red(x, y, z) ? kInfoRed : 0u);
        uint32_t cnt = 0u;
        for (int n = 0; n < 6; ++n) {
            int d[3];
            nb_offset(n, d);
            const int nx = x + d[0], ny = y + d[1], nz = z + d[2];
            if (nx < 0 || ny < 0 || nz < 0 || nx >= L.dim[0] || ny >= L.dim[1] || nz >= L.dim[2]) continue;
            const uint32_t t = L.type[(uint32_t)(nx + L.dim[0] * (ny + L.dim[1] * nz))];
            if (t != kCellSolid) cnt++;
            if (t == kCellFluid) info |= 1u << n;
        }
        info |= cnt << 8;
    }
    L.info[i] = (uint16_t)info;
}

// SloshPressureMG.usf:311-338 (MGSmoothCS), one colour (0 red, 1 black).
__device__ __forceinline__ void smooth_cell(const DenseLevel &L, uint32_t colour, uint32_t i)
{
    const uint32_t info = L.info[i];
    if (!(info & kInfoSelf) || ((info & kInfoRed) ? 0u : 1u) != colour) return;
    L.x[i] = (L.b[i] + dsum(L, L.x, i, info)) / info_diag(info);
}

// SloshPressureMG.usf:343-381 (MGRestrictCS) with the fused red sweep x_c = b_c / diag_c.
__device__ __forceinline__ void restrict_cell(const DenseLevel &F, const DenseLevel &C, uint32_t ci)
{
    const uint32_t cinfo = C.info[ci];
    if (!(cinfo & kInfoSelf)) {
        C.b[ci] = 0.0f;
        C.x[ci] = 0.0f;
        return;
    }
    int cx, cy, cz;
    dcoord(C, ci, cx, cy, cz);
    float sum = 0.0f;
    for (uint32_t oi = 0; oi < 8u; ++oi) {
        int fx = 2 * cx + (int)(oi & 1u), fy = 2 * cy + (int)((oi >> 1) & 1u), fz = 2 * cz + (int)(oi >> 2);
        if (fx >= F.dim[0] || fy >= F.dim[1] || fz >= F.dim[2]) continue;
        const uint32_t fi = (uint32_t)(fx + F.dim[0] * (fy + F.dim[1] * fz));
        const uint32_t finfo = F.info[fi];
        if (!(finfo & kInfoSelf)) continue;
        sum += F.b[fi] - (info_diag(finfo) * F.x[fi] - dsum(F, F.x, fi, finfo));
    }
    const float bv = 0.25f * sum;
    C.b[ci] = bv;
    C.x[ci] = (cinfo & kInfoRed) ? bv / info_diag(cinfo) : 0.0f;
}

// SloshPressureMG.usf:418-433 (MGProlongCS)
__device__ __forceinline__ void prolong_cell(const DenseLevel &F, const DenseLevel &C, uint32_t i)
{
    if (!(F.info[i] & kInfoSelf)) return;
    int x, y, z;
    dcoord(F, i, x, y, z);
    uint32_t pi = (uint32_t)((x >> 1) + C.dim[0] * ((y >> 1) + C.dim[1] * (z >> 1)));
    if (C.info[pi] & kInfoSelf) F.x[i] += C.x[pi];
}

__global__ void k_dsmooth(DenseLevel L, uint32_t colour, const SolverState *st)
{
    if (st->done) return;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < L.cells) smooth_cell(L, colour, i);
}

__global__ void k_drestrict(DenseLevel F, DenseLevel C, const SolverState *st)
{
    if (st->done) return;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < C.cells) restrict_cell(F, C, i);
}

// ---- levels 2 and 3 fused: one workgroup per 4^3 tile of level 3 holding fluid (the tile list is built once per
// solve). Level 2 is pooled (2^3 cells per brick slot) and is gathered through the brick table; level 3 is dense by
// brick index. Each pass stages the halos its sweeps need and recomputes there what neighbouring tiles also compute,
// with the same expressions and order as the separate level-2 and level-3 passes.
constexpr int kT3 = 4;                          // level-3 tile edge
constexpr int kDB2 = 2 * kT3 + 12;              // down: level-2 box edge (the tile's 8^3 cells and a 6-cell halo)
constexpr int kDNB2 = kDB2 * kDB2 * kDB2;
constexpr int kDBB = kT3 + 6;                   // down: bricks per edge under the level-2 box
constexpr int kDR3 = kT3 + 4;                   // down: level-3 region edge (the tile and a 2-cell halo)
constexpr int kDR2 = 2 * kDR3;                  // down: level-2 children of the level-3 region
constexpr uint32_t kD23DownThreads = 1024;
constexpr int kUB3 = kT3 + 6;                   // up: level-3 box edge (a 3-cell halo)
constexpr int kUB2 = 2 * kT3 + 4;               // up: level-2 box edge (a 2-cell halo)
constexpr int kUBB = kT3 + 2;                   // up: bricks per edge under the level-2 box
constexpr uint32_t kD23UpThreads = 512;

__device__ __forceinline__ void tile_of(const DenseLevel &L3, uint32_t tile, int &tx, int &ty, int &tz)
{
    const int ntx = (L3.dim[0] + kT3 - 1) / kT3, nty = (L3.dim[1] + kT3 - 1) / kT3;
    tx = (int)tile % ntx;
    ty = ((int)tile / ntx) % nty;
    tz = (int)tile / (ntx * nty);
}

// Offset of neighbour n (order -x +x -y +y -z +z) in a box of edge E.
template <int E> __device__ __forceinline__ int box_off(int n)
{
    const int s = n < 2 ? 1 : (n < 4 ? E : E * E);
    return (n & 1) ? s : -s;
}

template <int E> __device__ __forceinline__ float box_nbsum(const float *v, int k, uint32_t info)
{
    float s = 0.0f;
    for (int n = 0; n < 6; ++n)
        if (info >> n & 1u) s += v[k + box_off<E>(n)];
    return s;
}

// Brick slots under a box of level-2 cells starting at brick (b0x, b0y, b0z), NB bricks per edge (0 outside the
// grid or unallocated).
template <int NB>
__device__ __forceinline__ void stage_slots(const Geo &g, int b0x, int b0y, int b0z, uint32_t *s_slot, uint32_t nt)
{
    for (uint32_t k = threadIdx.x; k < (uint32_t)(NB * NB * NB); k += nt) {
        const int bx = b0x + (int)(k % NB), by = b0y + (int)((k / NB) % NB), bz = b0z + (int)(k / (NB * NB));
        const bool in = bx >= 0 && by >= 0 && bz >= 0 && bx < g.bd[0] && by < g.bd[1] && bz < g.bd[2];
        s_slot[k] = in ? g.table[(uint32_t)(bx + g.bd[0] * (by + g.bd[1] * bz))] : 0u;
    }
}

// Pool index of level-2 cell (x, y, z) through the staged slots (bricks from (b0x, b0y, b0z), NB per edge); cells
// outside level 2 read the null slot, whose words are 0.
template <int NB>
__device__ __forceinline__ uint32_t l2_index(const PLevel &L2, const uint32_t *s_slot, int b0x, int b0y, int b0z, int x,
                                             int y, int z)
{
    if (!pinside(L2, x, y, z)) return 0u;
    const uint32_t slot = s_slot[((x >> 1) - b0x) + NB * (((y >> 1) - b0y) + NB * ((z >> 1) - b0z))];
    return slot * 8u + (uint32_t)((x & 1) + 2 * ((y & 1) + 2 * (z & 1)));
}

// Down pass of levels 2 and 3 for one tile: the level-2 black sweep (MGSmoothCS) and residual on the children of the
// tile's level-3 cells and their 2-cell halo, the restriction into level 3 with its fused red sweep (MGRestrictCS),
// the level-3 black sweep and residual, and the restriction into the tile's 2^3 level-4 cells. Stored: the level-2
// black values of the tile's cells, level 3's b and x on the tile, level 4's b and x.
__device__ __forceinline__ void d23down_body(const Blk &G, Geo g, const uint32_t *tiles, PLevel L2, DenseLevel L3,
                                             DenseLevel L4, const SolverState *st)
{
    constexpr uint32_t NT = kD23DownThreads;
    constexpr int kPer = (kDNB2 + (int)NT - 1) / (int)NT;
    __shared__ uint32_t s_slot[kDBB * kDBB * kDBB];
    __shared__ float s_x[kDNB2], s_b[kDNB2];
    __shared__ uint16_t s_info[kDNB2];
    __shared__ float s_r2[kDR2 * kDR2 * kDR2];
    __shared__ float s_x3[kDR3 * kDR3 * kDR3], s_b3[kDR3 * kDR3 * kDR3];
    __shared__ uint16_t s_i3[kDR3 * kDR3 * kDR3];
    __shared__ float s_r3[kT3 * kT3 * kT3];
    const uint32_t t = threadIdx.x;
    const uint32_t tile = tiles[G.b];
    const uint32_t done = st->done;
    issue_point();
    if (done) return;
    int tx, ty, tz;
    tile_of(L3, tile, tx, ty, tz);
    const int o2x = 8 * tx - 6, o2y = 8 * ty - 6, o2z = 8 * tz - 6;  // level-2 box origin
    const int bbx = 4 * tx - 3, bby = 4 * ty - 3, bbz = 4 * tz - 3;  // first brick under it
    const int o3x = 4 * tx - 2, o3y = 4 * ty - 2, o3z = 4 * tz - 2;  // level-3 region origin
    stage_slots<kDBB>(g, bbx, bby, bbz, s_slot, NT);
    __syncthreads();
    uint32_t vi[kPer];
    float vx[kPer], vb[kPer];
    for (int q = 0; q < kPer; ++q) {
        const int k = (int)t + (int)NT * q;
        vi[q] = 0u;
        vx[q] = vb[q] = 0.0f;
        if (k >= kDNB2) continue;
        const uint32_t gi = l2_index<kDBB>(L2, s_slot, bbx, bby, bbz, o2x + k % kDB2, o2y + (k / kDB2) % kDB2,
                                           o2z + k / (kDB2 * kDB2));
        vi[q] = L2.info[gi];
        vx[q] = L2.x[gi];
        vb[q] = L2.b[gi];
    }
    uint32_t wi3 = 0u;
    if (t < (uint32_t)(kDR3 * kDR3 * kDR3)) {
        const int x = o3x + (int)t % kDR3, y = o3y + ((int)t / kDR3) % kDR3, z = o3z + (int)t / (kDR3 * kDR3);
        if (x >= 0 && y >= 0 && z >= 0 && x < L3.dim[0] && y < L3.dim[1] && z < L3.dim[2])
            wi3 = L3.info[(uint32_t)(x + L3.dim[0] * (y + L3.dim[1] * z))];
    }
    // the tile's level-4 cells: thread t < 8 takes child block t
    const int cx = 2 * tx + (int)(t & 1u), cy = 2 * ty + (int)((t >> 1) & 1u), cz = 2 * tz + (int)((t >> 2) & 1u);
    const bool cin = t < 8u && cx < L4.dim[0] && cy < L4.dim[1] && cz < L4.dim[2];
    const uint32_t ci = cin ? (uint32_t)(cx + L4.dim[0] * (cy + L4.dim[1] * cz)) : 0u;
    const uint32_t cinfo = L4.info[ci];
    for (int q = 0; q < kPer; ++q) {
        const int k = (int)t + (int)NT * q;
        if (k >= kDNB2) continue;
        const bool f = (vi[q] & kInfoSelf) != 0u;
        s_info[k] = (uint16_t)vi[q];
        s_x[k] = f ? vx[q] : 0.0f;
        s_b[k] = f ? vb[q] : 0.0f;
    }
    if (t < (uint32_t)(kDR3 * kDR3 * kDR3)) s_i3[t] = (uint16_t)wi3;
    __syncthreads();
    // level 2: black sweep on the box less a 1-cell rim
    for (int k = (int)t; k < (kDB2 - 2) * (kDB2 - 2) * (kDB2 - 2); k += (int)NT) {
        constexpr int E = kDB2 - 2;
        const int kb = (1 + k % E) + kDB2 * ((1 + (k / E) % E) + kDB2 * (1 + k / (E * E)));
        const uint32_t info = s_info[kb];
        if (!(info & kInfoSelf) || (info & kInfoRed)) continue;
        s_x[kb] = (s_b[kb] + box_nbsum<kDB2>(s_x, kb, info)) / info_diag(info);
    }
    __syncthreads();
    // level 2: residual of the children of the level-3 region (the box less a 2-cell rim); black values of the
    // tile's own level-2 cells go back to memory
    for (int k = (int)t; k < kDR2 * kDR2 * kDR2; k += (int)NT) {
        const int u = 2 + k % kDR2, v = 2 + (k / kDR2) % kDR2, w = 2 + k / (kDR2 * kDR2);
        const int kb = u + kDB2 * (v + kDB2 * w);
        const uint32_t info = s_info[kb];
        float r = 0.0f;
        if (info & kInfoSelf) {
            if (!(info & kInfoRed) && u >= 6 && u < 14 && v >= 6 && v < 14 && w >= 6 && w < 14)
                L2.x[l2_index<kDBB>(L2, s_slot, bbx, bby, bbz, o2x + u, o2y + v, o2z + w)] = s_x[kb];
            r = s_b[kb] - (info_diag(info) * s_x[kb] - box_nbsum<kDB2>(s_x, kb, info));
        }
        s_r2[k] = r;
    }
    __syncthreads();
    // restriction into the level-3 region, in child order, with the fused red sweep
    if (t < (uint32_t)(kDR3 * kDR3 * kDR3)) {
        const int i = (int)t % kDR3, j = ((int)t / kDR3) % kDR3, l = (int)t / (kDR3 * kDR3);
        const uint32_t c3 = s_i3[t];
        float bv = 0.0f, xv = 0.0f;
        if (c3 & kInfoSelf) {
            float sum = 0.0f;
            for (uint32_t oi = 0; oi < 8u; ++oi) {
                const int fu = 2 * i + (int)(oi & 1u), fv = 2 * j + (int)((oi >> 1) & 1u), fw = 2 * l + (int)(oi >> 2);
                if (s_info[(2 + fu) + kDB2 * ((2 + fv) + kDB2 * (2 + fw))] & kInfoSelf)
                    sum += s_r2[fu + kDR2 * (fv + kDR2 * fw)];
            }
            bv = 0.25f * sum;
            xv = (c3 & kInfoRed) ? bv / info_diag(c3) : 0.0f;
        }
        s_b3[t] = bv;
        s_x3[t] = xv;
    }
    __syncthreads();
    // level 3: black sweep on the region less a 1-cell rim
    if (t < (uint32_t)((kDR3 - 2) * (kDR3 - 2) * (kDR3 - 2))) {
        constexpr int E = kDR3 - 2;
        const int k3 = (1 + (int)t % E) + kDR3 * ((1 + ((int)t / E) % E) + kDR3 * (1 + (int)t / (E * E)));
        const uint32_t info = s_i3[k3];
        if ((info & kInfoSelf) && !(info & kInfoRed))
            s_x3[k3] = (s_b3[k3] + box_nbsum<kDR3>(s_x3, k3, info)) / info_diag(info);
    }
    __syncthreads();
    // level 3: residual of the tile; the tile's b and x go back to memory
    if (t < (uint32_t)(kT3 * kT3 * kT3)) {
        const int u = (int)t % kT3, v = ((int)t / kT3) % kT3, w = (int)t / (kT3 * kT3);
        const int k3 = (2 + u) + kDR3 * ((2 + v) + kDR3 * (2 + w));
        const uint32_t info = s_i3[k3];
        const int x = 4 * tx + u, y = 4 * ty + v, z = 4 * tz + w;
        float r = 0.0f;
        if (x < L3.dim[0] && y < L3.dim[1] && z < L3.dim[2]) {
            const uint32_t gi = (uint32_t)(x + L3.dim[0] * (y + L3.dim[1] * z));
            L3.b[gi] = s_b3[k3];
            L3.x[gi] = s_x3[k3];
            if (info & kInfoSelf) r = s_b3[k3] - (info_diag(info) * s_x3[k3] - box_nbsum<kDR3>(s_x3, k3, info));
        }
        s_r3[t] = r;
    }
    __syncthreads();
    if (!cin) return;
    if (!(cinfo & kInfoSelf)) {
        L4.b[ci] = 0.0f;
        L4.x[ci] = 0.0f;
        return;
    }
    // children in child order; one outside level 3 has residual 0 (adding +0 leaves the sum, which starts at +0
    // and never becomes -0, unchanged)
    float sum = 0.0f;
    for (uint32_t oi = 0; oi < 8u; ++oi) {
        const int u = 2 * (int)(t & 1u) + (int)(oi & 1u), v = 2 * (int)((t >> 1) & 1u) + (int)((oi >> 1) & 1u),
                  w = 2 * (int)((t >> 2) & 1u) + (int)(oi >> 2);
        sum += s_r3[u + kT3 * (v + kT3 * w)];
    }
    const float bv = 0.25f * sum;
    L4.b[ci] = bv;
    L4.x[ci] = (cinfo & kInfoRed) ? bv / info_diag(cinfo) : 0.0f;
}

// Up pass of levels 3 and 2 for one tile (MGProlongCS, then MGSmoothCS black and red on each level): level 3's
// prolonged iterate x + x4(parent) on the tile and a 3-cell halo, its black sweep and red sweep, which give level 3's
// result on the tile and a 1-cell halo; level 2's prolonged iterate from it, its black sweep and red sweep; level 2's
// result x1 on the tile's cells goes back to memory.
__device__ __forceinline__ void d23up_body(const Blk &G, Geo g, const uint32_t *tiles, PLevel L2, DenseLevel L3,
                                           DenseLevel L4, const SolverState *st)
{
    constexpr uint32_t NT = kD23UpThreads;
    constexpr int kN3 = kUB3 * kUB3 * kUB3, kN2 = kUB2 * kUB2 * kUB2;
    constexpr int kPer3 = (kN3 + (int)NT - 1) / (int)NT, kPer2 = (kN2 + (int)NT - 1) / (int)NT;
    __shared__ uint32_t s_slot[kUBB * kUBB * kUBB];
    __shared__ float s_v3[kN3], s_b3[kN3];
    __shared__ uint16_t s_i3[kN3];
    __shared__ float s_v2[kN2], s_b2[kN2];
    __shared__ uint16_t s_i2[kN2];
    const uint32_t t = threadIdx.x;
    const uint32_t tile = tiles[G.b];
    const uint32_t done = st->done;
    issue_point();
    if (done) return;
    int tx, ty, tz;
    tile_of(L3, tile, tx, ty, tz);
    const int o3x = 4 * tx - 3, o3y = 4 * ty - 3, o3z = 4 * tz - 3;  // level-3 box origin
    const int o2x = 8 * tx - 2, o2y = 8 * ty - 2, o2z = 8 * tz - 2;  // level-2 box origin
    const int bbx = 4 * tx - 1, bby = 4 * ty - 1, bbz = 4 * tz - 1;  // first brick under it
    stage_slots<kUBB>(g, bbx, bby, bbz, s_slot, NT);
    uint32_t wi3[kPer3], wp3[kPer3];
    float wx3[kPer3], wb3[kPer3], wpx[kPer3];
    for (int q = 0; q < kPer3; ++q) {
        const int k = (int)t + (int)NT * q;
        wi3[q] = wp3[q] = 0u;
        wx3[q] = wb3[q] = wpx[q] = 0.0f;
        if (k >= kN3) continue;
        const int x = o3x + k % kUB3, y = o3y + (k / kUB3) % kUB3, z = o3z + k / (kUB3 * kUB3);
        const bool in = x >= 0 && y >= 0 && z >= 0 && x < L3.dim[0] && y < L3.dim[1] && z < L3.dim[2];
        const uint32_t gi = in ? (uint32_t)(x + L3.dim[0] * (y + L3.dim[1] * z)) : 0u;
        const uint32_t pi = in ? (uint32_t)((x >> 1) + L4.dim[0] * ((y >> 1) + L4.dim[1] * (z >> 1))) : 0u;
        wi3[q] = in ? (uint32_t)L3.info[gi] : 0u;
        wx3[q] = L3.x[gi];
        wb3[q] = L3.b[gi];
        wp3[q] = L4.info[pi];
        wpx[q] = L4.x[pi];
    }
    __syncthreads();
    uint32_t wi2[kPer2];
    float wx2[kPer2], wb2[kPer2];
    for (int q = 0; q < kPer2; ++q) {
        const int k = (int)t + (int)NT * q;
        wi2[q] = 0u;
        wx2[q] = wb2[q] = 0.0f;
        if (k >= kN2) continue;
        const uint32_t gi = l2_index<kUBB>(L2, s_slot, bbx, bby, bbz, o2x + k % kUB2, o2y + (k / kUB2) % kUB2,
                                           o2z + k / (kUB2 * kUB2));
        wi2[q] = L2.info[gi];
        wx2[q] = L2.x[gi];
        wb2[q] = L2.b[gi];
    }
    for (int q = 0; q < kPer3; ++q) {
        const int k = (int)t + (int)NT * q;
        if (k >= kN3) continue;
        float val = 0.0f, bv = 0.0f;
        if (wi3[q] & kInfoSelf) {
            val = wx3[q];
            if (wp3[q] & kInfoSelf) val += wpx[q];
            bv = wb3[q];
        }
        s_i3[k] = (uint16_t)wi3[q];
        s_v3[k] = val;
        s_b3[k] = bv;
    }
    for (int q = 0; q < kPer2; ++q) {
        const int k = (int)t + (int)NT * q;
        if (k >= kN2) continue;
        const bool f = (wi2[q] & kInfoSelf) != 0u;
        s_i2[k] = (uint16_t)wi2[q];
        s_v2[k] = f ? wx2[q] : 0.0f;
        s_b2[k] = f ? wb2[q] : 0.0f;
    }
    __syncthreads();
    // level 3: black sweep on the box less a 1-cell rim, then red sweep on the box less a 2-cell rim (in place:
    // a red cell reads only black neighbours)
    for (int k = (int)t; k < (kUB3 - 2) * (kUB3 - 2) * (kUB3 - 2); k += (int)NT) {
        constexpr int E = kUB3 - 2;
        const int kb = (1 + k % E) + kUB3 * ((1 + (k / E) % E) + kUB3 * (1 + k / (E * E)));
        const uint32_t info = s_i3[kb];
        if ((info & kInfoSelf) && !(info & kInfoRed))
            s_v3[kb] = (s_b3[kb] + box_nbsum<kUB3>(s_v3, kb, info)) / info_diag(info);
    }
    __syncthreads();
    for (int k = (int)t; k < (kUB3 - 4) * (kUB3 - 4) * (kUB3 - 4); k += (int)NT) {
        constexpr int E = kUB3 - 4;
        const int kb = (2 + k % E) + kUB3 * ((2 + (k / E) % E) + kUB3 * (2 + k / (E * E)));
        const uint32_t info = s_i3[kb];
        if ((info & kInfoSelf) && (info & kInfoRed))
            s_v3[kb] = (s_b3[kb] + box_nbsum<kUB3>(s_v3, kb, info)) / info_diag(info);
    }
    __syncthreads();
    // level 2: prolonged iterate (a parent contributes only when it is fluid), black sweep on the box less a
    // 1-cell rim, red sweep on the tile
    for (int k = (int)t; k < kN2; k += (int)NT) {
        if (!(s_i2[k] & kInfoSelf)) continue;
        const int x = o2x + k % kUB2, y = o2y + (k / kUB2) % kUB2, z = o2z + k / (kUB2 * kUB2);
        const int k3 = ((x >> 1) - o3x) + kUB3 * (((y >> 1) - o3y) + kUB3 * ((z >> 1) - o3z));
        if (s_i3[k3] & kInfoSelf) s_v2[k] += s_v3[k3];
    }
    __syncthreads();
    for (int k = (int)t; k < (kUB2 - 2) * (kUB2 - 2) * (kUB2 - 2); k += (int)NT) {
        constexpr int E = kUB2 - 2;
        const int kb = (1 + k % E) + kUB2 * ((1 + (k / E) % E) + kUB2 * (1 + k / (E * E)));
        const uint32_t info = s_i2[kb];
        if ((info & kInfoSelf) && !(info & kInfoRed))
            s_v2[kb] = (s_b2[kb] + box_nbsum<kUB2>(s_v2, kb, info)) / info_diag(info);
    }
    __syncthreads();
    for (int k = (int)t; k < (kUB2 - 4) * (kUB2 - 4) * (kUB2 - 4); k += (int)NT) {
        constexpr int E = kUB2 - 4;
        const int u = 2 + k % E, v = 2 + (k / E) % E, w = 2 + k / (E * E);
        const int kb = u + kUB2 * (v + kUB2 * w);
        const uint32_t info = s_i2[kb];
        if (!(info & kInfoSelf)) continue;
        L2.x1[l2_index<kUBB>(L2, s_slot, bbx, bby, bbz, o2x + u, o2y + v, o2z + w)] =
            (info & kInfoRed) ? (s_b2[kb] + box_nbsum<kUB2>(s_v2, kb, info)) / info_diag(info) : s_v2[kb];
    }
}

__global__ __launch_bounds__(kD23DownThreads) void k_d23down(Geo g, const uint32_t *tiles, PLevel L2, DenseLevel L3,
                                                             DenseLevel L4, const SolverState *st)
{
    d23down_body(this_blk(), g, tiles, L2, L3, L4, st);
}

__global__ __launch_bounds__(kD23UpThreads) void k_d23up(Geo g, const uint32_t *tiles, PLevel L2, DenseLevel L3,
                                                         DenseLevel L4, const SolverState *st)
{
    d23up_body(this_blk(), g, tiles, L2, L3, L4, st);
}

// The fused level-2/3 passes over a batch of tanks (one level-3 tile per workgroup).
__global__ __launch_bounds__(kD23DownThreads) void kb_d23down(const BatchTank *T, const uint32_t *range, uint32_t nt)
{
    Blk G;
    const BatchTank &B = T[batch_tank(range, nt, blockIdx.x, G)];
    d23down_body(G, B.geo, B.tile_list, B.L2, B.L3, B.L4, B.state);
}

__global__ __launch_bounds__(kD23UpThreads) void kb_d23up(const BatchTank *T, const uint32_t *range, uint32_t nt)
{
    Blk G;
    const BatchTank &B = T[batch_tank(range, nt, blockIdx.x, G)];
    d23up_body(G, B.geo, B.tile_list, B.L2, B.L3, B.L4, B.state);
}

// Tiles of level 3 whose bricks hold fluid cells (fluid on levels 2 and 3 needs a fluid cell below it in the same
// brick under either coarsening rule): flag per tile, set by the bricks (the flags start at 0).
__global__ void k_tile_flags(Geo g, uint32_t nactive, const uint32_t *fluid_flag, DenseLevel L3, uint32_t *flag)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nactive || !fluid_flag[e]) return;
    int bc[3];
    brick_coord(g, g.active[e], bc);
    const int ntx = (L3.dim[0] + kT3 - 1) / kT3, nty = (L3.dim[1] + kT3 - 1) / kT3;
    flag[(uint32_t)(bc[0] / kT3 + ntx * (bc[1] / kT3 + nty * (bc[2] / kT3)))] = 1u;
}

__global__ void k_tile_list(uint32_t ntiles, const uint32_t *flag, const uint32_t *rank, uint32_t *list)
{
    const uint32_t tile = blockIdx.x * blockDim.x + threadIdx.x;
    if (tile < ntiles && flag[tile]) list[rank[tile]] = tile;
}


__global__ void k_dprolong(DenseLevel F, DenseLevel C, const SolverState *st)
{
    if (st->done) return;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < F.cells) prolong_cell(F, C, i);
}


// Cells per thread of any tail pass (1024 threads).
constexpr int kTailPer = (int)((kTailCells + 1023u) / 1024u);

// LDS image of the tail: per cell the stencil word, packed coordinates, iterate, right-hand side, residual and
// diagonal (info_diag of the word, summed once at staging); per level the LDS indices of its fluid cells, red ones
// from the front of the level's segment of `list`, black ones from its back, and their counts cnt[2 level + colour].
// The passes visit fluid cells only. Which thread takes which cell does not matter: a colour sweep reads only the
// other colour, and every other pass writes each output from values it does not write.
struct TailLds {
    uint16_t *info;
    uint32_t *coord;
    float *x, *b, *r, *dg;
    uint16_t *list;
    uint32_t *cnt;
};

// A tail level as the LDS passes see it (uniform across the workgroup).
struct TailLevel {
    uint32_t off, cells;
    int dx, dy, dz;
    int k; // index in the tail
};

__device__ __forceinline__ TailLevel tail_level(const TailArgs &T, int k)
{
    TailLevel V;
    V.off = T.off[k];
    V.cells = T.L[k].cells;
    V.dx = T.L[k].dim[0];
    V.dy = T.L[k].dim[1];
    V.dz = T.L[k].dim[2];
    V.k = k;
    return V;
}

// LDS index of the j-th fluid cell of colour c (0 red, 1 black) of level V.
__device__ __forceinline__ uint32_t tail_cell(const TailLds &S, const TailLevel &V, uint32_t c, uint32_t j)
{
    return S.list[c ? V.off + V.cells - 1u - j : V.off + j];
}

// LDS index of neighbour n (order -x +x -y +y -z +z) of cell k, or k itself when the word has no fluid neighbour
// there (so every load address is valid and the value is masked out).
__device__ __forceinline__ uint32_t tail_nbr(const TailLevel &V, uint32_t k, uint32_t info, int n)
{
    const int s = n < 2 ? 1 : (n < 4 ? V.dx : V.dx * V.dy);
    const int d = (n & 1) ? s : -s;
    return (info >> n & 1u) ? (uint32_t)((int)k + d) : k;
}

// Sum over the fluid neighbours of cell k of the LDS iterate, in neighbour order.
__device__ __forceinline__ float tail_nbsum(const TailLds &S, const TailLevel &V, uint32_t k, uint32_t info)
{
    float v[6];
    for (int n = 0; n < 6; ++n) v[n] = S.x[tail_nbr(V, k, info, n)];
    float s = 0.0f;
    for (int n = 0; n < 6; ++n)
        if (info >> n & 1u) s += v[n];
    return s;
}

// One Gauss-Seidel colour sweep (SloshPressureMG.usf:311-338) over the level's fluid cells of that colour.
__device__ __forceinline__ void tail_sweep(const TailLds &S, const TailLevel &V, uint32_t colour)
{
    const uint32_t n = S.cnt[2 * V.k + (int)colour];
    for (uint32_t j = threadIdx.x; j < n; j += 1024u) {
        const uint32_t k = tail_cell(S, V, colour, j), info = S.info[k];
        S.x[k] = (S.b[k] + tail_nbsum(S, V, k, info)) / S.dg[k];
    }
}

// Residual b - A x of the level's fluid cells, the children's values of the next restriction.
__device__ __forceinline__ void tail_residual(const TailLds &S, const TailLevel &V)
{
    const uint32_t n0 = S.cnt[2 * V.k], n = n0 + S.cnt[2 * V.k + 1];
    for (uint32_t j = threadIdx.x; j < n; j += 1024u) {
        const uint32_t k = j < n0 ? tail_cell(S, V, 0u, j) : tail_cell(S, V, 1u, j - n0), info = S.info[k];
        S.r[k] = S.b[k] - (S.dg[k] * S.x[k] - tail_nbsum(S, V, k, info));
    }
}

// SloshPressureMG.usf:343-381 (MGRestrictCS) from the staged residuals of the fluid children, with the fused red
// sweep x_c = b_c / diag_c, for the coarse level's fluid cells (the values of other cells are never read).
__device__ __forceinline__ void tail_restrict(const TailLds &S, const TailLevel &F, const TailLevel &C)
{
    const uint32_t n0 = S.cnt[2 * C.k], n = n0 + S.cnt[2 * C.k + 1];
    for (uint32_t j = threadIdx.x; j < n; j += 1024u) {
        const uint32_t kc = j < n0 ? tail_cell(S, C, 0u, j) : tail_cell(S, C, 1u, j - n0);
        const uint32_t cinfo = S.info[kc], pc = S.coord[kc];
        const int cx = (int)(pc & 1023u), cy = (int)((pc >> 10) & 1023u), cz = (int)(pc >> 20);
        float rv[8];
        bool fl[8];
        for (uint32_t oi = 0; oi < 8u; ++oi) {
            const int fx = 2 * cx + (int)(oi & 1u), fy = 2 * cy + (int)((oi >> 1) & 1u), fz = 2 * cz + (int)(oi >> 2);
            const bool in = fx < F.dx && fy < F.dy && fz < F.dz;
            const uint32_t fi = in ? F.off + (uint32_t)(fx + F.dx * (fy + F.dy * fz)) : kc;
            rv[oi] = S.r[fi];
            fl[oi] = in && (S.info[fi] & kInfoSelf);
        }
        float sum = 0.0f;
        for (uint32_t oi = 0; oi < 8u; ++oi)
            if (fl[oi]) sum += rv[oi];
        const float bv = 0.25f * sum;
        S.b[kc] = bv;
        S.x[kc] = (cinfo & kInfoRed) ? bv / S.dg[kc] : 0.0f;
    }
}

// SloshPressureMG.usf:418-433 (MGProlongCS): x_f += x_c(parent) for the fluid cells with a fluid parent.
__device__ __forceinline__ void tail_prolong(const TailLds &S, const TailLevel &F, const TailLevel &C)
{
    const uint32_t n0 = S.cnt[2 * F.k], n = n0 + S.cnt[2 * F.k + 1];
    for (uint32_t j = threadIdx.x; j < n; j += 1024u) {
        const uint32_t k = j < n0 ? tail_cell(S, F, 0u, j) : tail_cell(S, F, 1u, j - n0), pc = S.coord[k];
        const int x = (int)(pc & 1023u), y = (int)((pc >> 10) & 1023u), z = (int)(pc >> 20);
        const uint32_t pi = C.off + (uint32_t)((x >> 1) + C.dx * ((y >> 1) + C.dy * (z >> 1)));
        if (S.info[pi] & kInfoSelf) S.x[k] += S.x[pi];
    }
}

// The dense levels of the V-cycle (SloshPressureMG.usf via SloshProxy.cpp:904-959) in one 1024-thread workgroup,
// entirely in LDS, from the first tail level's b and red-initialised x: `sweeps` red-black pairs on each level on
// the way down (the first red sweep fused into the restriction), residuals and restrictions down to the coarsest
// level, the exact FP64 coarsest solve, then prolongation with `sweeps` black-red pairs back up. Every global read
// is issued at the start (solver flag, stencil words, coordinates, the coarsest inverse, the first level's inputs);
// the passes then visit only fluid cells, from per-level lists built once; only the first level's result goes back
// to memory.
__device__ __forceinline__ void coarse_tail_body(const TailArgs &T, const uint32_t *cidx, const uint32_t *cn,
                                                 const double *inv, const SolverState *st)
{
    constexpr int N = (int)kCoarsestCells;
    __shared__ uint16_t s_info[kTailCells];
    __shared__ uint32_t s_coord[kTailCells];
    __shared__ float s_x[kTailCells], s_b[kTailCells], s_r[kTailCells], s_dg[kTailCells];
    __shared__ uint16_t s_list[kTailCells];
    __shared__ uint32_t s_cnt[2 * kMaxLevels];
    __shared__ double bs[N];
    const TailLds S{s_info, s_coord, s_x, s_b, s_r, s_dg, s_list, s_cnt};
    const uint32_t tid = threadIdx.x;
    const int m = T.n;
    const uint32_t done = st->done;
    const int row = (int)(tid >> 4), part = (int)(tid & 15u);
    const int n = (int)*cn;
    double invr[N / 16];
    for (int k = 0; k < N / 16; ++k) invr[k] = inv[(size_t)row * N + part + 16 * k];
    const uint32_t c_row = cidx[row], c_tid = cidx[tid & (uint32_t)(N - 1)];
    uint16_t vi[kTailPer];
    uint32_t vc[kTailPer];
    for (int u = 0; u < kTailPer; ++u) {
        const uint32_t f = tid + 1024u * (uint32_t)u;
        vi[u] = f < T.total ? T.info[f] : (uint16_t)0;
        vc[u] = f < T.total ? T.coord[f] : 0u;
    }
    // the first tail level's b and red-initialised x (cells tid + 1024 u)
    const TailLevel V0 = tail_level(T, 0);
    float gx[kTailPer], gb[kTailPer];
    for (int u = 0; u < kTailPer; ++u) {
        const uint32_t ci = tid + 1024u * (uint32_t)u;
        gx[u] = gb[u] = 0.0f;
        if (ci >= V0.cells) continue;
        gx[u] = T.L[0].x[ci];
        gb[u] = T.L[0].b[ci];
    }
    issue_point();
    if (done) return;
    if (tid < 2u * kMaxLevels) s_cnt[tid] = 0u;
    for (int u = 0; u < kTailPer; ++u) {
        const uint32_t f = tid + 1024u * (uint32_t)u;
        if (f >= T.total) continue;
        s_info[f] = vi[u];
        s_coord[f] = vc[u];
        s_dg[f] = info_diag(vi[u]);
    }
    for (int u = 0; u < kTailPer; ++u) {
        const uint32_t ci = tid + 1024u * (uint32_t)u;
        if (ci >= V0.cells) continue;
        s_b[ci] = gb[u];
        s_x[ci] = gx[u];
    }
    __syncthreads();
    // fluid cells of each level and colour (the order within a list does not change any result)
    for (int l = 0; l < m; ++l) {
        const TailLevel V = tail_level(T, l);
        for (uint32_t i = tid; i < V.cells; i += 1024u) {
            const uint32_t f = V.off + i, info = s_info[f];
            if (!(info & kInfoSelf)) continue;
            if (info & kInfoRed) s_list[V.off + atomicAdd(&s_cnt[2 * l], 1u)] = (uint16_t)f;
            else s_list[V.off + V.cells - 1u - atomicAdd(&s_cnt[2 * l + 1], 1u)] = (uint16_t)f;
        }
    }
    __syncthreads();
    for (int l = 0; l < m - 1; ++l) {
        const TailLevel F = tail_level(T, l), C = tail_level(T, l + 1);
        for (int k = 0; k < T.sweeps; ++k) {
            if (k > 0) {
                tail_sweep(S, F, 0u);
                __syncthreads();
            }
            tail_sweep(S, F, 1u);
            __syncthreads();
        }
        tail_residual(S, F);
        __syncthreads();
        tail_restrict(S, F, C);
        __syncthreads();
    }
    {
        // x = A^-1 b on the coarsest level: 16 lanes per row, partial sums combined by a fixed butterfly
        const TailLevel Vc = tail_level(T, m - 1);
        if ((int)tid < n) bs[tid] = (double)s_b[Vc.off + c_tid];
        __syncthreads();
        double s = 0.0;
        for (int k = 0; k < N / 16; ++k)
            if (row < n && part + 16 * k < n) s += invr[k] * bs[part + 16 * k];
        for (int w = 8; w > 0; w >>= 1) s += __shfl_xor(s, w, 16);
        if (row < n && part == 0) s_x[Vc.off + c_row] = (float)s;
        __syncthreads();
    }
    for (int l = m - 2; l >= 0; --l) {
        const TailLevel F = tail_level(T, l), C = tail_level(T, l + 1);
        tail_prolong(S, F, C);
        __syncthreads();
        for (int k = 0; k < T.sweeps; ++k) {
            tail_sweep(S, F, 1u);
            __syncthreads();
            tail_sweep(S, F, 0u);
            __syncthreads();
        }
    }
    for (int u = 0; u < kTailPer; ++u) {
        const uint32_t i = tid + 1024u * (uint32_t)u;
        if (i < V0.cells) T.L[0].x[i] = s_x[i];
    }
}

__global__ __launch_bounds__(1024) void k_coarse_tail(TailArgs T, const uint32_t *cidx, const uint32_t *cn,
                                                      const double *inv, const SolverState *st)
{
    coarse_tail_body(T, cidx, cn, inv, st);
}

// The coarse tail of every tank of a batch (workgroup t: tank t).
__global__ __launch_bounds__(1024) void kb_coarse_tail(const BatchTank *T)
{
    const BatchTank &B = T[blockIdx.x];
    coarse_tail_body(B.tail, B.coarse_idx, B.coarse_n, B.coarse_inv, B.state);
}

// ---- coarsest level: FP64 explicit inverse of the <= 64 fluid unknowns' operator (once per solve) ----
// The coarsest level is level 0 (mode 0, grids of <= 64 cells), a pooled level (mode 1) or a dense level (mode 2).
__global__ __launch_bounds__(256) void k_coarse_factor(Geo g, const uint32_t *ctype, const float *diag, int mode,
                                                       PLevel P, DenseLevel D, uint32_t *cidx, uint32_t *cn,
                                                       double *inv, uint32_t *sig)
{
    constexpr int N = (int)kCoarsestCells;
    __shared__ double a[N][2 * N + 1]; // [A | I], becoming [I | A^-1]
    __shared__ double col[N];
    __shared__ int px[N], py[N], pz[N];
    __shared__ uint32_t pidx[N];
    __shared__ uint32_t n_s;
    const int t = threadIdx.x;
    const int dims[3] = {mode == 0 ? g.n[0] : (mode == 1 ? P.dim[0] : D.dim[0]),
                         mode == 0 ? g.n[1] : (mode == 1 ? P.dim[1] : D.dim[1]),
                         mode == 0 ? g.n[2] : (mode == 1 ? P.dim[2] : D.dim[2])};
    const uint32_t cells = (uint32_t)(dims[0] * dims[1] * dims[2]);
    auto index_of = [&](int x, int y, int z) -> uint32_t {
        if (mode == 0) return cell_index(g, x, y, z);
        if (mode == 1) return pindex(g, P, x, y, z);
        return (uint32_t)(x + dims[0] * (y + dims[1] * z));
    };
    auto type_of = [&](int x, int y, int z) -> uint32_t {
        if (x < 0 || y < 0 || z < 0 || x >= dims[0] || y >= dims[1] || z >= dims[2]) return kCellSolid;
        uint32_t i = index_of(x, y, z);
        if (mode == 0) return ctype[i] & kCellTypeMask;
        return mode == 1 ? P.type[i] : D.type[i];
    };
    // the fluid cells in increasing index order (the level has at most N = 64 cells, which slosh_create checks: one
    // wavefront, compacted by a ballot)
    static_assert(N == 64, "one wavefront scans the coarsest level");
    // A coarse level's operator, fluid list and inverse follow from its cells' types and pool indices alone: when
    // they equal those of the last factorisation (sig[0..63], sig[64] = 1 once one was made), the inverse stands.
    // A level-0 coarsest level (grids of at most 64 cells) also reads the diagonal and is always refactored.
    __shared__ uint32_t s_same;
    if (mode != 0 && t < N) {
        uint32_t v = 0xFFFFFFFFu;
        if ((uint32_t)t < cells) {
            const int x = t % dims[0], y = (t / dims[0]) % dims[1], z = t / (dims[0] * dims[1]);
            v = type_of(x, y, z) | (index_of(x, y, z) << 2);
        }
        const bool differ = v != sig[t];
        const uint64_t dm = __ballot(differ);
        if (differ) sig[t] = v;
        if (t == 0) {
            s_same = (dm == 0ull && sig[N] == 1u) ? 1u : 0u;
            sig[N] = 1u;
        }
    }
    if (mode == 0 && t == 0) s_same = 0u;
    __syncthreads();
    if (s_same) return;
    if (t < N) {
        bool f = false;
        int x = 0, y = 0, z = 0;
        if ((uint32_t)t < cells) {
            x = t % dims[0];
            y = (t / dims[0]) % dims[1];
            z = t / (dims[0] * dims[1]);
            f = type_of(x, y, z) == kCellFluid;
        }
        const uint64_t m = __ballot(f);
        const uint32_t pos = (uint32_t)__popcll(m & ((1ull << t) - 1ull));
        if (f) {
            px[pos] = x;
            py[pos] = y;
            pz[pos] = z;
            pidx[pos] = index_of(x, y, z);
        }
        if (t == 0) {
            n_s = (uint32_t)__popcll(m);
            *cn = n_s;
        }
    }
    __syncthreads();
    const int n = (int)n_s;
    for (int e = t; e < N * 2 * N; e += blockDim.x) a[e / (2 * N)][e % (2 * N)] = (e % (2 * N)) == N + e / (2 * N) ? 1.0 : 0.0;
    __syncthreads();
    if (t < n) {
        cidx[t] = pidx[t];
        float dg = kDiagEps;
        for (int nb = 0; nb < 6; ++nb) {
            int d[3];
            nb_offset(nb, d);
            int nx = px[t] + d[0], ny = py[t] + d[1], nz = pz[t] + d[2];
            if (nx < 0 || ny < 0 || nz < 0 || nx >= dims[0] || ny >= dims[1] || nz >= dims[2]) continue;
            uint32_t ty = type_of(nx, ny, nz);
            if (ty != kCellSolid) dg += 1.0f;
            if (ty == kCellFluid)
                for (int j = 0; j < n; ++j)
                    if (px[j] == nx && py[j] == ny && pz[j] == nz) a[t][j] = -1.0;
        }
        if (mode == 0) dg = diag[pidx[t]] + kDiagEps;
        a[t][t] = (double)dg;
    }
    __syncthreads();
    // Gauss-Jordan elimination on [A | I] (A is symmetric positive definite: no pivoting), all threads per step
    for (int k = 0; k < n; ++k) {
        const double piv = a[k][k];
        __syncthreads();
        for (int j = t; j < 2 * N; j += blockDim.x) a[k][j] /= piv;
        __syncthreads();
        if (t < n) col[t] = a[t][k];
        __syncthreads();
        for (int e = t; e < n * 2 * N; e += blockDim.x) {
            const int i = e / (2 * N), j = e % (2 * N);
            if (i != k) a[i][j] -= col[i] * a[k][j];
        }
        __syncthreads();
    }
    for (int e = t; e < N * N; e += blockDim.x) {
        const int i = e / N, j = e % N;
        inv[(size_t)i * N + j] = (i < n && j < n) ? a[i][N + j] : 0.0;
    }
}

// Exact coarsest solve x = A^-1 b in FP64 (replaces MGCoarseSolveCS, SloshPressureMG.usf:386-413) when the
// coarsest level is level 0 or pooled; b and x are indexed by the factor's cell list.
__global__ __launch_bounds__(64) void k_coarse_solve(const uint32_t *cidx, const uint32_t *cn, const double *inv,
                                                     LVec b, LVec x, const SolverState *st)
{
    if (st->done) return;
    constexpr int N = (int)kCoarsestCells;
    __shared__ double bs[N];
    const int n = (int)*cn;
    const int t = threadIdx.x;
    if (t < n) bs[t] = (double)b[cidx[t]];
    __syncthreads();
    if (t >= n) return;
    double s = 0.0;
    for (int j = 0; j < n; ++j) s += inv[(size_t)t * N + j] * bs[j];
    x[cidx[t]] = (float)s;
}

__global__ void k_state_reset(SolverState *st, double tol, int maxit)
{
    st->tol = tol;
    st->maxit = maxit;
    st->done = 0;
    st->converged = 0;
    st->iters = 0;
    st->rho = 0.0;
    st->alpha = 0.0;
    st->beta = 0.0;
    st->rel = 0.0;
    st->rel_true = 0.0;
    st->err_rel = -1.0;
    st->stalled = 0;
    for (int j = 0; j < 4; ++j) st->hist[j] = 0.0;
}

// Reduction of the partials of the preceding level-0 pass: one per brick with fluid, except the initial residual,
// which covers every allocated brick.
void finalize(slosh_ctx *c, int kind, int iters = 0)
{
    SLOSH_TIMED(c, K_FINALIZE);
    uint32_t nblocks = c->nparts;
    const double *parts = c->partials;
    if (nblocks > kFinOneMax) {
        const uint32_t chunk = (nblocks + kFoldGroups - 1u) / kFoldGroups, groups = (nblocks + chunk - 1u) / chunk;
        hipLaunchKernelGGL(k_fold_partials, dim3(groups), dim3(256), 0, c->stream, kind, c->partials, nblocks, chunk,
                           c->fold, c->state);
        parts = c->fold;
        nblocks = groups;
    }
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFinThreads), 0, c->stream, kind, parts, nblocks, iters, c->state);
}

void coarse_solve(slosh_ctx *c, LVec b, LVec x)
{
    launch_timed(c, K_COARSE_SOLVE, k_coarse_solve, dim3(1), dim3(64), 0, c->coarse_idx, c->coarse_n, c->coarse_inv, b,
                 x, c->state);
}

// The coarse tail's view of dense levels t .. m-1.
TailArgs tail_args(const slosh_ctx *c, int t)
{
    const int m = c->nlevels;
    TailArgs T{};
    T.n = m - t;
    T.sweeps = coarse_sweeps(c);
    T.info = c->dense_info + c->lv[t].dense_off;
    T.coord = c->dense_coord + c->lv[t].dense_off;
    for (int k = 0; k < T.n; ++k) {
        T.L[k] = dlevel_of(c, t + k);
        T.off[k] = c->lv[t + k].dense_off - c->lv[t].dense_off; // levels are contiguous in the blocks
        T.total = T.off[k] + T.L[k].cells;
    }
    return T;
}

// The first dense level the coarse tail takes (the coarser ones that fit its LDS together), for m >= 5 levels.
int tail_start(const slosh_ctx *c)
{
    const int m = c->nlevels;
    int t = m - 1;
    uint32_t tail_cells = c->lv[m - 1].cells;
    while (t - 1 >= kFirstDense && tail_cells + c->lv[t - 1].cells <= kTailCells) tail_cells += c->lv[--t].cells;
    return t;
}

// Levels 2 .. m-1 (m >= 5): the fused level-2/3 down pass with the restriction into level 4, grid-wide passes for
// dense levels kFirstDense .. t-1 (only when the dense levels do not all fit the tail's LDS), the tail on levels
// t .. m-1, the grid-wide up passes, the fused level-3/2 up pass.
void dense_part(slosh_ctx *c)
{
    const DenseLevel L3 = dlevel_of(c, 3), L4 = dlevel_of(c, kFirstDense);
    const PLevel L2 = plevel_of(c, 2);
    {
        launch_timed(c, K_D3DOWN, k_d23down, dim3(c->ntiles_f), dim3(kD23DownThreads), 0, c->geo, c->tile_list, L2,
                     L3, L4, c->state);
    }
    const int t = tail_start(c);
    const int sweeps = coarse_sweeps(c);
    for (int l = kFirstDense; l < t; ++l) {
        DenseLevel F = dlevel_of(c, l), C = dlevel_of(c, l + 1);
        for (int k = 0; k < sweeps; ++k)
            for (uint32_t colour : {0u, 1u}) {
                if (k == 0 && colour == 0u) continue; // the first red sweep is fused into the restriction
                launch_timed(c, K_DSMOOTH, k_dsmooth, dim3(blocks_for(F.cells, 256)), dim3(256), 0, F, colour,
                             c->state);
            }
        launch_timed(c, K_DRESTRICT, k_drestrict, dim3(blocks_for(C.cells, 256)), dim3(256), 0, F, C, c->state);
    }
    const TailArgs T = tail_args(c, t);
    {
        launch_timed(c, K_COARSE_TAIL, k_coarse_tail, dim3(1), dim3(1024), 0, T, c->coarse_idx, c->coarse_n,
                     c->coarse_inv, c->state);
    }
    for (int l = t - 1; l >= kFirstDense; --l) {
        DenseLevel F = dlevel_of(c, l), C = dlevel_of(c, l + 1);
        {
            launch_timed(c, K_DPROLONG, k_dprolong, dim3(blocks_for(F.cells, 256)), dim3(256), 0, F, C, c->state);
        }
        for (int k = 0; k < sweeps; ++k)
            for (uint32_t colour : {1u, 0u}) {
                launch_timed(c, K_DSMOOTH, k_dsmooth, dim3(blocks_for(F.cells, 256)), dim3(256), 0, F, colour,
                             c->state);
            }
    }
    launch_timed(c, K_D3UP, k_d23up, dim3(c->ntiles_f), dim3(kD23UpThreads), 0, c->geo, c->tile_list, L2, L3, L4,
                 c->state);
}

// SloshProxy.cpp:904-959 (VCycle) with the exact coarsest solve, from the level-0 down-sweep iterate (red
// initialisation and black sweep already in z) to the red sweep of the result z1 and the r.z1 reduction.
void vcycle(slosh_ctx *c, const float *r)
{
    const int m = c->nlevels;
    if (m == 1) {
        coarse_solve(c, LVec{(void *)r, 0}, LVec{c->z1, 0});
    } else {
        l0_restrict(c, r);
        // pooled levels with a down pass of their own: 1 and 2 when level 3 is the coarsest, else 1 (level 2 is
        // fused with level 3 in the dense part)
        const int last_pooled = m - 2 < 2 ? m - 2 : (m - 1 <= 3 ? 2 : 1);
        for (int l = 1; l <= last_pooled; ++l) pooled_down(c, l);
        if (m - 1 <= 3) coarse_solve(c, LVec{c->lv[m - 1].b, c->lv[m - 1].half}, LVec{c->lv[m - 1].x1, c->lv[m - 1].half});
        else dense_part(c);
        for (int l = last_pooled; l >= 1; --l) pooled_up(c, l);
        l0_prolong_black(c, r);
    }
    l0_red_dot(c, r);
    finalize(c, kFinRho);
}

// Per-solve set-up: brick table of the allocated bricks, coarse types and stencil words, coarsest factor.
void mg_setup(slosh_ctx *c)
{
    l0_brick_info(c, false);
    const int m = c->nlevels;
    for (int l = 1; l < m; ++l) {
        if (c->lv[l].pooled) {
            pooled_build_type(c, l);
        } else if (c->lv[l - 1].pooled) {
            dense_build_type_from_pooled(c, l);
        } else {
            DenseLevel C = dlevel_of(c, l);
            launch_timed(c, K_MG_SETUP, k_build_type, dim3(blocks_for(C.cells, 256)), dim3(256), 0, dlevel_of(c, l - 1),
                         C, fluid_coarsening(c));
        }
    }
    for (int l = 1; l < m; ++l) {
        if (c->lv[l].pooled) {
            pooled_build_info(c, l);
        } else {
            DenseLevel D = dlevel_of(c, l);
            launch_timed(c, K_MG_SETUP, k_dinfo, dim3(blocks_for(D.cells, 256)), dim3(256), 0, D);
        }
    }
    // tiles of level 3 for the fused level-2/3 passes (levels 3 and 4 exist below a coarser level)
    c->ntiles_f = 0;
    if (m >= 5) {
        const DenseLevel L3 = dlevel_of(c, 3);
        SLOSH_TIMED(c, K_MG_SETUP);
        launch_fill_u32(c->stream, c->tile_flag, 0u, c->ntiles);
        hipLaunchKernelGGL(k_tile_flags, dim3((c->nactive + 255u) / 256u), dim3(256), 0, c->stream, c->geo, c->nactive,
                           c->fluid_flag, L3, c->tile_flag);
        scan_exclusive(c, c->tile_flag, c->tile_rank, c->ntiles);
        hipLaunchKernelGGL(k_tile_list, dim3((c->ntiles + 255u) / 256u), dim3(256), 0, c->stream, c->ntiles,
                           c->tile_flag, c->tile_rank, c->tile_list);
        SLOSH_HIP(hipMemcpyAsync(c->h_nsolve + 1, c->tile_rank + c->ntiles, sizeof(uint32_t), hipMemcpyDeviceToHost,
                                 c->stream));
    }
    // one readback for the counts of bricks with fluid and of tiles; the host waits for it (mg_counts) while the
    // coarsest factor and the residual initialisation run
    SLOSH_HIP(hipEventRecord(c->ev_counts, c->stream));
    const int mode = m == 1 ? 0 : (c->lv[m - 1].pooled ? 1 : 2);
    PLevel P = mode == 1 ? plevel_of(c, m - 1) : PLevel{};
    DenseLevel D = mode == 2 ? dlevel_of(c, m - 1) : DenseLevel{};
    launch_timed(c, K_COARSE_FACTOR, k_coarse_factor, dim3(1), dim3(256), 0, c->geo, c->ctype, c->diag, mode, P, D,
                 c->coarse_idx, c->coarse_n, c->coarse_inv, c->coarse_sig);
}

// The counts read back by mg_setup; every pass launched over the bricks with fluid or the level-3 tiles comes after.
void mg_counts(slosh_ctx *c)
{
    SLOSH_HIP(hipEventSynchronize(c->ev_counts));
    c->nsolve = c->h_nsolve[0];
    c->ntiles_f = c->nlevels >= 5 ? c->h_nsolve[1] : 0u;
}

// Captures the launches of `body` on the context's stream into c->mg_graph (updating the executable graph in
// place when the topology is unchanged, else instantiating it anew). On any capture failure the launches are
// enqueued directly instead, graphs are switched off for the context, and false is returned.
template <typename F> bool capture_chunk_graph(slosh_ctx *c, hipGraphExec_t &exec, F body)
{
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        c->use_graphs = false;
        body();
        return false;
    }
    body();
    if (hipStreamEndCapture(c->stream, &g) != hipSuccess || !g) {
        c->use_graphs = false;
        (void)hipGetLastError();
        body();
        return false;
    }
    bool ok = false;
    if (exec) {
        hipGraphNode_t err_node = nullptr;
        hipGraphExecUpdateResult res;
        ok = hipGraphExecUpdate(exec, g, &err_node, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
        if (!ok) {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(exec);
            exec = nullptr;
        }
    }
    if (!ok) ok = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) == hipSuccess;
    (void)hipGraphDestroy(g);
    if (!ok) {
        (void)hipGetLastError();
        exec = nullptr;
        c->use_graphs = false;
        body();
    }
    return ok;
}

} // namespace

// The MGPCG iterations of a solve whose residual r = b - A p is in c->r and whose state holds the initial
// reduction: chunks of iterations between reads of the state, then the true-residual check, restarting from
// r = b - A p up to max_restarts times. Residuals are captured for training when capture_on and the settings ask.
void mgpcg_iterate(slosh_ctx *c, int maxit, int max_restarts, bool capture_on, bool resumed = false)
{
    float *R[2] = {c->r, c->r2};
    float *D[2] = {c->d, c->d2};
    float *const cap_buf = capture_on ? c->capture : nullptr;
    // SLOSH_FLAG_CAPTURE_VCYCLE_RESIDUALS: capture r - A V(r), the residual the V-cycle leaves, after each V-cycle
    // instead of the residual r before it
    const bool after_vcycle = (c->s.flags & SLOSH_FLAG_CAPTURE_VCYCLE_RESIDUALS) != 0;
    // residuals kept for training the learned preconditioner: the initial one and those after iterations 1, 2,
    // 3, 4, 6, 8, 11, 15, 20, ... (slots past the solve's last iteration are dropped afterwards)
    uint32_t ncap = 0;
    int next_cap = 0, cap_gap = 1;
    auto capture = [&](const float *r, int iteration, bool vcycled) {
        if (!cap_buf || vcycled != after_vcycle || ncap >= c->capture_cap || iteration != next_cap) return;
        float *slot = cap_buf + (size_t)ncap * c->geo.pool_cells;
        if (vcycled) l0_fcg_resid(c, r, c->z1, slot);
        else SLOSH_HIP(hipMemcpyAsync(slot, r, (size_t)c->geo.pool_cells * 4u, hipMemcpyDeviceToDevice, c->stream));
        c->capture_iter[ncap++] = iteration;
        next_cap += cap_gap;
        if (iteration >= 4) cap_gap++;
    };
    if (!resumed) capture(R[0], 0, false);
    int rc = 1, dc = 0;
    // the first iteration's update (no alpha yet), V-cycle and direction, from the residual in R[0]
    auto first_iteration = [&]() {
        l0_update_black(c, R[0], R[1], D[0], 1);
        rc = 1;
        dc = 0;
        vcycle(c, R[rc]);
        capture(R[rc], 0, true);
        l0_direction(c, D[1], D[0], R[rc], 1);
    };
    if (!resumed) first_iteration();
    SLOSH_HIP(hipGetLastError());
    // Iterations are enqueued in chunks between reads of the device state (an even count, so the ping-pong
    // buffers return to the same roles). Without residual capture or kernel timing, a chunk is captured once per
    // solve as a HIP graph and replayed from the second chunk on: every chunk of a solve launches the same kernels
    // on the same buffers.
    const int chunk = c->solver_chunk >= 2 ? c->solver_chunk & ~1 : 4;
    const bool use_graph = c->use_graphs && !cap_buf && c->timing_mask == 0;
    int done_iters = 0; // iterations of the chunks enqueued so far
    auto enqueue_chunk = [&](int it0) {
        for (int k = 0; k < chunk; ++k) {
            // alpha and the stopping test of the update it makes (done == 2: only that update still runs)
            finalize(c, kFinAlphaStop);
            l0_update_black(c, R[rc], R[rc ^ 1], D[dc], 0);
            rc ^= 1;
            capture(R[rc], it0 + k + 1, false);
            vcycle(c, R[rc]);
            capture(R[rc], it0 + k + 1, true);
            l0_direction(c, D[dc], D[dc ^ 1], R[rc], 0);
            dc ^= 1;
        }
    };
    bool graph_ready = false;
    // Before the first read of the state, as many chunks as the previous solve's iteration count fills (less one
    // iteration): consecutive solves of a simulation take similar counts, and a chunk enqueued past convergence
    // only costs its early-exit launches. Afterwards the state is read after every chunk.
    int ahead = c->last_mg_iters > 1 ? (c->last_mg_iters + chunk - 1) / chunk : 1;
    if (ahead < 1) ahead = 1;
    // The recurrence residual that stops the iteration can differ from the FP64 true residual ||b - A p|| / ||b||
    // of the pressure in the last bits; the solve is accepted only when the true residual meets the tolerance too,
    // else it restarts from r = b - A p of the current pressure (a few times at most, within max_iterations).
    // resumed: the iterations ran in a lockstep batch, whose last read of the state is in h_state
    if (resumed) done_iters = maxit + 1;
    for (int restart = 0;; ++restart) {
        while (done_iters <= maxit) {
            for (int k = 0; k < ahead && done_iters <= maxit; ++k, done_iters += chunk) {
                // the solve's first chunk is enqueued directly, so that the GPU runs it while the host captures the
                // graph and updates the executable one (which takes longer than the first iteration)
                if (use_graph && c->use_graphs && done_iters > 0) {
                    // on a capture failure the helper has enqueued the chunk directly and switched graphs off
                    if (!graph_ready) graph_ready = capture_chunk_graph(c, c->mg_graph, [&] { enqueue_chunk(done_iters); });
                    if (graph_ready) SLOSH_HIP(hipGraphLaunch(c->mg_graph, c->stream));
                } else {
                    enqueue_chunk(done_iters);
                }
                SLOSH_HIP(hipGetLastError());
            }
            ahead = 1;
            // the true residual of the FP64 pressure decides convergence and is reported with the results; its
            // check rides along with every read of the state (both passes return at once while the solve
            // iterates), so a finished solve costs no second host round trip
            l0_true_residual_if_done(c);
            finalize(c, kFinCheckDone);
            SLOSH_HIP(hipMemcpyAsync(c->h_state, c->state, sizeof(SolverState), hipMemcpyDeviceToHost, c->stream));
            SLOSH_HIP(hipStreamSynchronize(c->stream));
            if (c->h_state->done) break;
        }
        if (!c->h_state->done) {
            // out of enqueued iterations without the device's stop (not expected: the limit stops it first)
            l0_true_residual(c);
            finalize(c, kFinCheck);
            SLOSH_HIP(hipMemcpyAsync(c->h_state, c->state, sizeof(SolverState), hipMemcpyDeviceToHost, c->stream));
            SLOSH_HIP(hipStreamSynchronize(c->stream));
        }
        if (c->h_state->converged || c->h_state->iters >= maxit || restart >= max_restarts) break;
        l0_pcg_init(c, R[0]);
        finalize(c, kFinRestart);
        first_iteration();
        done_iters = c->h_state->iters;
        ahead = 1;
    }
    if (cap_buf) {
        // the last update's residual gets no V-cycle (the solve stops): its capture after the V-cycle is dropped
        uint32_t valid = 0;
        while (valid < ncap && (after_vcycle ? c->capture_iter[valid] < c->h_state->iters
                                             : c->capture_iter[valid] <= c->h_state->iters))
            valid++;
        c->capture_n = valid;
    }
}

// SLOSH_FLAG_REPORT_FP64_ERROR: the error of the solution against an FP64-accurate solve of the same system. The
// solution is kept, the system solved again from it to a true residual of 1e-12 (each restart forms r = b - A p
// in FP64, so neither the FP32 vectors nor the FP16 levels limit the accuracy), the relative L2 difference reduced
// over the fluid cells, and the first solution and its solver state restored.
void reference_error(slosh_ctx *c, int maxit)
{
    constexpr double kReferenceTol = 1e-12;
    const size_t pc = c->geo.pool_cells;
    SLOSH_HIP(hipMemcpyAsync(c->pres_lp, c->pres, pc * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    SLOSH_HIP(hipMemcpyAsync(c->state_save, c->state, sizeof(SolverState), hipMemcpyDeviceToDevice, c->stream));
    const SolverState kept = *c->h_state;
    const int last = c->last_mg_iters;
    hipLaunchKernelGGL(k_state_reset, dim3(1), dim3(1), 0, c->stream, c->state, kReferenceTol, 4 * maxit);
    l0_pcg_init(c, c->r);
    finalize(c, kFinInit);
    mgpcg_iterate(c, 4 * maxit, 12, false);
    l0_error_partials(c, c->pres_lp);
    SLOSH_HIP(hipMemcpyAsync(c->pres, c->pres_lp, pc * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    SLOSH_HIP(hipMemcpyAsync(c->state, c->state_save, sizeof(SolverState), hipMemcpyDeviceToDevice, c->stream));
    finalize(c, kFinErr);
    *c->h_state = kept;
    c->last_mg_iters = last;
}

// The set-up of an MGPCG solve (state reset, multigrid set-up, r = b - A p, its reduction, the counts of bricks and
// tiles with fluid). Returns whether the solve needs iterations: false without fluid (b = 0, converged as it
// stands).
bool mgpcg_begin(slosh_ctx *c)
{
    if (!c->nactive) return false;
    const int maxit = c->s.max_iterations > 0 ? c->s.max_iterations : 200;
    hipLaunchKernelGGL(k_state_reset, dim3(1), dim3(1), 0, c->stream, c->state, (double)c->s.tolerance, maxit);
    mg_setup(c);
    l0_pcg_init(c, c->r);
    finalize(c, kFinInit);
    mg_counts(c);
    if (!c->nsolve) {
        c->capture_n = 0;
        return false;
    }
    return true;
}

// SloshProxy.cpp:807-1020 (AddPressureSolveMGPCG), to ||r|| <= tol ||b||.
void run_mgpcg(slosh_ctx *c)
{
    if (!mgpcg_begin(c)) return;
    const int maxit = c->s.max_iterations > 0 ? c->s.max_iterations : 200;
    mgpcg_iterate(c, maxit, 3, true);
    c->last_mg_iters = c->h_state->iters;
    if (c->s.flags & SLOSH_FLAG_REPORT_FP64_ERROR) reference_error(c, maxit);
}

// Flexible CG (Notay 2000, FCG(1)) with the hybrid learned preconditioner M(r) = V(r) + N(r - A V(r)): the MGPCG
// V-cycle, and the network's correction of the residual the V-cycle leaves. z = M(r), d = z - (z.q_prev /
// d_prev.q_prev) d_prev, alpha = d.r / d.q. Iterations are enqueued in chunks between reads of the device state
// (every kernel, the network's included, returns at once once the solve is done), the chunk replayed as a captured
// HIP graph. MGPCG is the fallback and the judge: when the residual stops halving over four iterations, is not
// finite or the iteration limit is reached, MGPCG solves from the current pressure; a solve whose recurrence
// converged restarts from r = b - A p until its FP64 true residual meets the tolerance too (a few times at most,
// then MGPCG). Every accepted solve therefore has a true residual within the tolerance.
void run_fcg(slosh_ctx *c)
{
    c->fcg_iters = 0;
    c->fcg_fallback = 0;
    if (!c->nactive) return;
    if (!c->unet.loaded) {
        c->fcg_fallback = 2; // no network weights: the solve is plain MGPCG
        run_mgpcg(c);
        return;
    }
    const int maxit = c->s.max_iterations > 0 ? c->s.max_iterations : 200;
    hipLaunchKernelGGL(k_state_reset, dim3(1), dim3(1), 0, c->stream, c->state, (double)c->s.tolerance, maxit);
    mg_setup(c);
    l0_pcg_init(c, c->r);
    finalize(c, kFinInit);
    mg_counts(c);
    if (!c->nsolve) return;
    UNet &u = c->unet;
    struct SkipGuard {
        UNet &u;
        ~SkipGuard() { u.skip = u.zero_flag; }
    } guard{u};
    u.skip = (const int32_t *)((const char *)c->state + offsetof(SolverState, done));
    float *D[2] = {c->d, c->d2};
    int dc = 0;
    auto iteration = [&](bool first) {
        l0_update_black(c, c->r, c->r2, D[dc], 1); // r2 = r; z: red r / diag, then the black sweep
        vcycle(c, c->r2);                          // z1 = V(r)
        l0_fcg_resid(c, c->r2, c->z1, c->fcg_rp);  // r' = r - A z1
        unet_apply(c, c->fcg_rp, c->fcg_z, c->z1); // z = z1 + N(r')
        if (!first) {
            l0_fcg_zq(c, c->fcg_z);
            finalize(c, kFinBetaF);
        }
        l0_fcg_direction(c, c->fcg_z, D[dc], D[dc ^ 1], c->r, first);
        dc ^= 1;
        finalize(c, kFinAlphaF);
        l0_fcg_update(c, c->r, D[dc]);
        finalize(c, kFinRr);
    };
    const int chunk = 2; // even: the direction buffers return to the same roles
    const bool use_graph = c->use_graphs && c->timing_mask == 0;
    bool graph_ready = false;
    auto read_state = [&]() {
        SLOSH_HIP(hipMemcpyAsync(c->h_state, c->state, sizeof(SolverState), hipMemcpyDeviceToHost, c->stream));
        SLOSH_HIP(hipStreamSynchronize(c->stream));
    };
    bool fallback = false;
    for (int restart = 0;; ++restart) {
        dc = 0;
        iteration(true);
        int enqueued = 1;
        int ahead = c->last_fcg_iters > 1 ? (c->last_fcg_iters + chunk - 2) / chunk : 1;
        for (;;) {
            for (int k = 0; k < ahead && enqueued < maxit; ++k, enqueued += chunk) {
                if (use_graph && c->use_graphs) {
                    if (!graph_ready)
                        graph_ready = capture_chunk_graph(c, c->fcg_graph, [&] {
                            for (int j = 0; j < chunk; ++j) iteration(false);
                        });
                    if (graph_ready) SLOSH_HIP(hipGraphLaunch(c->fcg_graph, c->stream));
                } else {
                    for (int j = 0; j < chunk; ++j) iteration(false);
                }
                SLOSH_HIP(hipGetLastError());
            }
            ahead = 1;
            read_state();
            if (c->h_state->done || enqueued >= maxit) break;
        }
        c->fcg_iters = c->h_state->iters;
        if (c->h_state->stalled || !c->h_state->converged) {
            fallback = true;
            break;
        }
        l0_true_residual(c);
        finalize(c, kFinCheck);
        read_state();
        if (c->h_state->converged) break;
        if (restart >= 3 || c->h_state->iters >= maxit) {
            fallback = true;
            break;
        }
        l0_pcg_init(c, c->r);
        finalize(c, kFinRestart);
    }
    c->last_fcg_iters = c->fcg_iters;
    if (fallback) {
        if (!std::isfinite(c->h_state->rel)) launch_fill_f64(c->stream, c->pres, 0.0, c->geo.pool_cells);
        c->fcg_fallback = 1;
        run_mgpcg(c);
    } else if (c->s.flags & SLOSH_FLAG_REPORT_FP64_ERROR) {
        reference_error(c, maxit);
    }
}

// SloshProxy.cpp:656-683: the density projection's solve, run by the pressure solver on the same operator with
// Divergence := the projection's right-hand side and Pressure := q (from 0): two MGPCG iterations, or
// density_projection_iterations RB-SOR sweeps at omega 1.7. MGPCG (and the flexible CG, whose V-cycle hierarchy is
// the same) reuse the multigrid set-up of the step's pressure solve. No stopping test (tolerance 0), as in the
// reference.
void run_density_solve(slosh_ctx *c)
{
    if (!c->nactive || !c->nsolve) return;
    struct Swap {
        slosh_ctx *c;
        float *div;
        double *pres;
        int last;
        ~Swap()
        {
            c->div = div;
            c->pres = pres;
            c->last_mg_iters = last;
        }
    } keep{c, c->div, c->pres, c->last_mg_iters};
    c->div = c->dens_rhs;
    c->pres = c->dens_q;
    if (c->s.pressure_solver == SLOSH_SOLVER_RBSOR) {
        const int iters = c->s.density_projection_iterations > 0 ? c->s.density_projection_iterations : 1;
        hipLaunchKernelGGL(k_state_reset, dim3(1), dim3(1), 0, c->stream, c->state, 0.0, iters);
        sor_iterate(c, iters, (double)1.7f);
        SLOSH_HIP(hipGetLastError());
        return;
    }
    constexpr int kIters = 2;
    hipLaunchKernelGGL(k_state_reset, dim3(1), dim3(1), 0, c->stream, c->state, 0.0, kIters);
    l0_pcg_init(c, c->r);
    finalize(c, kFinInit);
    c->last_mg_iters = kIters;
    mgpcg_iterate(c, kIters, 0, false);
}

// Slosh.usf:856-892 (PressureRBGSCS) iterated to the tolerance, residual checked every 8 iterations.
// `iters` red-black SOR iterations on c->pres, each one two-colour pass from one of c->pres and c->sor_tmp into the
// other, the last one into c->pres (an odd count starts from a copy in c->sor_tmp). A pass writes every fluid cell
// and reads the other cells' values only through masks that exclude them, so what c->sor_tmp holds there is never
// used. The passes of a solve that is already done return at once, leaving c->pres as it is.
void sor_iterate(slosh_ctx *c, int iters, double w)
{
    if (iters <= 0) return;
    double *a = c->pres, *b = c->sor_tmp;
    if (iters & 1) {
        SLOSH_HIP(hipMemcpyAsync(b, a, (size_t)c->geo.pool_cells * sizeof(double), hipMemcpyDeviceToDevice,
                                 c->stream));
        std::swap(a, b);
    }
    for (int k = 0; k < iters; ++k) {
        l0_sor_pair(c, a, b, w);
        std::swap(a, b);
    }
}

void run_rbsor(slosh_ctx *c)
{
    if (!c->nactive) return;
    const int maxit = c->s.max_iterations > 0 ? c->s.max_iterations : 1000;
    hipLaunchKernelGGL(k_state_reset, dim3(1), dim3(1), 0, c->stream, c->state, (double)c->s.tolerance, maxit);
    l0_brick_info(c);
    if (!c->nsolve) {
        // no fluid cell: b = 0, the solve is converged as it stands
        l0_pcg_init(c, c->r);
        finalize(c, kFinInit);
        return;
    }
    const double w = (double)c->s.sor_weight;
    int it = 0;
    auto residual = [&](int iters) {
        l0_sor_residual(c);
        finalize(c, kFinSor, iters);
    };
    residual(0);
    while (it < maxit) {
        for (int chunk = 0; chunk < 16 && it < maxit; ++chunk) {
            const int k = maxit - it < 8 ? maxit - it : 8;
            sor_iterate(c, k, w);
            it += k;
            residual(it);
        }
        SLOSH_HIP(hipGetLastError());
        SLOSH_HIP(hipMemcpyAsync(c->h_state, c->state, sizeof(SolverState), hipMemcpyDeviceToHost, c->stream));
        SLOSH_HIP(hipStreamSynchronize(c->stream));
        if (c->h_state->done) break;
    }
}

// ---- batched MGPCG: the iterations of many tanks in lockstep, each pass one launch over all of them ----------
// A tank's solve joins when it is MGPCG without residual capture, FP64 error report or kernel timing, on a hierarchy
// of at least five levels whose dense levels all fit the coarse tail (the V-cycle is then the same pass sequence for
// every tank: level 0, level 1, the fused level-2/3 passes, the tail).
bool mgpcg_batchable(const slosh_ctx *c)
{
    if (c->s.pressure_solver != SLOSH_SOLVER_MGPCG || c->timing_mask) return false;
    if (c->s.flags & (SLOSH_FLAG_CAPTURE_RESIDUALS | SLOSH_FLAG_REPORT_FP64_ERROR)) return false;
    // the batch's finalize reads a tank's partials in one workgroup (at most one set per allocated brick)
    if (c->nactive > kFinOneMax) return false;
    return c->nlevels >= 5 && tail_start(c) == kFirstDense;
}

void batch_tank_of(const slosh_ctx *c, BatchTank &T)
{
    T = BatchTank{};
    T.binfo_f = c->binfo_f;
    T.fmask = c->fmask;
    T.diag = c->diag;
    T.div = c->div;
    T.pres = c->pres;
    T.r[0] = c->r;
    T.r[1] = c->r2;
    T.z = c->z;
    T.z1 = c->z1;
    T.d[0] = c->d;
    T.d[1] = c->d2;
    T.q = c->q;
    T.partials = c->partials;
    T.state = c->state;
    T.nsolve = c->nsolve;
    T.ntiles_f = c->ntiles_f;
    T.geo = c->geo;
    T.tile_list = c->tile_list;
    T.L1 = plevel_of(c, 1);
    T.L2 = plevel_of(c, 2);
    T.L3 = dlevel_of(c, 3);
    T.L4 = dlevel_of(c, kFirstDense);
    T.tail = tail_args(c, kFirstDense);
    T.coarse_idx = c->coarse_idx;
    T.coarse_n = c->coarse_n;
    T.coarse_inv = c->coarse_inv;
}

void b_d23down(const BatchRun &B)
{
    hipLaunchKernelGGL(kb_d23down, dim3(B.total[kRangeTile]), dim3(kD23DownThreads), 0u, B.stream, B.d_tanks,
                       B.range(kRangeTile), B.ntanks);
}

void b_d23up(const BatchRun &B)
{
    hipLaunchKernelGGL(kb_d23up, dim3(B.total[kRangeTile]), dim3(kD23UpThreads), 0u, B.stream, B.d_tanks,
                       B.range(kRangeTile), B.ntanks);
}

void b_coarse_tail(const BatchRun &B)
{
    hipLaunchKernelGGL(kb_coarse_tail, dim3(B.ntanks), dim3(1024), 0u, B.stream, B.d_tanks);
}

void b_finalize(const BatchRun &B, int kind)
{
    hipLaunchKernelGGL(kb_finalize, dim3(B.ntanks), dim3(kFinThreads), 0u, B.stream, kind, B.d_tanks);
}

namespace {
// Copies every tank's solver state into one array (one read back for the batch).
__global__ void kb_gather_states(const BatchTank *T, SolverState *out)
{
    static_assert(sizeof(SolverState) % 4u == 0u, "word copy");
    const uint32_t *src = (const uint32_t *)T[blockIdx.x].state;
    uint32_t *dst = (uint32_t *)(out + blockIdx.x);
    for (uint32_t i = threadIdx.x; i < sizeof(SolverState) / 4u; i += blockDim.x) dst[i] = src[i];
}

// mgpcg_iterate's graph capture for the batch's stream.
template <typename F> bool capture_batch_graph(const BatchRun &B, hipGraphExec_t &exec, bool &graphs, F body)
{
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        graphs = false;
        body();
        return false;
    }
    body();
    if (hipStreamEndCapture(B.stream, &g) != hipSuccess || !g) {
        graphs = false;
        (void)hipGetLastError();
        body();
        return false;
    }
    bool ok = false;
    if (exec) {
        hipGraphNode_t err_node = nullptr;
        hipGraphExecUpdateResult res;
        ok = hipGraphExecUpdate(exec, g, &err_node, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
        if (!ok) {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(exec);
            exec = nullptr;
        }
    }
    if (!ok) ok = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) == hipSuccess;
    (void)hipGraphDestroy(g);
    if (!ok) {
        (void)hipGetLastError();
        exec = nullptr;
        graphs = false;
        body();
    }
    return ok;
}
} // namespace

// mgpcg_iterate's iterations for every tank of the batch at once, on the batch's stream (which the caller has made
// wait for every tank's set-up, with the descriptors and ranges uploaded): the first iteration, then chunks of
// `chunk` iterations (replayed as a captured graph) between reads of all tanks' states, until every tank's solve is
// done. Each pass is the single-tank pass applied to every tank, so each tank computes what its own solve would.
// The last read of each tank's state is left in its h_state.
void batch_iterate(BatchRun &B, slosh_ctx *const *tanks, uint32_t n, int chunk, bool graphs, hipGraphExec_t &graph,
                   SolverState *d_states, SolverState *h_states)
{
    chunk = chunk >= 2 ? chunk & ~1 : 4;
    int maxit = 0, ahead = 1;
    for (uint32_t k = 0; k < n; ++k) {
        const slosh_ctx *c = tanks[k];
        maxit = std::max(maxit, c->s.max_iterations > 0 ? c->s.max_iterations : 200);
        if (c->last_mg_iters > 1) ahead = std::max(ahead, (c->last_mg_iters + chunk - 1) / chunk);
    }
    int rc = 1, dc = 0;
    auto vcycle_b = [&](int ri) {
        b_restrict(B, ri);
        b_pooled_down1(B);
        b_d23down(B);
        b_coarse_tail(B);
        b_d23up(B);
        b_pooled_up1(B);
        b_prolong_black(B, ri);
        b_red_dot(B, ri);
        b_finalize(B, kFinRho);
    };
    // the first iteration's update (no alpha yet), V-cycle and direction, from the residual in r[0]
    b_update_black(B, 0, 1, 0, true);
    vcycle_b(rc);
    b_direction(B, 1, 0, rc, true);
    auto chunk_body = [&]() {
        for (int k = 0; k < chunk; ++k) {
            b_finalize(B, kFinAlphaStop);
            b_update_black(B, rc, rc ^ 1, dc, false);
            rc ^= 1;
            vcycle_b(rc);
            b_direction(B, dc, dc ^ 1, rc, false);
            dc ^= 1;
        }
    };
    SLOSH_HIP(hipGetLastError());
    bool graph_ready = false;
    int done_iters = 0;
    for (;;) {
        for (int k = 0; k < ahead && done_iters <= maxit; ++k, done_iters += chunk) {
            if (graphs && done_iters > 0) {
                if (!graph_ready) graph_ready = capture_batch_graph(B, graph, graphs, chunk_body);
                if (graph_ready) SLOSH_HIP(hipGraphLaunch(graph, B.stream));
            } else {
                chunk_body();
            }
            SLOSH_HIP(hipGetLastError());
        }
        ahead = 1;
        b_true_residual_if_done(B);
        b_finalize(B, kFinCheckDone);
        hipLaunchKernelGGL(kb_gather_states, dim3(n), dim3(64), 0u, B.stream, B.d_tanks, d_states);
        SLOSH_HIP(hipMemcpyAsync(h_states, d_states, n * sizeof(SolverState), hipMemcpyDeviceToHost, B.stream));
        SLOSH_HIP(hipStreamSynchronize(B.stream));
        bool all = true;
        for (uint32_t k = 0; k < n; ++k) all = all && h_states[k].done != 0;
        if (all || done_iters > maxit) break;
    }
    for (uint32_t k = 0; k < n; ++k) *tanks[k]->h_state = h_states[k];
}

// After batched iterations: restarts from r = b - A p while the FP64 true residual misses the tolerance (on the
// tank's own stream), as mgpcg_iterate does after its own iterations.
void mgpcg_resume(slosh_ctx *c)
{
    const int maxit = c->s.max_iterations > 0 ? c->s.max_iterations : 200;
    mgpcg_iterate(c, maxit, 3, false, true);
    c->last_mg_iters = c->h_state->iters;
}

} // namespace slosh
