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

enum FinKind { kFinInit = 0, kFinRho = 1, kFinAlpha = 2, kFinRr = 3, kFinSor = 4, kFinBetaF = 5, kFinAlphaF = 6,
               kFinCheck = 7 };

__host__ __device__ constexpr int fin_values(int kind)
{
    return kind == kFinInit ? 3 : ((kind == kFinSor || kind == kFinAlphaF || kind == kFinCheck) ? 2 : 1);
}

// Sums the partials of the preceding reduction in a fixed order (per thread: its strided run in increasing order;
// then a wavefront butterfly and the 16 wavefront sums in order) and applies them to the solver state.
__global__ __launch_bounds__(1024) void k_finalize(int kind, const double *partials, uint32_t nblocks, int iters,
                                                   SolverState *st)
{
    const int nv = fin_values(kind);
    __shared__ double sh[3][16];
    const uint32_t t = threadIdx.x;
    // the first 8 partials of this thread (b = t + 1024 k) are read together with the done flag; a thread sums its
    // partials in increasing b (an absent one adds +0, which leaves the sum unchanged: it starts at +0 and never
    // becomes -0)
    constexpr int kFirst = 8;
    const SolverState s0 = *st; // read with the partials; thread 0 updates the state from it
    const uint32_t done = (uint32_t)s0.done;
    double v[kFirst][3];
    for (int k = 0; k < kFirst; ++k) {
        const uint32_t b = t + 1024u * (uint32_t)k;
        for (int j = 0; j < 3; ++j) v[k][j] = (b < nblocks && j < nv) ? partials[(size_t)b * nv + j] : 0.0;
    }
    issue_point();
    if (kind != kFinInit && kind != kFinCheck && done) return;
    double a[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < kFirst; ++k)
        for (int j = 0; j < 3; ++j) a[j] += v[k][j];
    for (uint32_t b = t + 1024u * kFirst; b < nblocks; b += 1024u)
        for (int j = 0; j < nv; ++j) a[j] += partials[(size_t)b * nv + j];
    for (int w = 32; w > 0; w >>= 1)
        for (int j = 0; j < 3; ++j) a[j] += __shfl_xor(a[j], w);
    if ((t & 63u) == 0u)
        for (int j = 0; j < 3; ++j) sh[j][t >> 6] = a[j];
    __syncthreads();
    if (t != 0) return;
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
    for (int w = 0; w < 16; ++w) {
        t0 += sh[0][w];
        t1 += sh[1][w];
        t2 += sh[2][w];
    }
    if (kind == kFinInit) {
        st->nf = t2;
        st->mg_scale = t1 > 0.0 ? sqrt(t2 / t1) : 1.0;
    }
    if (kind == kFinInit || kind == kFinSor) {
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
    } else if (kind == kFinAlpha) {
        st->dq = t0;
        st->alpha = fabs(t0) > 1e-30 ? s0.rho / t0 : 0.0;
    } else if (kind == kFinBetaF) {
        // flexible CG: beta = -(z . q_prev) / (d_prev . q_prev)
        st->beta = fabs(s0.dq) > 1e-30 ? -t0 / s0.dq : 0.0;
    } else if (kind == kFinAlphaF) {
        // flexible CG: alpha = (d . r) / (d . q)
        st->dq = t0;
        st->alpha = fabs(t0) > 1e-30 ? t1 / t0 : 0.0;
    } else if (kind == kFinCheck) {
        st->rel_true = t0 > 0.0 ? sqrt(t1) / sqrt(t0) : 0.0;
    } else {
        st->rr = t0;
        st->mg_scale = t0 > 0.0 ? sqrt(s0.nf / t0) : 1.0;
        const double rel = sqrt(t0) / sqrt(s0.bb);
        st->rel = rel;
        const int it = s0.iters + 1;
        st->iters = it;
        if (rel <= s0.tol) {
            st->converged = 1;
            st->done = 1;
        } else if (it >= s0.maxit) {
            st->done = 1;
        }
    }
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

// ---- level 3 (one cell per brick, dense by brick index): one wavefront per 4^3 tile, the tile and a two-cell halo
// staged in LDS (the passes move little data; all of a tile's loads are issued at once, addresses from arithmetic).
constexpr int kD3T = 4;                       // tile edge
constexpr int kD3B = kD3T + 4;                // box edge: the tile and a two-cell halo
constexpr int kD3NB = kD3B * kD3B * kD3B;     // box cells
constexpr int kD3In = kD3T + 2;               // the tile and a one-cell halo
constexpr uint32_t kD3Threads = 64;
constexpr int kD3Per = kD3NB / (int)kD3Threads;

__device__ __forceinline__ int d3_box(int u, int v, int w) { return (u + 2) + kD3B * ((v + 2) + kD3B * (w + 2)); }

__device__ __forceinline__ float d3_box_sum(const float *v, int k, uint32_t info)
{
    float s = 0.0f;
    for (int n = 0; n < 6; ++n)
        if (info >> n & 1u) s += v[k + ((n < 2) ? 1 : (n < 4 ? kD3B : kD3B * kD3B)) * ((n & 1) ? 1 : -1)];
    return s;
}

// Level-3 cell of box cell k of tile (tx, ty, tz): its coordinates and whether it lies in the level.
__device__ __forceinline__ bool d3_cell(const DenseLevel &L, int tx, int ty, int tz, int k, int &x, int &y, int &z)
{
    x = tx * kD3T + (k % kD3B) - 2;
    y = ty * kD3T + ((k / kD3B) % kD3B) - 2;
    z = tz * kD3T + k / (kD3B * kD3B) - 2;
    return x >= 0 && y >= 0 && z >= 0 && x < L.dim[0] && y < L.dim[1] && z < L.dim[2];
}

__device__ __forceinline__ void d3_tile(const DenseLevel &L, int &tx, int &ty, int &tz)
{
    const int ntx = (L.dim[0] + kD3T - 1) / kD3T, nty = (L.dim[1] + kD3T - 1) / kD3T;
    tx = (int)blockIdx.x % ntx;
    ty = ((int)blockIdx.x / ntx) % nty;
    tz = (int)blockIdx.x / (ntx * nty);
}

// Level 3, down pass fused with the restriction into level 4 (MGSmoothCS black, then MGRestrictCS with the fused red
// sweep): the black sweep on the tile and its one-cell halo (from the staged red values), the residual of every
// tile cell (0 where not fluid), then b_p = 1/4 of the children's residuals summed in child order for the tile's 2^3
// level-4 cells and x_p = b_p / diag_p on red ones. Black values of the tile go back to memory.
__global__ __launch_bounds__(kD3Threads) void k_d3down(DenseLevel L, DenseLevel C, const SolverState *st)
{
    __shared__ float s_x[kD3NB], s_b[kD3NB];
    __shared__ uint16_t s_info[kD3NB];
    __shared__ float s_res[kD3T * kD3T * kD3T];
    int tx, ty, tz;
    d3_tile(L, tx, ty, tz);
    const uint32_t t = threadIdx.x;
    const uint32_t done = st->done;
    uint32_t vi[kD3Per];
    float vx[kD3Per], vb[kD3Per];
    for (int q = 0; q < kD3Per; ++q) {
        int x, y, z;
        const bool in = d3_cell(L, tx, ty, tz, (int)t + (int)kD3Threads * q, x, y, z);
        const uint32_t gi = in ? (uint32_t)(x + L.dim[0] * (y + L.dim[1] * z)) : 0u;
        vi[q] = in ? (uint32_t)L.info[gi] : 0u;
        vx[q] = L.x[gi];
        vb[q] = L.b[gi];
    }
    // the tile's level-4 cells: thread t < 8 takes child block t
    const int cx = 2 * tx + (int)(t & 1u), cy = 2 * ty + (int)((t >> 1) & 1u), cz = 2 * tz + (int)((t >> 2) & 1u);
    const bool cin = t < 8u && cx < C.dim[0] && cy < C.dim[1] && cz < C.dim[2];
    const uint32_t ci = cin ? (uint32_t)(cx + C.dim[0] * (cy + C.dim[1] * cz)) : 0u;
    const uint32_t cinfo = C.info[ci];
    issue_point();
    if (done) return;
    for (int q = 0; q < kD3Per; ++q) {
        const int k = (int)t + (int)kD3Threads * q;
        const bool f = (vi[q] & kInfoSelf) != 0u;
        s_info[k] = (uint16_t)vi[q];
        s_x[k] = f ? vx[q] : 0.0f;
        s_b[k] = f ? vb[q] : 0.0f;
    }
    __syncthreads();
    for (int k = (int)t; k < kD3In * kD3In * kD3In; k += (int)kD3Threads) {
        const int kb = d3_box(k % kD3In - 1, (k / kD3In) % kD3In - 1, k / (kD3In * kD3In) - 1);
        const uint32_t info = s_info[kb];
        if (!(info & kInfoSelf) || (info & kInfoRed)) continue;
        s_x[kb] = (s_b[kb] + d3_box_sum(s_x, kb, info)) / info_diag(info);
    }
    __syncthreads();
    {
        const int u = (int)(t % kD3T), v = (int)((t / kD3T) % kD3T), w = (int)(t / (kD3T * kD3T));
        const int kb = d3_box(u, v, w);
        const uint32_t info = s_info[kb];
        float r = 0.0f;
        if (info & kInfoSelf) {
            if (!(info & kInfoRed)) {
                const int x = tx * kD3T + u, y = ty * kD3T + v, z = tz * kD3T + w;
                L.x[(uint32_t)(x + L.dim[0] * (y + L.dim[1] * z))] = s_x[kb];
            }
            r = s_b[kb] - (info_diag(info) * s_x[kb] - d3_box_sum(s_x, kb, info));
        }
        s_res[t] = r;
    }
    __syncthreads();
    if (!cin) return;
    if (!(cinfo & kInfoSelf)) {
        C.b[ci] = 0.0f;
        C.x[ci] = 0.0f;
        return;
    }
    // children in child order; one outside level 3 has residual 0 (adding +0 leaves the sum, which starts at +0 and
    // never becomes -0, unchanged)
    float sum = 0.0f;
    for (uint32_t oi = 0; oi < 8u; ++oi) {
        const int u = 2 * (int)(t & 1u) + (int)(oi & 1u), v = 2 * (int)((t >> 1) & 1u) + (int)((oi >> 1) & 1u),
                  w = 2 * (int)((t >> 2) & 1u) + (int)(oi >> 2);
        sum += s_res[u + kD3T * (v + kD3T * w)];
    }
    const float bv = 0.25f * sum;
    C.b[ci] = bv;
    C.x[ci] = (cinfo & kInfoRed) ? bv / info_diag(cinfo) : 0.0f;
}

// Level 3, up pass from level 4 (MGProlongCS, then MGSmoothCS black and red): the prolonged down-sweep iterate
// x + x4(parent) (a parent contributes only when it is fluid) on the box, the black sweep on the tile and its
// one-cell halo, the red sweep on the tile; result in x1.
__global__ __launch_bounds__(kD3Threads) void k_d3up(DenseLevel L, DenseLevel P, const SolverState *st)
{
    __shared__ float s_v[kD3NB], s_b[kD3NB];
    __shared__ uint16_t s_info[kD3NB];
    int tx, ty, tz;
    d3_tile(L, tx, ty, tz);
    const uint32_t t = threadIdx.x;
    const uint32_t done = st->done;
    uint32_t vi[kD3Per], pinf[kD3Per];
    float vx[kD3Per], vb[kD3Per], px[kD3Per];
    for (int q = 0; q < kD3Per; ++q) {
        int x, y, z;
        const bool in = d3_cell(L, tx, ty, tz, (int)t + (int)kD3Threads * q, x, y, z);
        const uint32_t gi = in ? (uint32_t)(x + L.dim[0] * (y + L.dim[1] * z)) : 0u;
        const uint32_t pi = in ? (uint32_t)((x >> 1) + P.dim[0] * ((y >> 1) + P.dim[1] * (z >> 1))) : 0u;
        vi[q] = in ? (uint32_t)L.info[gi] : 0u;
        vx[q] = L.x[gi];
        vb[q] = L.b[gi];
        pinf[q] = P.info[pi];
        px[q] = P.x[pi];
    }
    issue_point();
    if (done) return;
    for (int q = 0; q < kD3Per; ++q) {
        const int k = (int)t + (int)kD3Threads * q;
        float val = 0.0f, bv = 0.0f;
        if (vi[q] & kInfoSelf) {
            val = vx[q];
            if (pinf[q] & kInfoSelf) val += px[q];
            bv = vb[q];
        }
        s_info[k] = (uint16_t)vi[q];
        s_v[k] = val;
        s_b[k] = bv;
    }
    __syncthreads();
    for (int k = (int)t; k < kD3In * kD3In * kD3In; k += (int)kD3Threads) {
        const int kb = d3_box(k % kD3In - 1, (k / kD3In) % kD3In - 1, k / (kD3In * kD3In) - 1);
        const uint32_t info = s_info[kb];
        if (!(info & kInfoSelf) || (info & kInfoRed)) continue;
        s_v[kb] = (s_b[kb] + d3_box_sum(s_v, kb, info)) / info_diag(info);
    }
    __syncthreads();
    const int u = (int)(t % kD3T), v = (int)((t / kD3T) % kD3T), w = (int)(t / (kD3T * kD3T));
    const int x = tx * kD3T + u, y = ty * kD3T + v, z = tz * kD3T + w;
    if (x >= L.dim[0] || y >= L.dim[1] || z >= L.dim[2]) return;
    const int kb = d3_box(u, v, w);
    const uint32_t info = s_info[kb];
    if (!(info & kInfoSelf)) return;
    L.x1[(uint32_t)(x + L.dim[0] * (y + L.dim[1] * z))] =
        (info & kInfoRed) ? (s_b[kb] + d3_box_sum(s_v, kb, info)) / info_diag(info) : s_v[kb];
}

// Tiles of the level-3 passes.
uint32_t d3_tiles(const DenseLevel &L)
{
    return (uint32_t)(((L.dim[0] + kD3T - 1) / kD3T) * ((L.dim[1] + kD3T - 1) / kD3T) * ((L.dim[2] + kD3T - 1) / kD3T));
}

__global__ void k_dprolong(DenseLevel F, DenseLevel C, const SolverState *st)
{
    if (st->done) return;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < F.cells) prolong_cell(F, C, i);
}

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
__global__ __launch_bounds__(1024) void k_coarse_tail(TailArgs T, const uint32_t *cidx,
                                                      const uint32_t *cn, const double *inv, const SolverState *st)
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

// ---- coarsest level: FP64 explicit inverse of the <= 64 fluid unknowns' operator (once per solve) ----
// The coarsest level is level 0 (mode 0, grids of <= 64 cells), a pooled level (mode 1) or a dense level (mode 2).
__global__ __launch_bounds__(256) void k_coarse_factor(Geo g, const uint32_t *ctype, const float *diag, int mode,
                                                       PLevel P, DenseLevel D, uint32_t *cidx, uint32_t *cn,
                                                       double *inv)
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
}

// Reduction of the partials of the preceding level-0 pass: one per brick with fluid, except the initial residual,
// which covers every allocated brick.
void finalize(slosh_ctx *c, int kind, int iters = 0)
{
    SLOSH_TIMED(c, K_FINALIZE);
    const uint32_t nblocks = kind == kFinInit ? c->nactive : c->nsolve;
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, c->stream, kind, c->partials, nblocks, iters, c->state);
    (void)fin_values;
}

void coarse_solve(slosh_ctx *c, LVec b, LVec x)
{
    SLOSH_TIMED(c, K_COARSE_SOLVE);
    hipLaunchKernelGGL(k_coarse_solve, dim3(1), dim3(64), 0, c->stream, c->coarse_idx, c->coarse_n, c->coarse_inv, b,
                       x, c->state);
}

// Levels 3 .. m-1 (m >= 5): level 3's down pass with the restriction into level 4, grid-wide passes for dense
// levels kFirstDense .. t-1 (only when the dense levels do not all fit the tail's LDS), the tail on levels
// t .. m-1, the grid-wide up passes, level 3's up pass.
void dense_part(slosh_ctx *c)
{
    const int m = c->nlevels;
    const DenseLevel L3 = dlevel_of(c, 3);
    {
        SLOSH_TIMED(c, K_D3DOWN);
        const DenseLevel L4 = dlevel_of(c, kFirstDense);
        hipLaunchKernelGGL(k_d3down, dim3(d3_tiles(L3)), dim3(kD3Threads), 0, c->stream, L3, L4, c->state);
    }
    int t = m - 1;
    uint32_t tail_cells = c->lv[m - 1].cells;
    while (t - 1 >= kFirstDense && tail_cells + c->lv[t - 1].cells <= kTailCells) tail_cells += c->lv[--t].cells;
    const int sweeps = coarse_sweeps(c);
    for (int l = kFirstDense; l < t; ++l) {
        DenseLevel F = dlevel_of(c, l), C = dlevel_of(c, l + 1);
        for (int k = 0; k < sweeps; ++k)
            for (uint32_t colour : {0u, 1u}) {
                if (k == 0 && colour == 0u) continue; // the first red sweep is fused into the restriction
                SLOSH_TIMED(c, K_DSMOOTH);
                hipLaunchKernelGGL(k_dsmooth, dim3(blocks_for(F.cells, 256)), dim3(256), 0, c->stream, F, colour,
                                   c->state);
            }
        SLOSH_TIMED(c, K_DRESTRICT);
        hipLaunchKernelGGL(k_drestrict, dim3(blocks_for(C.cells, 256)), dim3(256), 0, c->stream, F, C, c->state);
    }
    TailArgs T{};
    T.n = m - t;
    T.sweeps = sweeps;
    T.info = c->dense_info + c->lv[t].dense_off;
    T.coord = c->dense_coord + c->lv[t].dense_off;
    for (int k = 0; k < T.n; ++k) {
        T.L[k] = dlevel_of(c, t + k);
        T.off[k] = c->lv[t + k].dense_off - c->lv[t].dense_off; // levels are contiguous in the blocks
        T.total = T.off[k] + T.L[k].cells;
    }
    {
        SLOSH_TIMED(c, K_COARSE_TAIL);
        hipLaunchKernelGGL(k_coarse_tail, dim3(1), dim3(1024), 0, c->stream, T, c->coarse_idx, c->coarse_n,
                           c->coarse_inv, c->state);
    }
    for (int l = t - 1; l >= kFirstDense; --l) {
        DenseLevel F = dlevel_of(c, l), C = dlevel_of(c, l + 1);
        {
            SLOSH_TIMED(c, K_DPROLONG);
            hipLaunchKernelGGL(k_dprolong, dim3(blocks_for(F.cells, 256)), dim3(256), 0, c->stream, F, C, c->state);
        }
        for (int k = 0; k < sweeps; ++k)
            for (uint32_t colour : {1u, 0u}) {
                SLOSH_TIMED(c, K_DSMOOTH);
                hipLaunchKernelGGL(k_dsmooth, dim3(blocks_for(F.cells, 256)), dim3(256), 0, c->stream, F, colour,
                                   c->state);
            }
    }
    SLOSH_TIMED(c, K_D3UP);
    hipLaunchKernelGGL(k_d3up, dim3(d3_tiles(L3)), dim3(kD3Threads), 0, c->stream, L3, dlevel_of(c, kFirstDense),
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
        const int last_pooled = m - 2 < 2 ? m - 2 : 2; // pooled levels with a down pass (1, 2)
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
    l0_brick_info(c);
    const int m = c->nlevels;
    for (int l = 1; l < m; ++l) {
        if (c->lv[l].pooled) {
            pooled_build_type(c, l);
        } else if (c->lv[l - 1].pooled) {
            dense_build_type_from_pooled(c, l);
        } else {
            DenseLevel C = dlevel_of(c, l);
            SLOSH_TIMED(c, K_MG_SETUP);
            hipLaunchKernelGGL(k_build_type, dim3(blocks_for(C.cells, 256)), dim3(256), 0, c->stream, dlevel_of(c, l - 1),
                               C, fluid_coarsening(c));
        }
    }
    for (int l = 1; l < m; ++l) {
        if (c->lv[l].pooled) {
            pooled_build_info(c, l);
        } else {
            DenseLevel D = dlevel_of(c, l);
            SLOSH_TIMED(c, K_MG_SETUP);
            hipLaunchKernelGGL(k_dinfo, dim3(blocks_for(D.cells, 256)), dim3(256), 0, c->stream, D);
        }
    }
    const int mode = m == 1 ? 0 : (c->lv[m - 1].pooled ? 1 : 2);
    PLevel P = mode == 1 ? plevel_of(c, m - 1) : PLevel{};
    DenseLevel D = mode == 2 ? dlevel_of(c, m - 1) : DenseLevel{};
    SLOSH_TIMED(c, K_COARSE_FACTOR);
    hipLaunchKernelGGL(k_coarse_factor, dim3(1), dim3(256), 0, c->stream, c->geo, c->ctype, c->diag, mode, P, D,
                       c->coarse_idx, c->coarse_n, c->coarse_inv);
}

// Captures the launches of `body` on the context's stream into c->mg_graph (updating the executable graph in
// place when the topology is unchanged, else instantiating it anew). On any capture failure the launches are
// enqueued directly instead, graphs are switched off for the context, and false is returned.
template <typename F> bool capture_chunk_graph(slosh_ctx *c, F body)
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
    if (c->mg_graph) {
        hipGraphNode_t err_node = nullptr;
        hipGraphExecUpdateResult res;
        ok = hipGraphExecUpdate(c->mg_graph, g, &err_node, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
        if (!ok) {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(c->mg_graph);
            c->mg_graph = nullptr;
        }
    }
    if (!ok) ok = hipGraphInstantiate(&c->mg_graph, g, nullptr, nullptr, 0) == hipSuccess;
    (void)hipGraphDestroy(g);
    if (!ok) {
        (void)hipGetLastError();
        c->mg_graph = nullptr;
        c->use_graphs = false;
        body();
    }
    return ok;
}

} // namespace

// SloshProxy.cpp:807-1020 (AddPressureSolveMGPCG), to ||r|| <= tol ||b||.
void run_mgpcg(slosh_ctx *c)
{
    if (!c->nactive) return;
    const int maxit = c->s.max_iterations > 0 ? c->s.max_iterations : 200;
    hipLaunchKernelGGL(k_state_reset, dim3(1), dim3(1), 0, c->stream, c->state, (double)c->s.tolerance, maxit);
    mg_setup(c);
    float *R[2] = {c->r, c->r2};
    float *D[2] = {c->d, c->d2};
    l0_pcg_init(c, R[0]);
    finalize(c, kFinInit);
    if (!c->nsolve) {
        // no fluid cell: b = 0, the solve is converged as it stands
        c->capture_n = 0;
        return;
    }
    // residuals kept for training the learned preconditioner: the initial one and those after iterations 1, 2,
    // 3, 4, 6, 8, 11, 15, 20, ... (slots past the solve's last iteration are dropped afterwards)
    uint32_t ncap = 0;
    int next_cap = 0, cap_gap = 1;
    auto capture = [&](const float *r, int iteration) {
        if (!c->capture || ncap >= c->capture_cap || iteration != next_cap) return;
        SLOSH_HIP(hipMemcpyAsync(c->capture + (size_t)ncap * c->geo.pool_cells, r, (size_t)c->geo.pool_cells * 4u,
                                 hipMemcpyDeviceToDevice, c->stream));
        c->capture_iter[ncap++] = iteration;
        next_cap += cap_gap;
        if (iteration >= 4) cap_gap++;
    };
    capture(R[0], 0);
    l0_update_black(c, R[0], R[1], D[0], 1);
    int rc = 1, dc = 0;
    vcycle(c, R[rc]);
    l0_direction(c, D[1], D[0], 1);
    SLOSH_HIP(hipGetLastError());
    // Iterations are enqueued in chunks between reads of the device state (an even count, so the ping-pong
    // buffers return to the same roles). Without residual capture or kernel timing, a chunk is captured once per
    // solve as a HIP graph and replayed: every chunk of a solve launches the same kernels on the same buffers.
    const int chunk = c->solver_chunk >= 2 ? c->solver_chunk & ~1 : 4;
    const bool use_graph = c->use_graphs && !c->capture && c->timing_mask == 0;
    auto enqueue_chunk = [&](int it0) {
        for (int k = 0; k < chunk; ++k) {
            finalize(c, kFinAlpha);
            l0_update_black(c, R[rc], R[rc ^ 1], D[dc], 0);
            rc ^= 1;
            finalize(c, kFinRr);
            capture(R[rc], it0 + k + 1);
            vcycle(c, R[rc]);
            l0_direction(c, D[dc], D[dc ^ 1], 0);
            dc ^= 1;
        }
    };
    bool graph_ready = false;
    for (int it = 0; it <= maxit; it += chunk) {
        if (use_graph && c->use_graphs) {
            // on a capture failure the helper has enqueued the chunk directly and switched graphs off
            if (!graph_ready) graph_ready = capture_chunk_graph(c, [&] { enqueue_chunk(it); });
            if (graph_ready) SLOSH_HIP(hipGraphLaunch(c->mg_graph, c->stream));
        } else {
            enqueue_chunk(it);
        }
        SLOSH_HIP(hipGetLastError());
        SLOSH_HIP(hipMemcpyAsync(c->h_state, c->state, sizeof(SolverState), hipMemcpyDeviceToHost, c->stream));
        SLOSH_HIP(hipStreamSynchronize(c->stream));
        if (c->h_state->done) break;
    }
    // the true residual ||b - A p|| / ||b|| of the FP64 pressure, reported with the step's results
    l0_true_residual(c);
    finalize(c, kFinCheck);
    if (c->capture) {
        uint32_t valid = 0;
        while (valid < ncap && c->capture_iter[valid] <= c->h_state->iters) valid++;
        c->capture_n = valid;
    }
}

// Flexible CG (Notay 2000, FCG(1)) preconditioned by the network: z = M(r), d = z - (z.q_prev / d_prev.q_prev) d_prev,
// alpha = d.r / d.q. MGPCG is the fallback and the judge: when the relative residual has not halved over the last
// four iterations, or an iteration limit is reached, or the recurrence converges but the true residual
// ||b - A p|| / ||b|| of the FP64 pressure exceeds twice the tolerance, MGPCG continues from the current pressure
// to the tolerance. Every accepted solve therefore meets the MGPCG stopping rule.
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
    l0_brick_info(c);
    float *D[2] = {c->d, c->d2};
    float *r = c->r;
    l0_pcg_init(c, r);
    finalize(c, kFinInit);
    if (!c->nsolve) return;
    int dc = 0;
    double hist[4] = {0, 0, 0, 0};
    bool fallback = false;
    for (int it = 0; it < maxit; ++it) {
        unet_apply(c, r, c->fcg_z);
        if (it > 0) {
            l0_fcg_zq(c, c->fcg_z);
            finalize(c, kFinBetaF);
        }
        l0_fcg_direction(c, c->fcg_z, D[dc], D[dc ^ 1], r, it == 0);
        dc ^= 1;
        finalize(c, kFinAlphaF);
        l0_fcg_update(c, r, D[dc]);
        finalize(c, kFinRr);
        SLOSH_HIP(hipGetLastError());
        SLOSH_HIP(hipMemcpyAsync(c->h_state, c->state, sizeof(SolverState), hipMemcpyDeviceToHost, c->stream));
        SLOSH_HIP(hipStreamSynchronize(c->stream));
        const SolverState &h = *c->h_state;
        c->fcg_iters = h.iters;
        if (!std::isfinite(h.rel)) {
            launch_fill_f64(c->stream, c->pres, 0.0, c->geo.pool_cells);
            fallback = true;
            break;
        }
        if (h.done) {
            fallback = !h.converged;
            break;
        }
        if (it >= 4 && h.rel > 0.5 * hist[it % 4]) {
            fallback = true;
            break;
        }
        hist[it % 4] = h.rel;
        if (it + 1 == maxit) fallback = true;
    }
    if (!fallback) {
        // MGPCG as the judge: the true residual of the accepted pressure
        l0_true_residual(c);
        finalize(c, kFinCheck);
        SLOSH_HIP(hipMemcpyAsync(c->h_state, c->state, sizeof(SolverState), hipMemcpyDeviceToHost, c->stream));
        SLOSH_HIP(hipStreamSynchronize(c->stream));
        if (!(c->h_state->rel_true <= 2.0 * (double)c->s.tolerance)) fallback = true;
    }
    if (fallback) {
        c->fcg_fallback = 1;
        run_mgpcg(c);
    }
}

// Slosh.usf:856-892 (PressureRBGSCS) iterated to the tolerance, residual checked every 8 iterations.
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
            for (int k = 0; k < 8 && it < maxit; ++k, ++it) {
                l0_sor_sweep(c, 0u, w);
                l0_sor_sweep(c, 1u, w);
            }
            residual(it);
        }
        SLOSH_HIP(hipGetLastError());
        SLOSH_HIP(hipMemcpyAsync(c->h_state, c->state, sizeof(SolverState), hipMemcpyDeviceToHost, c->stream));
        SLOSH_HIP(hipStreamSynchronize(c->stream));
        if (c->h_state->done) break;
    }
}

} // namespace slosh
