// grid.hip -- grid kernels of the solver step on gfx950: brick pool, normalise, divergence, projection,
// extrapolation, post-projection divergence statistics, and the scan / fill utilities they use.
//
// Grid passes run one 512-thread workgroup (8 waves) per allocated brick: thread l owns cell l of the brick's
// pool slot, so every per-cell array is read and written as contiguous 2 KB (FP32) runs.
#include "slosh_ctx.h"

namespace slosh {

namespace {

inline uint32_t blocks_for(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_zero3_u32(uint32_t *a, uint32_t *b, uint32_t *c, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        a[i] = 0u;
        b[i] = 0u;
        c[i] = 0u;
    }
}
__global__ void k_fill_u8(uint8_t *p, uint8_t v, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_f64(double *p, double v, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---- exclusive scan of u32: 1024 elements per workgroup, block sums scanned by one workgroup ----------------
constexpr uint32_t kScanBlock = 1024;

__device__ uint32_t block_scan_excl_256(uint32_t v, uint32_t *sh, uint32_t *total)
{
    const uint32_t t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (uint32_t off = 1; off < 256; off <<= 1) {
        uint32_t add = t >= off ? sh[t - off] : 0u;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    uint32_t incl = sh[t];
    *total = sh[255];
    __syncthreads();
    return incl - v;
}

// Exclusive scan of n <= kScanOneMax values by one workgroup, out[n] = the total. Values are staged in LDS with one
// pad word after every 32 (value i at i + i/32), so the 32 consecutive values a thread scans lie in distinct banks;
// each thread scans its run serially, the 1024 run totals are scanned by wavefront shuffles and one LDS stage.
constexpr uint32_t kScanOneThreads = 1024u;
constexpr uint32_t kScanOneMax = 33u * kScanOneThreads;
constexpr uint32_t kScanOneLds = kScanOneMax + kScanOneMax / 32u;

__device__ __forceinline__ uint32_t scan_pad(uint32_t i) { return i + (i >> 5); }

// Two independent scans can share a launch: workgroup 1 scans (in1, n1) into out1.
__global__ __launch_bounds__(kScanOneThreads) void k_scan_one(const uint32_t *in0, uint32_t n0, uint32_t *out0,
                                                              const uint32_t *in1, uint32_t n1, uint32_t *out1)
{
    const uint32_t *const in = blockIdx.x ? in1 : in0;
    const uint32_t n = blockIdx.x ? n1 : n0;
    uint32_t *const out = blockIdx.x ? out1 : out0;
    constexpr uint32_t T = kScanOneThreads, U = kScanOneMax / kScanOneThreads, W = 64u;
    __shared__ uint32_t v[kScanOneLds];
    __shared__ uint32_t wsum[T / W];
    const uint32_t t = threadIdx.x;
    {
        // every load in flight before the first store
        uint32_t x[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) x[u] = t + T * u < n ? in[t + T * u] : 0u;
#pragma unroll
        for (uint32_t u = 0; u < U; ++u)
            if (t + T * u < n) v[scan_pad(t + T * u)] = x[u];
    }
    __syncthreads();
    const uint32_t per = (n + T - 1u) / T;
    const uint32_t beg = min(t * per, n), end = min(beg + per, n);
    uint32_t sum = 0u;
    for (uint32_t i = beg; i < end; ++i) sum += v[scan_pad(i)];
    // inclusive scan over the wavefront, then over the wavefront totals
    const uint32_t lane = t % W, wave = t / W;
    uint32_t inc = sum;
    for (uint32_t off = 1; off < W; off <<= 1) {
        const uint32_t y = __shfl_up(inc, off, W);
        if (lane >= off) inc += y;
    }
    if (lane == W - 1u) wsum[wave] = inc;
    __syncthreads();
    if (wave == 0) {
        uint32_t w = lane < T / W ? wsum[lane] : 0u;
        for (uint32_t off = 1; off < T / W; off <<= 1) {
            const uint32_t y = __shfl_up(w, off, W);
            if (lane >= off) w += y;
        }
        if (lane < T / W) wsum[lane] = w; // inclusive
    }
    __syncthreads();
    uint32_t run = inc - sum + (wave ? wsum[wave - 1u] : 0u);
    for (uint32_t i = beg; i < end; ++i) {
        const uint32_t x = v[scan_pad(i)];
        v[scan_pad(i)] = run;
        run += x;
    }
    if (t == 0) out[n] = wsum[T / W - 1u];
    __syncthreads();
    for (uint32_t i = t; i < n; i += T) out[i] = v[scan_pad(i)];
}

__global__ __launch_bounds__(256) void k_scan_reduce(const uint32_t *in, uint32_t n, uint32_t *block_sums)
{
    __shared__ uint32_t sh[256];
    uint32_t base = blockIdx.x * kScanBlock + threadIdx.x * 4u;
    uint32_t s = 0;
    for (uint32_t k = 0; k < 4; ++k)
        if (base + k < n) s += in[base + k];
    uint32_t total;
    block_scan_excl_256(s, sh, &total);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(256) void k_scan_blocks(uint32_t *block_sums, uint32_t nb)
{
    __shared__ uint32_t sh[256];
    const uint32_t per = (nb + 255u) / 256u;
    const uint32_t beg = threadIdx.x * per;
    uint32_t s = 0;
    for (uint32_t k = beg; k < beg + per && k < nb; ++k) s += block_sums[k];
    uint32_t total;
    uint32_t run = block_scan_excl_256(s, sh, &total);
    for (uint32_t k = beg; k < beg + per && k < nb; ++k) {
        uint32_t v = block_sums[k];
        block_sums[k] = run;
        run += v;
    }
}

__global__ __launch_bounds__(256) void k_scan_down(const uint32_t *in, uint32_t n, const uint32_t *block_sums,
                                                   uint32_t *out)
{
    __shared__ uint32_t sh[256];
    uint32_t base = blockIdx.x * kScanBlock + threadIdx.x * 4u;
    uint32_t v[4], s = 0;
    for (uint32_t k = 0; k < 4; ++k) {
        v[k] = base + k < n ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t run = block_scan_excl_256(s, sh, &total) + block_sums[blockIdx.x];
    for (uint32_t k = 0; k < 4; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 255) out[n] = run;
}

// ---- brick pool (Slosh.usf:1719-1831), deterministic slot assignment -----------------------------------------
// Want = any touched brick in the 3x3x3 brick neighbourhood; allocated bricks no longer wanted free their slot.
__global__ void k_pool_want(Geo g, const uint8_t *touch, uint32_t *table, uint32_t *slot_brick, uint32_t *need)
{
    uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= g.nbricks) return;
    int c[3];
    brick_coord(g, b, c);
    bool want = false;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                int x = c[0] + dx, y = c[1] + dy, z = c[2] + dz;
                if (x < 0 || y < 0 || z < 0 || x >= g.bd[0] || y >= g.bd[1] || z >= g.bd[2]) continue;
                want = want || touch[(uint32_t)(x + g.bd[0] * (y + g.bd[1] * z))] != 0u;
            }
    uint32_t slot = table[b];
    if (!want && slot != 0u) {
        table[b] = 0u;
        slot_brick[slot] = kNoBrick;
        slot = 0u;
    }
    need[b] = (want && slot == 0u) ? 1u : 0u;
}

__global__ void k_pool_freeflags(const uint32_t *slot_brick, uint32_t npool, uint32_t *freeflag)
{
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= npool) return;
    freeflag[s] = (s >= 1u && slot_brick[s] == kNoBrick) ? 1u : 0u;
}

__global__ void k_pool_free_list(const uint32_t *freeflag, const uint32_t *free_rank, uint32_t npool, uint32_t *free_list)
{
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= npool) return;
    if (freeflag[s]) free_list[free_rank[s]] = s;
}

// Wanted bricks without storage take the lowest free slots in ascending brick order; the rest overflow.
__global__ void k_pool_assign(Geo g, const uint32_t *need, const uint32_t *need_rank, const uint32_t *free_list,
                              const uint32_t *free_rank, uint32_t npool, uint32_t *table, uint32_t *slot_brick,
                              uint8_t *new_slot, uint8_t *touch, uint32_t *alloc_flag, StepStats *stats)
{
    uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= g.nbricks) return;
    touch[b] = 0u;
    uint32_t slot = table[b];
    if (need[b]) {
        uint32_t j = need_rank[b];
        if (j < free_rank[npool]) {
            slot = free_list[j];
            table[b] = slot;
            slot_brick[slot] = b;
            new_slot[slot] = 1u;
        } else {
            atomicAdd(&stats->overflow, 1u);
        }
    }
    alloc_flag[b] = slot != 0u ? 1u : 0u;
}

__global__ void k_pool_active(Geo g, const uint32_t *alloc_flag, const uint32_t *alloc_rank, uint32_t *active,
                              StepStats *stats)
{
    uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b == 0) stats->nactive = alloc_rank[g.nbricks];
    if (b >= g.nbricks) return;
    if (alloc_flag[b]) active[alloc_rank[b]] = b;
}

// Per allocated brick (active order), once per step after the pool: the slots of its 3x3x3 brick neighbourhood
// (word (dx+1) + 3(dy+1) + 9(dz+1); 0 outside the grid or unallocated, as the brick table gives) and, in word 27,
// its brick index. One thread per word.
__global__ void k_brick_rows(Geo g, uint32_t first, const uint32_t *nactive, uint32_t *rows)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t e = first + t / kBinfoWords, k = t % kBinfoWords;
    if (e >= *nactive) return;
    const uint32_t b = g.active[e];
    uint32_t v = 0u;
    if (k == 27u) {
        v = b;
    } else if (k < 27u) {
        int bc[3];
        brick_coord(g, b, bc);
        const int x = bc[0] + (int)(k % 3u) - 1, y = bc[1] + (int)(k / 3u % 3u) - 1, z = bc[2] + (int)(k / 9u) - 1;
        if (x >= 0 && y >= 0 && z >= 0 && x < g.bd[0] && y < g.bd[1] && z < g.bd[2])
            v = g.table[(uint32_t)(x + g.bd[0] * (y + g.bd[1] * z))];
    }
    rows[(size_t)e * kBinfoWords + k] = v;
}

// ---- per-brick helpers -------------------------------------------------------------------------------------
// One workgroup per allocated brick, one thread per cell. The brick's row of neighbour slots is staged in LDS, so
// a neighbour's pool index costs one LDS read instead of a brick-table load that the data load has to wait for.
struct CellPos {
    int x, y, z;
    uint32_t idx;
    uint32_t slot;
    int bc[3];
    const uint32_t *row; // LDS copy of the brick's neighbourhood row
};

// Workgroup -> active brick e (bricks of one XCD contiguous); stages the row (ends with a barrier).
__device__ __forceinline__ bool cell_of_block(const Geo &g, const uint32_t *rows, uint32_t *s_row, CellPos &c,
                                              uint32_t &e)
{
    e = xcd_remap(blockIdx.x, gridDim.x);
    if (threadIdx.x < 28u) s_row[threadIdx.x] = rows[(size_t)e * kBinfoWords + threadIdx.x];
    __syncthreads();
    const uint32_t b = s_row[27];
    const uint32_t l = threadIdx.x;
    brick_coord(g, b, c.bc);
    c.x = c.bc[0] * 8 + (int)(l & 7u);
    c.y = c.bc[1] * 8 + (int)((l >> 3) & 7u);
    c.z = c.bc[2] * 8 + (int)(l >> 6);
    c.slot = s_row[13];
    c.idx = c.slot * kBrickCells + l;
    c.row = s_row;
    return in_grid(g, c.x, c.y, c.z);
}

// Slot of the brick holding cell (x, y, z), which lies in the 3x3x3 neighbourhood of c's brick.
__device__ __forceinline__ uint32_t nb_slot(const CellPos &c, int x, int y, int z)
{
    const int k = ((x >> 3) - c.bc[0] + 1) + 3 * ((y >> 3) - c.bc[1] + 1) + 9 * ((z >> 3) - c.bc[2] + 1);
    return c.row[k];
}

// CellIndex (Slosh.usf:136-144) for a cell of the neighbourhood: the same slot the brick table gives.
__device__ __forceinline__ uint32_t cell_index(const Geo &g, const CellPos &c, int x, int y, int z)
{
    const uint32_t slot = in_grid(g, x, y, z) ? nb_slot(c, x, y, z) : 0u;
    return slot * kBrickCells + local_of_cell(x, y, z);
}

__device__ __forceinline__ uint32_t type_at(const Geo &g, const CellPos &c, const uint32_t *ctype, int x, int y, int z)
{
    return in_grid(g, x, y, z) ? ctype[cell_index(g, c, x, y, z)] : kCellSolid;
}

__device__ __forceinline__ uint32_t face_index(const Geo &g, const CellPos &c, int axis, int x, int y, int z)
{
    return (uint32_t)axis * g.pool_cells + cell_index(g, c, x, y, z);
}

// Slosh.usf:232-241
__device__ __forceinline__ float solid_face_velocity(const slosh_collider *col, uint32_t packed, int axis,
                                                     const float fp[3])
{
    uint32_t id = (packed >> kColliderShift) & 0xFFu;
    if (id == 0u) return 0.0f;
    return collider_velocity_axis(col[id - 1u], fp, axis);
}

// Slosh.usf:604-651 (NormalizeFace, APIC) over the P2G tiles: each face sums the tiles of the (up to 8) bricks
// whose node range covers it, then ratio, weight, gravity and solid boundary velocities as the reference.
__global__ __launch_bounds__(512) void k_normalize(Geo g, const uint32_t *rows, const int32_t *tiles,
                                                   const uint8_t *tile_valid, const uint32_t *ctype, float *vel,
                                                   float *fweight, uint8_t *fflags, float gdt0, float gdt1, float gdt2,
                                                   const slosh_collider *col)
{
    __shared__ uint32_t s_row[28];
    CellPos c;
    uint32_t e;
    if (!cell_of_block(g, rows, s_row, c, e)) return;
    const int lc[3] = {c.x & 7, c.y & 7, c.z & 7};
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = lc[a] == 0 ? -1 : 0;
        hi[a] = lc[a] == 7 ? 1 : 0;
    }
    int32_t mom[3] = {0, 0, 0}, wt[3] = {0, 0, 0};
    for (int dz = lo[2]; dz <= hi[2]; ++dz)
        for (int dy = lo[1]; dy <= hi[1]; ++dy)
            for (int dx = lo[0]; dx <= hi[0]; ++dx) {
                int nb[3] = {c.bc[0] + dx, c.bc[1] + dy, c.bc[2] + dz};
                const uint32_t ns = c.row[(dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)]; // 0 outside the grid
                if (ns == 0u || !tile_valid[ns]) continue;
                int tx = c.x - (nb[0] * 8 - 1), ty = c.y - (nb[1] * 8 - 1), tz = c.z - (nb[2] * 8 - 1);
                const int32_t *t = tiles + (size_t)ns * kTileInts + (tx + kTileDim * (ty + kTileDim * tz));
                for (int a = 0; a < 3; ++a) {
                    mom[a] += t[a * kTileNodes];
                    wt[a] += t[(3 + a) * kTileNodes];
                }
            }
    const float gdt[3] = {gdt0, gdt1, gdt2};
    const uint32_t tb = ctype[c.idx];
    for (int a = 0; a < 3; ++a) {
        uint32_t f = (uint32_t)a * g.pool_cells + c.idx;
        float u = wt[a] > 0 ? (float)mom[a] / (float)wt[a] : 0.0f;
        fweight[f] = (float)wt[a] * kFixedInv;
        uint32_t flags = wt[a] > 0 ? kFaceHasWeight : 0u;
        int ca[3] = {c.x, c.y, c.z};
        ca[a] -= 1;
        uint32_t ta = type_at(g, c, ctype, ca[0], ca[1], ca[2]);
        uint32_t sa = ta & kCellTypeMask, sb = tb & kCellTypeMask;
        float fp[3] = {(float)c.x + 0.5f, (float)c.y + 0.5f, (float)c.z + 0.5f};
        fp[a] = (float)(a == 0 ? c.x : (a == 1 ? c.y : c.z));
        for (int k = 0; k < 3; ++k) fp[k] = fp[k] * g.h;
        if (sa == kCellSolid && sb == kCellSolid) {
            u = solid_face_velocity(col, tb, a, fp);
        } else if (sa == kCellSolid || sb == kCellSolid) {
            u = solid_face_velocity(col, sb == kCellSolid ? tb : ta, a, fp);
            flags |= kFaceValid | kFaceSolid;
        } else if (sa == kCellFluid || sb == kCellFluid) {
            u += gdt[a];
            flags |= kFaceValid;
        } else {
            u = 0.0f;
        }
        vel[f] = u;
        fflags[f] = (uint8_t)flags;
    }
}

// Slosh.usf:211-224
__device__ __forceinline__ float cell_density(const Geo &g, const CellPos &c, const float *fw, int x, int y, int z)
{
    if (!in_grid(g, x, y, z)) return 0.0f;
    float ws = 0.0f;
    for (int a = 0; a < 3; ++a) {
        int p[3] = {x, y, z};
        p[a] += 1;
        ws += fw[face_index(g, c, a, x, y, z)] + fw[face_index(g, c, a, p[0], p[1], p[2])];
    }
    return ws / (6.0f * kTargetPpc);
}

__device__ __forceinline__ float air_face_weight(const Geo &g, const CellPos &c, const float *fw, int ghost,
                                                 float rho_fluid, int x, int y, int z)
{
    return ghost ? 1.0f / ghost_theta(rho_fluid, cell_density(g, c, fw, x, y, z)) : 1.0f;
}

// Slosh.usf:692-741 (DivergenceCS)
// Also writes the solver's per-cell stencil mask: bit n (0..5, order -x +x -y +y -z +z) = neighbour n is fluid,
// bit 6 = the cell itself is fluid (0 for every other cell).
__global__ __launch_bounds__(512) void k_divergence(Geo g, const uint32_t *rows, const float *vel, const float *fw,
                                                    const uint32_t *ctype, float *div, float *diag, uint8_t *fmask,
                                                    int ghost, float dc, float deadband, float h, float dt)
{
    __shared__ uint32_t s_row[28];
    CellPos c;
    uint32_t e;
    if (!cell_of_block(g, rows, s_row, c, e) || (ctype[c.idx] & kCellTypeMask) != kCellFluid) {
        // cells of a partial brick beyond the grid get a zero mask too (their slot may be recycled)
        div[c.idx] = 0.0f;
        diag[c.idx] = 0.0f;
        fmask[c.idx] = 0u;
        return;
    }
    float dv = 0.0f;
    for (int a = 0; a < 3; ++a) {
        int p[3] = {c.x, c.y, c.z};
        p[a] += 1;
        dv += vel[face_index(g, c, a, p[0], p[1], p[2])] - vel[(uint32_t)a * g.pool_cells + c.idx];
    }
    float ws = 0.0f;
    for (int a = 0; a < 3; ++a) {
        int p[3] = {c.x, c.y, c.z};
        p[a] += 1;
        ws += fw[(uint32_t)a * g.pool_cells + c.idx] + fw[face_index(g, c, a, p[0], p[1], p[2])];
    }
    float rho = ws / (6.0f * kTargetPpc);
    float excess = fmaxf(rho - (1.0f + deadband), 0.0f);
    float target = fminf(dc * excess, 0.5f) * h / dt;
    div[c.idx] = dv - target;
    float dg = 0.0f;
    uint32_t mask = 1u << 6;
    for (int nb = 0; nb < 6; ++nb) {
        int d[3];
        nb_offset(nb, d);
        int nx = c.x + d[0], ny = c.y + d[1], nz = c.z + d[2];
        uint32_t t = type_at(g, c, ctype, nx, ny, nz) & kCellTypeMask;
        if (t == kCellFluid) mask |= 1u << nb;
        dg += t == kCellFluid ? 1.0f : (t == kCellAir ? air_face_weight(g, c, fw, ghost, rho, nx, ny, nz) : 0.0f);
    }
    diag[c.idx] = dg;
    fmask[c.idx] = (uint8_t)mask;
}

// Slosh.usf:995-1052 (ProjectFace) on the FP64 pressure, without collider-coupling accumulation.
__global__ __launch_bounds__(512) void k_project(Geo g, const uint32_t *rows, float *vel, const float *fw,
                                                 const uint32_t *ctype, const double *pres, uint8_t *fflags, int ghost)
{
    __shared__ uint32_t s_row[28];
    CellPos c;
    uint32_t e;
    if (!cell_of_block(g, rows, s_row, c, e)) return;
    const uint32_t tb = ctype[c.idx] & kCellTypeMask;
    for (int a = 0; a < 3; ++a) {
        uint32_t f = (uint32_t)a * g.pool_cells + c.idx;
        int ca[3] = {c.x, c.y, c.z};
        ca[a] -= 1;
        uint32_t ta = type_at(g, c, ctype, ca[0], ca[1], ca[2]) & kCellTypeMask;
        uint32_t flags = (uint32_t)fflags[f] & kFaceHasWeight;
        if (ta == kCellSolid && tb == kCellSolid) {
        } else if (ta == kCellSolid || tb == kCellSolid) {
            flags |= kFaceValid | kFaceSolid;
        } else if (ta == kCellFluid || tb == kCellFluid) {
            double pa, pb;
            if (ta == kCellFluid && tb == kCellFluid) {
                pa = pres[cell_index(g, c, ca[0], ca[1], ca[2])];
                pb = pres[c.idx];
            } else if (ta == kCellFluid) {
                pa = pres[cell_index(g, c, ca[0], ca[1], ca[2])];
                pb = pa - pa * (double)air_face_weight(g, c, fw, ghost, cell_density(g, c, fw, ca[0], ca[1], ca[2]),
                                                       c.x, c.y, c.z);
            } else {
                pb = pres[c.idx];
                pa = pb - pb * (double)air_face_weight(g, c, fw, ghost, cell_density(g, c, fw, c.x, c.y, c.z), ca[0],
                                                       ca[1], ca[2]);
            }
            vel[f] -= (float)(pb - pa);
            flags |= kFaceValid;
        }
        fflags[f] = (uint8_t)flags;
    }
}

// Slosh.usf:657-686 (ViscosityJacobiFace): one Jacobi iteration of (I - nu dt Laplacian) u = u* on the valid, non-solid
// faces (ustar = u*, src the current iterate); a neighbour face outside the grid or not valid reads as the centre.
__global__ __launch_bounds__(512) void k_viscosity(Geo g, const uint32_t *rows, const float *ustar, const float *src,
                                                   float *dst, const uint8_t *fflags, float alpha)
{
    __shared__ uint32_t s_row[28];
    CellPos c;
    uint32_t e;
    const bool in = cell_of_block(g, rows, s_row, c, e);
    for (int a = 0; a < 3; ++a) {
        const uint32_t f = (uint32_t)a * g.pool_cells + c.idx;
        const float center = src[f];
        const uint32_t flags = fflags[f];
        if (!in || !(flags & kFaceValid) || (flags & kFaceSolid)) {
            dst[f] = center;
            continue;
        }
        float sum = 0.0f;
        for (int nb = 0; nb < 6; ++nb) {
            int d[3];
            nb_offset(nb, d);
            const int nx = c.x + d[0], ny = c.y + d[1], nz = c.z + d[2];
            float value = center;
            if (in_grid(g, nx, ny, nz)) {
                const uint32_t ni = face_index(g, c, a, nx, ny, nz);
                if (fflags[ni] & kFaceValid) value = src[ni];
            }
            sum += value;
        }
        dst[f] = (ustar[f] + alpha * sum) / (1.0f + 6.0f * alpha);
    }
}

// ---- collider coupling (Slosh.usf:909-1216, 1316-1370) ----------------------------------------------------------
// Per two-way collider k, 16 int32 sums in fixed point (1024 per unit; integer adds, so any order gives the same
// sums): 0-2 pressure force (N), 3-5 torque (N m), 6 wetted faces, 7 submerged cells, 8-10 fluid velocity over the
// wetted faces, 11 solid cells, 12-14 submerged-cell offset from the centre.
constexpr float kCouplingScale = 1024.0f;

__device__ __forceinline__ uint32_t collider_of(uint32_t packed) { return (packed >> kColliderShift) & 0xFFu; }

// Cell type of any cell of the grid (walks leave the brick neighbourhood): through the brick table, SOLID outside.
__device__ __forceinline__ uint32_t type_global(const Geo &g, const uint32_t *ctype, int x, int y, int z)
{
    if (!in_grid(g, x, y, z)) return kCellSolid;
    const uint32_t b = (uint32_t)((x >> 3) + g.bd[0] * ((y >> 3) + g.bd[1] * (z >> 3)));
    return ctype[g.table[b] * kBrickCells + local_of_cell(x, y, z)];
}
__device__ __forceinline__ uint32_t index_global(const Geo &g, int x, int y, int z)
{
    const uint32_t b = (uint32_t)((x >> 3) + g.bd[0] * ((y >> 3) + g.bd[1] * (z >> 3)));
    return g.table[b] * kBrickCells + local_of_cell(x, y, z);
}

__device__ __forceinline__ void add_fixed(int32_t *p, float v) { atomicAdd(p, (int32_t)rintf(v * kCouplingScale)); }

// Slosh.usf:913-926 (AccumulatePressureForce): force p_hat * PressureToPascal * h^2 along the direction into the
// solid (`sign` times axis `axis`), torque about the collider's centre.
__device__ __forceinline__ void pressure_force(const Geo &g, const slosh_collider &col, int32_t *acc, const float fp[3],
                                               int axis, float sign, float p_hat, float p2pa)
{
    const float pa = p_hat * p2pa;
    float f[3], r[3], t[3];
    for (int a = 0; a < 3; ++a) {
        f[a] = pa * g.h * g.h * (a == axis ? sign : 0.0f);
        r[a] = fp[a] - col.center[a];
    }
    cross3(r, f, t);
    for (int a = 0; a < 3; ++a) add_fixed(&acc[a], f[a]);
    for (int a = 0; a < 3; ++a) add_fixed(&acc[3 + a], t[a]);
}

// Slosh.usf:951-983 (EstimateSeepagePressure): the mean pressure of the liquid met first by walks of up to 16 cells
// along +x, -x, +y, -y out of collider `id`'s solid region (the reference's horizontal axes); false when dry.
__device__ __forceinline__ bool seepage_pressure(const Geo &g, const uint32_t *ctype, const double *pres, int x, int y, int z,
                                 uint32_t id, float &p_out)
{
    // directions +x, -x, +y, -y
    float sum = 0.0f, cnt = 0.0f;
    for (int d = 0; d < 4; ++d) {
        int px = x, py = y;
        for (int s = 0; s < 16; ++s) {
            px += d == 0 ? 1 : (d == 1 ? -1 : 0);
            py += d == 2 ? 1 : (d == 3 ? -1 : 0);
            const uint32_t t = type_global(g, ctype, px, py, z);
            if ((t & kCellTypeMask) != kCellSolid || collider_of(t) != id) {
                if ((t & kCellTypeMask) == kCellFluid) {
                    sum += (float)pres[index_global(g, px, py, z)];
                    cnt += 1.0f;
                }
                break;
            }
        }
    }
    p_out = cnt > 0.0f ? sum / cnt : 0.0f;
    return cnt > 0.0f;
}

// Slosh.usf:985-998 (AccumulateContact): the seepage pressure on a contact face of a two-way body with another solid.
__device__ __forceinline__ void contact(const Geo &g, const uint32_t *ctype, const double *pres, const slosh_collider *col,
                        int32_t *coupling, uint32_t packed, int x, int y, int z, const float fp[3], int axis, float sign,
                        float p2pa)
{
    const uint32_t id = collider_of(packed);
    if (id == 0u || !(col[id - 1u].flags & 1u)) return;
    float p;
    if (seepage_pressure(g, ctype, pres, x, y, z, id, p))
        pressure_force(g, col[id - 1u], coupling + (id - 1u) * 16u, fp, axis, sign, p, p2pa);
}

// ProjectFace's coupling terms (Slosh.usf:1003-1028) after the projection: contact faces between different solids
// get the seepage pressure; a fluid face of a two-way body adds to its wetted faces and fluid velocity (the fluid
// cell's centre velocity after the projection) and, with the binary estimator, its face pressure force.
__global__ __launch_bounds__(512) void k_coupling_faces(Geo g, const uint32_t *rows, const uint32_t *ctype,
                                                        const double *pres, const float *vel, const slosh_collider *col,
                                                        int32_t *coupling, float p2pa, int binary)
{
    __shared__ uint32_t s_row[28];
    CellPos c;
    uint32_t e;
    if (!cell_of_block(g, rows, s_row, c, e)) return;
    const uint32_t B = ctype[c.idx], tb = B & kCellTypeMask;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        int ca[3] = {c.x, c.y, c.z};
        ca[a] -= 1;
        const uint32_t A = type_at(g, c, ctype, ca[0], ca[1], ca[2]), ta = A & kCellTypeMask;
        float fp[3] = {((float)c.x + 0.5f) * g.h, ((float)c.y + 0.5f) * g.h, ((float)c.z + 0.5f) * g.h};
        fp[a] = (float)(a == 0 ? c.x : (a == 1 ? c.y : c.z)) * g.h;
        if (ta == kCellSolid && tb == kCellSolid) {
            if (collider_of(A) != collider_of(B)) {
                contact(g, ctype, pres, col, coupling, B, c.x, c.y, c.z, fp, a, 1.0f, p2pa);
                contact(g, ctype, pres, col, coupling, A, ca[0], ca[1], ca[2], fp, a, -1.0f, p2pa);
            }
        } else if (ta == kCellSolid || tb == kCellSolid) {
            uint32_t id = 0u;
            int fc[3] = {c.x, c.y, c.z};
            bool body_above = true; // the body is cell B, on the face's +axis side
            if (ta == kCellFluid && collider_of(B) != 0u) {
                id = collider_of(B);
                fc[a] -= 1;
            } else if (tb == kCellFluid && collider_of(A) != 0u) {
                id = collider_of(A);
                body_above = false;
            }
            if (id == 0u || !(col[id - 1u].flags & 1u)) continue;
            int32_t *acc = coupling + (id - 1u) * 16u;
            const uint32_t fi = cell_index(g, c, fc[0], fc[1], fc[2]);
            if (binary) pressure_force(g, col[id - 1u], acc, fp, a, body_above ? 1.0f : -1.0f, (float)pres[fi], p2pa);
            atomicAdd(&acc[6], 1);
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                int up[3] = {fc[0], fc[1], fc[2]};
                up[b] += 1;
                const float v = 0.5f * (vel[(uint32_t)b * g.pool_cells + fi] + vel[face_index(g, c, b, up[0], up[1], up[2])]);
                add_fixed(&acc[8 + b], v);
            }
        }
    }
}

// Slosh.usf:1094-1111 (FluidPressureGradient): per metre, central where both neighbours are fluid, else one-sided.
__device__ void fluid_pressure_gradient(const Geo &g, const CellPos &c, const uint32_t *ctype, const double *pres,
                                        float pc, float gr[3])
{
    for (int a = 0; a < 3; ++a) {
        int p[3] = {c.x, c.y, c.z}, m[3] = {c.x, c.y, c.z};
        p[a] += 1;
        m[a] -= 1;
        const bool bp = (type_at(g, c, ctype, p[0], p[1], p[2]) & kCellTypeMask) == kCellFluid;
        const bool bm = (type_at(g, c, ctype, m[0], m[1], m[2]) & kCellTypeMask) == kCellFluid;
        const float pp = bp ? (float)pres[cell_index(g, c, p[0], p[1], p[2])] : 0.0f;
        const float pm = bm ? (float)pres[cell_index(g, c, m[0], m[1], m[2])] : 0.0f;
        gr[a] = bp && bm ? (pp - pm) * 0.5f : (bp ? pp - pc : (bm ? pc - pm : 0.0f));
    }
    for (int a = 0; a < 3; ++a) gr[a] = gr[a] * g.inv_h;
}

// Slosh.usf:1115-1136 (BodySurfacePressure): from each fluid neighbour, extrapolated linearly to the shared face.
__device__ bool body_surface_pressure(const Geo &g, const CellPos &c, const uint32_t *ctype, const double *pres,
                                      float &p_out)
{
    float sum = 0.0f, cnt = 0.0f;
    for (int nb = 0; nb < 6; ++nb) {
        int d[3];
        nb_offset(nb, d);
        const int nx = c.x + d[0], ny = c.y + d[1], nz = c.z + d[2];
        if ((type_at(g, c, ctype, nx, ny, nz) & kCellTypeMask) != kCellFluid) continue;
        const float pn = (float)pres[cell_index(g, c, nx, ny, nz)];
        const int fx = nx + d[0], fy = ny + d[1], fz = nz + d[2];
        const float slope = (type_at(g, c, ctype, fx, fy, fz) & kCellTypeMask) == kCellFluid
                                ? (float)pres[cell_index(g, c, fx, fy, fz)] - pn
                                : 0.0f;
        sum += pn - 0.5f * slope;
        cnt += 1.0f;
    }
    p_out = cnt > 0.0f ? sum / cnt : 0.0f;
    return cnt > 0.0f;
}

// Slosh.usf:1138-1216 (CouplingForceCS): the smoothed pressure force on two-way colliders, -p h^2 sum_faces w_f n_f with
// w_f = saturate(0.5 + phi(x_f) / h), from fluid cells within a cell of the surface (their pressure moved to the
// surface along the gradient) and the body's own surface cells; faces a body cell shares with another solid are
// left to the contact rule.
__global__ __launch_bounds__(512) void k_coupling_force(Geo g, const uint32_t *rows, const uint32_t *ctype,
                                                        const double *pres, const slosh_collider *col, uint32_t ncol,
                                                        int32_t *coupling, float p2pa)
{
    __shared__ uint32_t s_row[28];
    CellPos c;
    uint32_t e;
    if (!cell_of_block(g, rows, s_row, c, e)) return;
    const uint32_t packed = ctype[c.idx], type = packed & kCellTypeMask;
    if (type == kCellAir) return;
    const float X[3] = {((float)c.x + 0.5f) * g.h, ((float)c.y + 0.5f) * g.h, ((float)c.z + 0.5f) * g.h};
    float ph = 0.0f, gr[3] = {0.0f, 0.0f, 0.0f};
    bool have = false;
    if (type == kCellFluid) {
        ph = (float)pres[c.idx];
        fluid_pressure_gradient(g, c, ctype, pres, ph, gr);
        have = true;
    } else if (collider_of(packed) != 0u) {
        have = body_surface_pressure(g, c, ctype, pres, ph);
    }
    if (!have) return;
    for (uint32_t k = 0; k < ncol; ++k) {
        const slosh_collider cc = col[k];
        if (!(cc.flags & 1u) || fabsf(collider_sdf(cc, X)) > g.h) continue;
        if (type == kCellSolid && collider_of(packed) != k + 1u) continue;
        float sp = ph;
        if (type == kCellFluid) {
            const float phi = collider_sdf(cc, X);
            float n[3];
            collider_normal(cc, X, 0.25f * g.h, n);
            sp += gr[0] * (-phi * n[0]) + gr[1] * (-phi * n[1]) + gr[2] * (-phi * n[2]);
        }
        const float scale = -sp * p2pa * g.h * g.h;
        float F[3] = {0.0f, 0.0f, 0.0f}, T[3] = {0.0f, 0.0f, 0.0f};
        for (int nb = 0; nb < 6; ++nb) {
            float nrm[3] = {0.0f, 0.0f, 0.0f};
            nrm[nb >> 1] = (nb & 1) ? 1.0f : -1.0f;
            float fpos[3];
            for (int a = 0; a < 3; ++a) fpos[a] = X[a] + 0.5f * g.h * nrm[a];
            if (type == kCellSolid) {
                int d[3];
                nb_offset(nb, d);
                const uint32_t nbp = type_at(g, c, ctype, c.x + d[0], c.y + d[1], c.z + d[2]);
                if ((nbp & kCellTypeMask) == kCellSolid && collider_of(nbp) != k + 1u) continue;
            }
            const float w = clampf(0.5f + collider_sdf(cc, fpos) * g.inv_h, 0.0f, 1.0f);
            float df[3], r[3], t[3];
            for (int a = 0; a < 3; ++a) {
                df[a] = scale * w * nrm[a];
                r[a] = fpos[a] - cc.center[a];
            }
            cross3(r, df, t);
            for (int a = 0; a < 3; ++a) {
                F[a] += df[a];
                T[a] += t[a];
            }
        }
        int32_t *acc = coupling + k * 16u;
        for (int a = 0; a < 3; ++a) add_fixed(&acc[a], F[a]);
        for (int a = 0; a < 3; ++a) add_fixed(&acc[3 + a], T[a]);
    }
}

// Adds each lane's values v[k] to coupling[(id - 1) * 16 + slot[k]] for the lanes with `on`: one atomic per value
// and collider present in the wavefront instead of one per lane (integer sums, whose grouping changes nothing; the
// additions wrap like the atomics). Every lane of the wavefront must call it.
template <int N>
__device__ __forceinline__ void wave_coupling_add(int32_t *coupling, uint32_t id, bool on, const int (&slot)[N],
                                                  const int32_t (&v)[N])
{
    uint64_t pending = __ballot(on);
    const int lane = (int)(threadIdx.x & 63u);
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t lid = (uint32_t)__shfl((int)id, leader, 64);
        const bool mine = on && id == lid;
        for (int k = 0; k < N; ++k) {
            uint32_t sum = mine ? (uint32_t)v[k] : 0u;
            for (int o = 32; o > 0; o >>= 1) sum += (uint32_t)__shfl_xor((int)sum, o, 64);
            if (lane == leader) atomicAdd(&coupling[(lid - 1u) * 16u + (uint32_t)slot[k]], (int32_t)sum);
        }
        pending &= ~(uint64_t)__ballot(mine);
    }
}

// Slosh.usf:1316-1370 (SubmergedCS): per collider the count of its solid cells and of the submerged ones, with their
// offsets from the collider's centre. A solid cell counts as submerged when a walk of up to 12 cells along +x, -x,
// +y, -y out of the collider first meets dense fluid with no air in the three cells below (-z, the reference's down
// axis).
__global__ __launch_bounds__(512) void k_submerged(Geo g, const uint32_t *rows, const uint32_t *ctype,
                                                   const slosh_collider *col, int32_t *coupling)
{
    __shared__ uint32_t s_row[28];
    CellPos c;
    uint32_t e;
    const bool inside = cell_of_block(g, rows, s_row, c, e);
    const uint32_t packed = inside ? ctype[c.idx] : 0u, id = collider_of(packed);
    const bool body = inside && (packed & kCellTypeMask) == kCellSolid && id != 0u;
    // directions +x, -x, +y, -y
    bool sub = false;
    if (body) {
        for (int d = 0; d < 4; ++d) {
            int px = c.x, py = c.y;
            for (int s = 0; s < 12; ++s) {
                px += d == 0 ? 1 : (d == 1 ? -1 : 0);
                py += d == 2 ? 1 : (d == 3 ? -1 : 0);
                const uint32_t t = type_global(g, ctype, px, py, c.z);
                if ((t & kCellTypeMask) != kCellSolid || collider_of(t) != id) {
                    bool liquid = (t & kCellTypeMask) == kCellFluid && (t & kCellDense) != 0u;
                    for (int below = 1; below <= 3; ++below)
                        liquid = liquid && (type_global(g, ctype, px, py, c.z - below) & kCellTypeMask) != kCellAir;
                    sub = sub || liquid;
                    break;
                }
            }
        }
    }
    int32_t fx[3] = {0, 0, 0};
    if (sub) {
        const slosh_collider cc = col[id - 1u];
        const float off[3] = {((float)c.x + 0.5f) * g.h - cc.center[0], ((float)c.y + 0.5f) * g.h - cc.center[1],
                              ((float)c.z + 0.5f) * g.h - cc.center[2]};
        for (int a = 0; a < 3; ++a) fx[a] = (int32_t)rintf(off[a] * kCouplingScale);
    }
    wave_coupling_add<1>(coupling, id, body, {11}, {1});
    wave_coupling_add<4>(coupling, id, sub, {7, 12, 13, 14}, {1, fx[0], fx[1], fx[2]});
}

// ---- two-sided density projection (Slosh.usf:743-854) --------------------------------------------------------
// Slosh.usf:753-773 (CellDensityInterior): the P2G weight of the cell's faces that no solid shares, per rest
// particle count of those faces; 1 when a solid shares every face.
__device__ __forceinline__ float cell_density_interior(const Geo &g, const CellPos &c, const float *fw,
                                                       const uint32_t *ctype)
{
    float ws = 0.0f, faces = 0.0f;
    for (int a = 0; a < 3; ++a) {
        int m[3] = {c.x, c.y, c.z}, p[3] = {c.x, c.y, c.z};
        m[a] -= 1;
        p[a] += 1;
        if ((type_at(g, c, ctype, m[0], m[1], m[2]) & kCellTypeMask) != kCellSolid) {
            ws += fw[(uint32_t)a * g.pool_cells + c.idx];
            faces += 1.0f;
        }
        if ((type_at(g, c, ctype, p[0], p[1], p[2]) & kCellTypeMask) != kCellSolid) {
            ws += fw[face_index(g, c, a, p[0], p[1], p[2])];
            faces += 1.0f;
        }
    }
    return faces > 0.0f ? ws / (faces * kTargetPpc) : 1.0f;
}

// Slosh.usf:775-805 (DensityProjectionRhsCS): the solver's "divergence" -relax * clamp(rho - 1, -0.5, 0.5) of fluid
// cells, where a cell touching air may only expand (rho - 1 >= 0); q starts from 0 (kept when keep_q: the
// pressure solver is off and q was written by the caller).
__global__ __launch_bounds__(512) void k_density_rhs(Geo g, const uint32_t *rows, const float *fw, const uint32_t *ctype,
                                                     float *rhs, double *q, float relax, int keep_q)
{
    __shared__ uint32_t s_row[28];
    CellPos c;
    uint32_t e;
    const bool in = cell_of_block(g, rows, s_row, c, e);
    if (!keep_q) q[c.idx] = 0.0;
    if (!in || (ctype[c.idx] & kCellTypeMask) != kCellFluid) {
        rhs[c.idx] = 0.0f;
        return;
    }
    float dev = cell_density_interior(g, c, fw, ctype) - 1.0f;
    bool air = false;
    for (int nb = 0; nb < 6; ++nb) {
        int d[3];
        nb_offset(nb, d);
        air = air || (type_at(g, c, ctype, c.x + d[0], c.y + d[1], c.z + d[2]) & kCellTypeMask) == kCellAir;
    }
    if (air) dev = fmaxf(dev, 0.0f);
    rhs[c.idx] = -relax * clampf(dev, -0.5f, 0.5f);
}

// Slosh.usf:808-828 (DensityFaceDisplacement, DensityDisplacementFace): -(q_b - q_a) on each face, q of a non-fluid
// cell 0, and 0 on faces that touch a solid.
__global__ __launch_bounds__(512) void k_density_disp(Geo g, const uint32_t *rows, const uint32_t *ctype,
                                                      const double *q, float *disp)
{
    __shared__ uint32_t s_row[28];
    CellPos c;
    uint32_t e;
    const bool in = cell_of_block(g, rows, s_row, c, e);
    const uint32_t tb = ctype[c.idx] & kCellTypeMask;
    for (int a = 0; a < 3; ++a) {
        const uint32_t f = (uint32_t)a * g.pool_cells + c.idx;
        if (!in) {
            disp[f] = 0.0f;
            continue;
        }
        int ca[3] = {c.x, c.y, c.z};
        ca[a] -= 1;
        const uint32_t ta = type_at(g, c, ctype, ca[0], ca[1], ca[2]) & kCellTypeMask;
        if (ta == kCellSolid || tb == kCellSolid) {
            disp[f] = 0.0f;
            continue;
        }
        const double qa = ta == kCellFluid ? q[cell_index(g, c, ca[0], ca[1], ca[2])] : 0.0;
        const double qb = tb == kCellFluid ? q[c.idx] : 0.0;
        disp[f] = (float)(-(qb - qa));
    }
}

// Slosh.usf:1222-1240 (ResidualCS): max and sum of |div|/h over fluid cells after projection, one partial per
// brick (reduced with the other step statistics by k_step_stats).
__global__ __launch_bounds__(512) void k_div_stats(Geo g, const uint32_t *rows, const float *vel, const uint32_t *ctype,
                                                   uint32_t *bp_max, uint32_t *bp_cnt, double *bp_sum)
{
    constexpr uint32_t kWaves = 512u / 64u;
    __shared__ double ssum[kWaves];
    __shared__ uint32_t smax[kWaves], scount[kWaves];
    __shared__ uint32_t s_row[28];
    CellPos c;
    uint32_t e;
    double rate_d = 0.0;
    uint32_t rate_bits = 0u; // rates are non-negative, so their bits order like the values
    bool fluid = false;
    if (cell_of_block(g, rows, s_row, c, e) && (ctype[c.idx] & kCellTypeMask) == kCellFluid) {
        float dv = 0.0f;
        for (int a = 0; a < 3; ++a) {
            int p[3] = {c.x, c.y, c.z};
            p[a] += 1;
            dv += vel[face_index(g, c, a, p[0], p[1], p[2])] - vel[(uint32_t)a * g.pool_cells + c.idx];
        }
        float rate = fabsf(dv) * g.inv_h;
        rate_bits = __float_as_uint(rate);
        rate_d = (double)rate;
        fluid = true;
    }
    // per wavefront: the sum by a fixed shuffle tree, the maximum and the count in any order; then the eight
    // wavefront values in wavefront order
    for (int s = 32; s > 0; s >>= 1) {
        rate_d += __shfl_down(rate_d, s);
        rate_bits = max(rate_bits, (uint32_t)__shfl_down((int)rate_bits, s));
    }
    const uint32_t n = (uint32_t)__popcll(__ballot(fluid));
    if ((threadIdx.x & 63u) == 0u) {
        ssum[threadIdx.x >> 6] = rate_d;
        smax[threadIdx.x >> 6] = rate_bits;
        scount[threadIdx.x >> 6] = n;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sum = 0.0;
        uint32_t mx = 0u, cnt = 0u;
        for (uint32_t w = 0; w < kWaves; ++w) {
            sum += ssum[w];
            mx = max(mx, smax[w]);
            cnt += scount[w];
        }
        bp_sum[e] = sum;
        bp_max[e] = mx;
        bp_cnt[e] = cnt;
    }
}

// End of step: one workgroup reduces the per-workgroup statistics of P2G, G2P and the divergence pass (counts and
// maxima in any order, the FP64 divergence sum in a fixed order: per thread in increasing index, then a tree).
// Each thread keeps eight loads in flight per loop trip.
constexpr uint32_t kStatsThreads = 1024;
__global__ __launch_bounds__(kStatsThreads) void k_step_stats(const uint32_t *bp_fluid, uint32_t n_p2g,
                                                              const uint32_t *bp_div_max, const uint32_t *bp_div_cnt,
                                                              const double *bp_div_sum, uint32_t n_div,
                                                              const uint32_t *bp_speed, const uint32_t *bp_clamp,
                                                              const uint32_t *bp_drain, uint32_t n_g2p, StepStats *stats)
{
    constexpr uint32_t T = kStatsThreads, U = 8;
    __shared__ uint32_t su[6][T];
    __shared__ double sd[T];
    const uint32_t t = threadIdx.x;
    uint32_t fluid = 0, dmax = 0, dcnt = 0, smax = 0, clamp = 0, drain = 0;
    double dsum = 0.0;
    for (uint32_t b0 = t; b0 < n_p2g; b0 += T * U) {
        uint32_t v[U];
        for (uint32_t u = 0; u < U; ++u) v[u] = b0 + T * u < n_p2g ? bp_fluid[b0 + T * u] : 0u;
        for (uint32_t u = 0; u < U; ++u) fluid += v[u];
    }
    for (uint32_t b0 = t; b0 < n_div; b0 += T * U) {
        uint32_t m[U], k[U];
        double d[U];
        for (uint32_t u = 0; u < U; ++u) {
            const uint32_t b = b0 + T * u;
            const bool in = b < n_div;
            m[u] = in ? bp_div_max[b] : 0u;
            k[u] = in ? bp_div_cnt[b] : 0u;
            d[u] = in ? bp_div_sum[b] : 0.0;
        }
        for (uint32_t u = 0; u < U; ++u) {
            dmax = max(dmax, m[u]);
            dcnt += k[u];
            dsum += d[u];
        }
    }
    for (uint32_t b0 = t; b0 < n_g2p; b0 += T * U) {
        uint32_t sp[U], cl[U], dr[U];
        for (uint32_t u = 0; u < U; ++u) {
            const uint32_t b = b0 + T * u;
            const bool in = b < n_g2p;
            sp[u] = in ? bp_speed[b] : 0u;
            cl[u] = in ? bp_clamp[b] : 0u;
            dr[u] = in ? bp_drain[b] : 0u;
        }
        for (uint32_t u = 0; u < U; ++u) {
            smax = max(smax, sp[u]);
            clamp += cl[u];
            drain += dr[u];
        }
    }
    su[0][t] = fluid;
    su[1][t] = dmax;
    su[2][t] = dcnt;
    su[3][t] = smax;
    su[4][t] = clamp;
    su[5][t] = drain;
    sd[t] = dsum;
    __syncthreads();
    for (uint32_t s = T / 2; s > 0; s >>= 1) {
        if (t < s) {
            su[0][t] += su[0][t + s];
            su[1][t] = max(su[1][t], su[1][t + s]);
            su[2][t] += su[2][t + s];
            su[3][t] = max(su[3][t], su[3][t + s]);
            su[4][t] += su[4][t + s];
            su[5][t] += su[5][t + s];
            sd[t] += sd[t + s];
        }
        __syncthreads();
    }
    if (t == 0) {
        stats->fluid_cells = su[0][0];
        stats->max_div_bits = su[1][0];
        stats->div_count = su[2][0];
        stats->max_speed_bits = su[3][0];
        stats->clamped = su[4][0];
        stats->drained = su[5][0];
        stats->div_sum = sd[0];
    }
}

// Slosh.usf:1263-1310 (ExtrapolateFace) for layer `layer`: src -> dst, marking filled faces with bit 4+layer.
__global__ __launch_bounds__(512) void k_extrapolate(Geo g, const uint32_t *rows, const float *src, float *dst,
                                                     uint8_t *fflags, uint32_t layer)
{
    __shared__ uint32_t s_row[28];
    CellPos c;
    uint32_t e;
    if (!cell_of_block(g, rows, s_row, c, e)) return;
    uint32_t known = kFaceValid;
    for (uint32_t k = 0; k < layer; ++k) known |= 1u << (kFaceExtrapShift + k);
    for (int a = 0; a < 3; ++a) {
        uint32_t f = (uint32_t)a * g.pool_cells + c.idx;
        uint32_t flags = fflags[f];
        float center = src[f];
        if (flags & known) {
            dst[f] = center;
            continue;
        }
        float sum = 0.0f, cnt = 0.0f;
        for (int nb = 0; nb < 6; ++nb) {
            int d[3];
            nb_offset(nb, d);
            int nx = c.x + d[0], ny = c.y + d[1], nz = c.z + d[2];
            if (!in_grid(g, nx, ny, nz)) continue;
            uint32_t ni = face_index(g, c, a, nx, ny, nz);
            uint32_t nf = fflags[ni];
            if ((nf & known) != 0u && (nf & kFaceSolid) == 0u) {
                sum += src[ni];
                cnt += 1.0f;
            }
        }
        if (cnt > 0.0f) {
            dst[f] = sum / cnt;
            fflags[f] = (uint8_t)(flags | (1u << (kFaceExtrapShift + layer)));
        } else {
            dst[f] = 0.0f;
        }
    }
}

} // namespace

void launch_fill_u32(hipStream_t st, uint32_t *p, uint32_t v, size_t n)
{
    if (!n) return;
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(n, 256)), dim3(256), 0, st, p, v, n);
    SLOSH_HIP(hipGetLastError());
}
void launch_zero3_u32(hipStream_t st, uint32_t *a, uint32_t *b, uint32_t *c, size_t n)
{
    if (!n) return;
    hipLaunchKernelGGL(k_zero3_u32, dim3(blocks_for(n, 256)), dim3(256), 0, st, a, b, c, n);
    SLOSH_HIP(hipGetLastError());
}
void launch_fill_u8(hipStream_t st, uint8_t *p, uint8_t v, size_t n)
{
    if (!n) return;
    hipLaunchKernelGGL(k_fill_u8, dim3(blocks_for(n, 256)), dim3(256), 0, st, p, v, n);
    SLOSH_HIP(hipGetLastError());
}
void launch_fill_f64(hipStream_t st, double *p, double v, size_t n)
{
    if (!n) return;
    hipLaunchKernelGGL(k_fill_f64, dim3(blocks_for(n, 256)), dim3(256), 0, st, p, v, n);
    SLOSH_HIP(hipGetLastError());
}

// out[0..n] = exclusive prefix sums of in[0..n), out[n] = total. Supports n <= 256 * 1024 * 1024.
void scan_exclusive(slosh_ctx *c, const uint32_t *in, uint32_t *out, uint32_t n)
{
    if (n <= kScanOneMax) {
        launch_timed(c, K_SCAN, k_scan_one, dim3(1), dim3(kScanOneThreads), 0, in, n, out, in, n, out);
        SLOSH_HIP(hipGetLastError());
        return;
    }
    const uint32_t nb = blocks_for(n, kScanBlock);
    {
        launch_timed(c, K_SCAN, k_scan_reduce, dim3(nb), dim3(256), 0, in, n, c->scan_blocks);
    }
    {
        launch_timed(c, K_SCAN, k_scan_blocks, dim3(1), dim3(256), 0, c->scan_blocks, nb);
    }
    {
        launch_timed(c, K_SCAN, k_scan_down, dim3(nb), dim3(256), 0, in, n, c->scan_blocks, out);
    }
    SLOSH_HIP(hipGetLastError());
}

void scan_exclusive2(slosh_ctx *c, const uint32_t *in0, uint32_t *out0, uint32_t n0, const uint32_t *in1, uint32_t *out1,
                     uint32_t n1)
{
    if (n0 <= kScanOneMax && n1 <= kScanOneMax) {
        launch_timed(c, K_SCAN, k_scan_one, dim3(2), dim3(kScanOneThreads), 0, in0, n0, out0, in1, n1, out1);
        SLOSH_HIP(hipGetLastError());
        return;
    }
    scan_exclusive(c, in0, out0, n0);
    scan_exclusive(c, in1, out1, n1);
}

void run_pool(slosh_ctx *c)
{
    const Geo &g = c->geo;
    uint32_t *table = c->table;
    {
        launch_timed(c, K_POOL, k_pool_want, dim3(blocks_for(g.nbricks, 256)), dim3(256), 0, g, c->touch, table,
                     c->slot_brick, c->need);
    }
    {
        launch_timed(c, K_POOL, k_pool_freeflags, dim3(blocks_for(g.npool, 256)), dim3(256), 0, c->slot_brick, g.npool,
                     c->freeflag);
    }
    SLOSH_HIP(hipGetLastError());
    scan_exclusive2(c, c->need, c->need_rank, g.nbricks, c->freeflag, c->free_rank, g.npool);
    {
        launch_timed(c, K_POOL, k_pool_free_list, dim3(blocks_for(g.npool, 256)), dim3(256), 0, c->freeflag,
                     c->free_rank, g.npool, c->free_list);
    }
    {
        launch_timed(c, K_POOL, k_pool_assign, dim3(blocks_for(g.nbricks, 256)), dim3(256), 0, g, c->need,
                     c->need_rank, c->free_list, c->free_rank, g.npool, table, c->slot_brick,
                     c->new_slot, c->touch, c->alloc_flag,
                     c->stats);
    }
    SLOSH_HIP(hipGetLastError());
    scan_exclusive(c, c->alloc_flag, c->alloc_rank, g.nbricks);
    {
        launch_timed(c, K_POOL, k_pool_active, dim3(blocks_for(g.nbricks, 256)), dim3(256), 0, g, c->alloc_flag,
                     c->alloc_rank, c->active, c->stats);
    }
    SLOSH_HIP(hipGetLastError());
}

// The neighbourhood rows of the allocated bricks (once per step, after the pool; every per-brick pass uses them).
const uint32_t *device_nactive(const slosh_ctx *c)
{
    return reinterpret_cast<const uint32_t *>(reinterpret_cast<const char *>(c->stats) + offsetof(StepStats, nactive));
}

void build_brick_rows(slosh_ctx *c, uint32_t first, uint32_t count)
{
    if (!count) return;
    launch_timed(c, K_POOL, k_brick_rows, dim3(blocks_for((size_t)count * kBinfoWords, 256)), dim3(256), 0, c->geo,
                 first, device_nactive(c), c->binfo);
    SLOSH_HIP(hipGetLastError());
}

void launch_normalize(slosh_ctx *c, float dt, const float g[3])
{
    if (!c->nactive) return;
    c->vcur = 0;
    {
        launch_timed(c, K_NORMALIZE, k_normalize, dim3(c->nactive), dim3(512), 0, c->geo, c->binfo, c->tiles,
                     c->tile_valid, c->ctype, c->vel[0], c->fweight, c->fflags, g[0] * dt, g[1] * dt, g[2] * dt,
                     c->d_col);
    }
    SLOSH_HIP(hipGetLastError());
}

void launch_divergence(slosh_ctx *c, float dt)
{
    if (!c->nactive) return;
    {
        launch_timed(c, K_DIVERGENCE, k_divergence, dim3(c->nactive), dim3(512), 0, c->geo, c->binfo,
                     c->vel[c->vcur], c->fweight, c->ctype, c->div, c->diag, c->fmask, c->s.ghost_fluid,
                     (c->s.flags & SLOSH_FLAG_DENSITY_PROJECTION) ? 0.0f : c->s.density_correction,
                     c->s.density_deadband, c->geo.h, dt);
    }
    SLOSH_HIP(hipGetLastError());
}

void launch_project(slosh_ctx *c)
{
    if (!c->nactive) return;
    {
        launch_timed(c, K_PROJECT, k_project, dim3(c->nactive), dim3(512), 0, c->geo, c->binfo, c->vel[c->vcur],
                     c->fweight, c->ctype, c->pres, c->fflags, c->s.ghost_fluid);
    }
    {
        launch_timed(c, K_DIV_STATS, k_div_stats, dim3(c->nactive), dim3(512), 0, c->geo, c->binfo, c->vel[c->vcur],
                     c->ctype, c->bp_div_max, c->bp_div_cnt, c->bp_div_sum);
    }
    c->n_div_blocks = c->nactive;
    SLOSH_HIP(hipGetLastError());
}

void launch_coupling(slosh_ctx *c, float dt)
{
    if (!c->nactive || !c->ncol) return;
    // the face and force terms accumulate only for two-way bodies (every other collider returns at once in them)
    bool two_way = false;
    for (uint32_t k = 0; k < c->ncol; ++k) two_way = two_way || (c->h_col[k].flags & 1u) != 0u;
    if (!two_way) return;
    // SloshProxy.cpp:562: PressureToPascal = density h / dt
    const float rho = c->s.fluid_density > 0.0f ? c->s.fluid_density : 1000.0f;
    const float p2pa = rho * c->geo.h / dt;
    const int binary = (c->s.flags & SLOSH_FLAG_BINARY_COUPLING) != 0;
    launch_timed(c, K_COUPLING, k_coupling_faces, dim3(c->nactive), dim3(512), 0, c->geo, c->binfo, c->ctype, c->pres,
                 c->vel[c->vcur], c->d_col, c->coupling, p2pa, binary);
    if (!binary)
        launch_timed(c, K_COUPLING, k_coupling_force, dim3(c->nactive), dim3(512), 0, c->geo, c->binfo, c->ctype,
                     c->pres, c->d_col, c->ncol, c->coupling, p2pa);
    SLOSH_HIP(hipGetLastError());
}

void launch_submerged(slosh_ctx *c)
{
    if (!c->nactive || !c->ncol) return;
    launch_timed(c, K_COUPLING, k_submerged, dim3(c->nactive), dim3(512), 0, c->geo, c->binfo, c->ctype, c->d_col,
                 c->coupling);
    SLOSH_HIP(hipGetLastError());
}

void run_viscosity(slosh_ctx *c, float alpha, int iterations)
{
    if (!c->nactive || iterations <= 0) return;
    // vel[0] keeps u*; the iterates alternate between vel[1] and vel[2]
    SLOSH_HIP(hipMemcpyAsync(c->vel[1], c->vel[0], (size_t)3 * c->geo.pool_cells * sizeof(float), hipMemcpyDeviceToDevice,
                             c->stream));
    int src = 1;
    for (int it = 0; it < iterations; ++it) {
        const int dst = src == 1 ? 2 : 1;
        launch_timed(c, K_VISCOSITY, k_viscosity, dim3(c->nactive), dim3(512), 0, c->geo, c->binfo, c->vel[0],
                     c->vel[src], c->vel[dst], c->fflags, alpha);
        src = dst;
    }
    SLOSH_HIP(hipGetLastError());
    c->vcur = src;
}

void launch_density_rhs(slosh_ctx *c, float relax, int keep_q)
{
    if (!c->nactive) return;
    launch_timed(c, K_DENSITY, k_density_rhs, dim3(c->nactive), dim3(512), 0, c->geo, c->binfo, c->fweight, c->ctype,
                 c->dens_rhs, c->dens_q, relax, keep_q);
    SLOSH_HIP(hipGetLastError());
}

void launch_density_disp(slosh_ctx *c)
{
    if (!c->nactive) return;
    launch_timed(c, K_DENSITY, k_density_disp, dim3(c->nactive), dim3(512), 0, c->geo, c->binfo, c->ctype, c->dens_q,
                 c->dens_disp);
    SLOSH_HIP(hipGetLastError());
}

void launch_step_stats(slosh_ctx *c)
{
    SLOSH_TIMED(c, K_STEP_STATS);
    hipLaunchKernelGGL(k_step_stats, dim3(1), dim3(kStatsThreads), 0, c->stream, c->bp_fluid, c->n_p2g_blocks, c->bp_div_max,
                       c->bp_div_cnt, c->bp_div_sum, c->n_div_blocks, c->bp_speed, c->bp_clamp, c->bp_drain, c->n_g2p_blocks,
                       c->stats);
    SLOSH_HIP(hipGetLastError());
}

void launch_extrapolate(slosh_ctx *c, uint32_t layer)
{
    if (!c->nactive) return;
    int dst = (c->vcur + 1) % 3;
    {
        launch_timed(c, K_EXTRAPOLATE, k_extrapolate, dim3(c->nactive), dim3(512), 0, c->geo, c->binfo,
                     c->vel[c->vcur], c->vel[dst], c->fflags, layer);
    }
    SLOSH_HIP(hipGetLastError());
    c->vcur = dst;
}

} // namespace slosh
