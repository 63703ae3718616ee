// particles.hip -- particle kernels of the solver step on gfx950: emission, stable brick sort, P2G, G2P.
//
// Launch shapes are wave64-native: 256-thread workgroups (4 waves) for particle passes and one workgroup per
// allocated 8^3 brick for the P2G gather, whose LDS tile holds every node the brick's particles touch.
#include <utility>
#include "slosh_ctx.h"

namespace slosh {

namespace {

constexpr uint32_t kTile = 256; // particles per sort tile (one workgroup)

__device__ __forceinline__ uint32_t key_of(const Geo &g, float4 p, uint32_t dead_key)
{
    if (!(p.w >= 0.0f)) return dead_key;
    int c[3];
    pos_cell(g, p.x, p.y, p.z, c);
    return brick_of_cell(g, c[0], c[1], c[2]);
}

__device__ __forceinline__ void touch_brick(const Geo &g, uint8_t *touch, float x, float y, float z)
{
    int c[3];
    pos_cell(g, x, y, z, c);
    touch[brick_of_cell(g, c[0], c[1], c[2])] = 1u;
}

// Slosh.usf:280-358 (EmitParticlesCS): mode 0 FillCells (particle I in cell I/8, sub-cell I%8, jittered), 1 a
// uniform point in the sphere of radius extent[0] (cube-root radius, uniform direction), 2 a uniform point in the
// rotated box; optional cone spread and speed jitter of the velocity; position clamped into [1.01h, (N-1.01)h].
__global__ void k_emit(Geo g, Parts p, uint32_t base, uint32_t count, slosh_emit_request r, uint8_t *touch)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint32_t h0 = pcg_hash(i * 3u + r.seed * 7919u);
    uint32_t h1 = pcg_hash(h0);
    uint32_t h2 = pcg_hash(h1);
    float rnd[3] = {hash_to_unit(h0), hash_to_unit(h1), hash_to_unit(h2)};
    float lo = g.h * 1.01f;
    float pos[3];
    if (r.shape == SLOSH_EMIT_FILL_CELLS) {
        uint32_t dx = (uint32_t)r.cell_dims[0], dy = (uint32_t)r.cell_dims[1];
        uint32_t cl = i / 8u, sub = i % 8u;
        int local[3] = {(int)(cl % dx), (int)((cl / dx) % dy), (int)(cl / (dx * dy))};
        float so[3] = {(float)(sub & 1u), (float)((sub >> 1u) & 1u), (float)((sub >> 2u) & 1u)};
        for (int a = 0; a < 3; ++a) {
            float in_cell = ((so[a] + 0.5f) + 0.8f * (rnd[a] - 0.5f)) * 0.5f;
            pos[a] = ((float)(r.cell_min[a] + local[a]) + in_cell) * g.h;
        }
    } else if (r.shape == SLOSH_EMIT_SPHERE) {
        uint32_t h3 = pcg_hash(h2);
        float cos_t = 2.0f * rnd[0] - 1.0f;
        float sin_t = sqrtf(clampf(1.0f - cos_t * cos_t, 0.0f, 1.0f));
        float phi = 6.28318530718f * rnd[1];
        float rad = r.extent[0] * powf(fmaxf(hash_to_unit(h3), 1e-6f), 1.0f / 3.0f);
        float dir[3] = {sin_t * cosf(phi), sin_t * sinf(phi), cos_t};
        for (int a = 0; a < 3; ++a) pos[a] = r.center[a] + rad * dir[a];
    } else {
        float l[3], w[3];
        for (int a = 0; a < 3; ++a) l[a] = (2.0f * rnd[a] - 1.0f) * r.extent[a];
        quat_rotate(r.rotation, l, w);
        for (int a = 0; a < 3; ++a) pos[a] = r.center[a] + w[a];
    }
    for (int a = 0; a < 3; ++a) pos[a] = clampf(pos[a], lo, ((float)g.n[a] - 1.01f) * g.h);
    float vel[3] = {r.velocity[0], r.velocity[1], r.velocity[2]};
    const float speed = len3(vel[0], vel[1], vel[2]);
    if (speed > 1e-6f && (r.velocity_spread > 0.0f || r.speed_jitter > 0.0f)) {
        uint32_t h4 = pcg_hash(h2 ^ 0x9E3779B9u), h5 = pcg_hash(h4), h6 = pcg_hash(h5);
        float dir[3] = {vel[0] / speed, vel[1] / speed, vel[2] / speed};
        const float up[3] = {fabsf(dir[2]) < 0.9f ? 0.0f : 1.0f, 0.0f, fabsf(dir[2]) < 0.9f ? 1.0f : 0.0f};
        float t0[3], t1[3];
        cross3(dir, up, t0);
        const float tl = len3(t0[0], t0[1], t0[2]);
        for (int a = 0; a < 3; ++a) t0[a] = t0[a] / tl;
        cross3(dir, t0, t1);
        const float cos_a = 1.0f + hash_to_unit(h4) * (cosf(r.velocity_spread) - 1.0f);
        const float sin_a = sqrtf(clampf(1.0f - cos_a * cos_a, 0.0f, 1.0f));
        const float phi = 6.28318530718f * hash_to_unit(h5);
        const float cp = cosf(phi), sp = sinf(phi);
        const float scale = fmaxf(1.0f + r.speed_jitter * (2.0f * hash_to_unit(h6) - 1.0f), 0.0f);
        for (int a = 0; a < 3; ++a) vel[a] = (cos_a * dir[a] + sin_a * (cp * t0[a] + sp * t1[a])) * speed * scale;
    }
    uint32_t j = base + i;
    p.pa[j] = make_float4(pos[0], pos[1], pos[2], 0.0f);
    for (int a = 0; a < 3; ++a) p.v[a][j] = vel[a];
    for (int k = 0; k < 9; ++k) p.c[k][j] = 0.0f;
    touch_brick(g, touch, pos[0], pos[1], pos[2]);
}

// Resume from 64-byte records: positions are clamped into the particle band [1.01h, (N-1.01)h] of Slosh.usf:1459.
__global__ void k_unpack(Geo g, Parts p, const slosh_particle *src, uint32_t n, uint8_t *touch)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    slosh_particle s = src[i];
    float lo = g.h * 1.01f;
    float pos[3];
    for (int a = 0; a < 3; ++a) pos[a] = clampf(s.pos[a], lo, ((float)g.n[a] - 1.01f) * g.h);
    p.pa[i] = make_float4(pos[0], pos[1], pos[2], s.age);
    for (int a = 0; a < 3; ++a) p.v[a][i] = s.vel[a];
    for (int k = 0; k < 9; ++k) p.c[k][i] = s.c[k];
    if (s.age >= 0.0f) touch_brick(g, touch, pos[0], pos[1], pos[2]);
}

// SloshCommon.ush:71-88 (PackHalf2): clamp to +-65000 then round to nearest even half.
__device__ __forceinline__ uint32_t pack_half2(float a, float b)
{
    _Float16 ha = (_Float16)clampf(a, -65000.0f, 65000.0f);
    _Float16 hb = (_Float16)clampf(b, -65000.0f, 65000.0f);
    return (uint32_t)__builtin_bit_cast(uint16_t, ha) | ((uint32_t)__builtin_bit_cast(uint16_t, hb) << 16);
}

__global__ void k_pack(Parts p, uint32_t n, void *dst, int layout)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 pa = p.pa[i];
    if (layout == SLOSH_LAYOUT_F32_64) {
        slosh_particle o;
        o.pos[0] = pa.x; o.pos[1] = pa.y; o.pos[2] = pa.z; o.age = pa.w;
        for (int a = 0; a < 3; ++a) o.vel[a] = p.v[a][i];
        for (int k = 0; k < 9; ++k) o.c[k] = p.c[k][i];
        ((slosh_particle *)dst)[i] = o;
    } else {
        slosh_particle_ref48 o;
        o.pos[0] = pa.x; o.pos[1] = pa.y; o.pos[2] = pa.z; o.age = pa.w;
        for (int a = 0; a < 3; ++a) o.vel[a] = p.v[a][i];
        float cm[9];
        for (int k = 0; k < 9; ++k) cm[k] = p.c[k][i];
        o.c[0] = pack_half2(cm[0], cm[1]);
        o.c[1] = pack_half2(cm[2], cm[3]);
        o.c[2] = pack_half2(cm[4], cm[5]);
        o.c[3] = pack_half2(cm[6], cm[7]);
        o.c[4] = pack_half2(cm[8], 0.0f);
        ((slosh_particle_ref48 *)dst)[i] = o;
    }
}

// Block-wide bitonic sort of `size` (power of two) u32 in LDS, ascending.
__device__ void bitonic_lds(uint32_t *s, uint32_t size)
{
    for (uint32_t k = 2; k <= size; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t t = threadIdx.x; t < size; t += blockDim.x) {
                uint32_t ixj = t ^ j;
                if (ixj > t) {
                    uint32_t a = s[t], b = s[ixj];
                    bool up = (t & k) == 0u;
                    if ((a > b) == up) {
                        s[t] = b;
                        s[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Ascending bitonic sort of one value per thread over a 256-thread workgroup: stages with partner distance < 64
// exchange through wavefront shuffles, the three with distance 64 or 128 through LDS.
__device__ __forceinline__ uint32_t bitonic_tile(uint32_t x, uint32_t *s)
{
    const uint32_t t = threadIdx.x;
    for (uint32_t k = 2; k <= kTile; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            uint32_t y;
            if (j >= 64u) {
                s[t] = x;
                __syncthreads();
                y = s[t ^ j];
                __syncthreads();
            } else {
                y = (uint32_t)__shfl_xor((int)x, (int)j);
            }
            const bool lower = (t & j) == 0u, up = (t & k) == 0u;
            x = (lower == up) ? min(x, y) : max(x, y);
        }
    }
    return x;
}

// Sort pass 1. Per tile of 256 consecutive particles: key = 8^3 brick of the clamped cell (dead: last key), or the
// given key when `keys` is set (the keys of G2P's first pass), a
// stable in-tile rank per key from a sort of (key << 8 | lane), and the tile's runs of equal keys. The values are
// distinct, so their sorted order is unique; a tile whose keys already ascend (the common case: particles were
// sorted last step and few cross a brick face) skips the sort. Run starts are counted with wavefront ballots.
// Reference counterpart: SortCountCS (SloshSort.usf:65-76), whose atomic rank is not stable.
__global__ __launch_bounds__(256) void k_sort_tiles(Geo g, const float4 *pa, const uint32_t *keys, uint32_t np,
                                                    uint32_t dead_key, uint32_t *packed, uint32_t *tile_nruns,
                                                    uint32_t *run_key, uint32_t *run_len, uint32_t *key_count,
                                                    uint32_t *key_runs)
{
    __shared__ uint32_t s[kTile];
    __shared__ uint32_t start[kTile];
    __shared__ uint32_t wruns[kTile / 64];
    __shared__ uint32_t unsorted;
    const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
    const uint32_t i = blockIdx.x * kTile + t;
    uint32_t v = 0xFFFFFFFFu;
    if (i < np) v = ((keys ? keys[i] : key_of(g, pa[i], dead_key)) << 8) | t;
    s[t] = v;
    if (t == 0u) unsorted = 0u;
    __syncthreads();
    if (t > 0u && s[t - 1u] > v) unsorted = 1u;
    __syncthreads();
    uint32_t e = v;
    if (unsorted) {
        e = bitonic_tile(v, s);
        s[t] = e;
        __syncthreads();
    }
    const bool valid = e != 0xFFFFFFFFu;
    const uint32_t k = e >> 8;
    const bool flag = valid && (t == 0u || (s[t - 1u] >> 8) != k);
    // run index of every entry = run starts at or before it, minus one
    const uint64_t m = __ballot(flag);
    const uint32_t upto = (uint32_t)__popcll(m & (lane == 63u ? ~0ull : ((1ull << (lane + 1u)) - 1ull)));
    if (lane == 0u) wruns[w] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t base = 0u, nruns = 0u;
    for (uint32_t q = 0; q < kTile / 64; ++q) {
        base += q < w ? wruns[q] : 0u;
        nruns += wruns[q];
    }
    const uint32_t run = base + upto - 1u;
    if (flag) start[run] = t;
    __syncthreads();
    const uint32_t nvalid = np - blockIdx.x * kTile < kTile ? np - blockIdx.x * kTile : kTile;
    if (valid) packed[blockIdx.x * kTile + (e & 255u)] = (run << 8) | (t - start[run]);
    if (flag) {
        uint32_t next = run + 1u < nruns ? start[run + 1u] : nvalid;
        uint32_t len = next - t;
        uint32_t gid = blockIdx.x * kTile + run;
        run_key[gid] = k;
        run_len[gid] = len;
        atomicAdd(&key_count[k], len);
        atomicAdd(&key_runs[k], 1u);
    }
    if (t == 0u) tile_nruns[blockIdx.x] = nruns;
}

// Sort pass 2: bucket each run under its key (arbitrary order within the key, fixed by pass 3).
__global__ __launch_bounds__(256) void k_sort_runs_scatter(const uint32_t *tile_nruns, const uint32_t *run_key,
                                                           const uint32_t *run_off, uint32_t *run_cursor,
                                                           uint32_t *run_sorted)
{
    uint32_t r = threadIdx.x;
    if (r >= tile_nruns[blockIdx.x]) return;
    uint32_t gid = blockIdx.x * kTile + r;
    uint32_t k = run_key[gid];
    uint32_t pos = run_off[k] + atomicAdd(&run_cursor[k], 1u);
    run_sorted[pos] = gid;
}

constexpr uint32_t kRunLds = 2048;

// Sort pass 3, one workgroup per key: order the key's runs by tile (run ids grow with the tile) and give each run
// its base = key offset + particles of this key in earlier tiles. Together with the in-tile rank this is the
// stable order by (key, old index).
__global__ __launch_bounds__(256) void k_sort_key_runs(const uint32_t *run_off, const uint32_t *key_off,
                                                       const uint32_t *run_len, uint32_t *run_sorted, uint32_t *run_base)
{
    __shared__ uint32_t s[kRunLds];
    __shared__ uint32_t part[256];
    const uint32_t k = blockIdx.x;
    const uint32_t b0 = run_off[k], m = run_off[k + 1] - b0;
    if (m == 0u) return;
    const uint32_t kb = key_off[k];
    if (m == 1u) {
        if (threadIdx.x == 0) run_base[run_sorted[b0]] = kb;
        return;
    }
    if (m <= kRunLds) {
        uint32_t size = 2;
        while (size < m) size <<= 1;
        for (uint32_t t = threadIdx.x; t < size; t += blockDim.x) s[t] = t < m ? run_sorted[b0 + t] : 0xFFFFFFFFu;
        __syncthreads();
        bitonic_lds(s, size);
        // exclusive scan of run lengths in sorted order: contiguous chunk per thread, block scan of chunk sums
        const uint32_t per = (m + blockDim.x - 1) / blockDim.x;
        const uint32_t beg = threadIdx.x * per;
        uint32_t sum = 0;
        for (uint32_t j = beg; j < beg + per && j < m; ++j) sum += run_len[s[j]];
        part[threadIdx.x] = sum;
        __syncthreads();
        for (uint32_t off = 1; off < blockDim.x; off <<= 1) {
            uint32_t add = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
            __syncthreads();
            part[threadIdx.x] += add;
            __syncthreads();
        }
        uint32_t run_acc = part[threadIdx.x] - sum;
        for (uint32_t j = beg; j < beg + per && j < m; ++j) {
            uint32_t gid = s[j];
            run_base[gid] = kb + run_acc;
            run_acc += run_len[gid];
        }
        return;
    }
    // Rare fallback for keys split over more than kRunLds tiles: base = sum of lengths of runs from earlier tiles.
    for (uint32_t j = threadIdx.x; j < m; j += blockDim.x) {
        uint32_t gid = run_sorted[b0 + j];
        uint32_t acc = 0;
        for (uint32_t q = 0; q < m; ++q) {
            uint32_t o = run_sorted[b0 + q];
            if (o < gid) acc += run_len[o];
        }
        run_base[gid] = kb + acc;
    }
}

// Sort pass 4: move each live particle's 64-byte record to its stable position (SortScatterCS, SloshSort.usf:135).
__global__ __launch_bounds__(256) void k_sort_scatter(Parts src, Parts dst, uint32_t np, const uint32_t *packed,
                                                      const uint32_t *run_key, const uint32_t *run_base,
                                                      uint32_t dead_key)
{
    uint32_t i = blockIdx.x * kTile + threadIdx.x;
    if (i >= np) return;
    uint32_t pk = packed[i];
    uint32_t gid = blockIdx.x * kTile + (pk >> 8);
    if (run_key[gid] == dead_key) return;
    uint32_t j = run_base[gid] + (pk & 255u);
    dst.pa[j] = src.pa[i];
    for (int a = 0; a < 3; ++a) dst.v[a][j] = src.v[a][i];
    for (int q = 0; q < 9; ++q) dst.c[q][j] = src.c[q][i];
}

// ------------------------------------------------------------------------------------------------------------
// Warm start of the pressure solve on fluid cells: the last solution times WarmStartScale = dt/dt_prev
// (SloshProxy.cpp:575, applied in MarkCellsCS, Slosh.usf:569-598). With `prev` set (SLOSH_FLAG_WARM_START_EXTRAPOLATE)
// a cell that was fluid in the last two solves starts from the linear extrapolation in time of its pressure per unit
// dt, P = dt (q1 + (q1 - q2) dt / dt_prev) with q1 = P_last / dt_prev and q2 the value before; any other fluid cell
// keeps the reference's start. `prev` holds q1 for the next step (NaN where there is no history).
struct WarmStart {
    double *prev;
    double dt, dt_prev;
};

// P2G (Slosh.usf:448-563) as a per-brick gather: one workgroup per allocated brick reads the brick's sorted
// particles, accumulates the int32 fixed-point momentum and weight of every node they touch into an LDS tile
// (nodes [brick_min - 1, brick_min + 9) per axis), and stores the tile; the normalise pass sums the up to 8 tiles
// that overlap each face. No global atomics; integer sums make the result independent of order. The same
// workgroup counts particles per cell and classifies its 512 cells (BuildSolidCS :365-395, MarkCellsCS :569-598).
__global__ __launch_bounds__(256) void k_p2g(Geo g, Parts p, const uint32_t *key_off, int32_t *tiles,
                                             uint8_t *tile_valid, uint32_t *ctype, double *pres, uint8_t *new_slot,
                                             float warm, WarmStart ws, const slosh_collider *col, uint32_t ncol,
                                             uint32_t *bp_fluid, uint32_t first, const uint32_t *nactive)
{
    __shared__ int32_t tile[kTileInts];
    __shared__ int32_t cnt[kBrickCells];
    __shared__ uint32_t fluid_sum;
    const uint32_t e = first + blockIdx.x;
    if (e >= *nactive) return;
    const uint32_t b = g.active[e];
    const uint32_t slot = g.table[b];
    const bool fresh = new_slot[slot] != 0u; // recycled slot: its old pressure is not a warm start (ClearNewSlotsCS)
    int bc[3];
    brick_coord(g, b, bc);
    const int ox = bc[0] * 8 - 1, oy = bc[1] * 8 - 1, oz = bc[2] * 8 - 1;
    for (uint32_t t = threadIdx.x; t < (uint32_t)kTileInts; t += blockDim.x) tile[t] = 0;
    for (uint32_t t = threadIdx.x; t < kBrickCells; t += blockDim.x) cnt[t] = 0;
    if (threadIdx.x == 0) fluid_sum = 0u;
    __syncthreads();
    const uint32_t i0 = key_off[b], i1 = key_off[b + 1];
    // each thread's next record is loaded while the current one is splatted
    auto load = [&](uint32_t i, float4 &pa, float v[3], float c[9]) {
        pa = p.pa[i];
        for (int a = 0; a < 3; ++a) v[a] = p.v[a][i];
        for (int k = 0; k < 9; ++k) c[k] = p.c[k][i];
    };
    float4 pa_n;
    float v_n[3], c_n[9];
    if (i0 + threadIdx.x < i1) load(i0 + threadIdx.x, pa_n, v_n, c_n);
    for (uint32_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        const float4 pa = pa_n;
        float vp[3], cp[9];
        for (int a = 0; a < 3; ++a) vp[a] = v_n[a];
        for (int k = 0; k < 9; ++k) cp[k] = c_n[k];
        if (i + blockDim.x < i1) load(i + blockDim.x, pa_n, v_n, c_n);
        float xp[3] = {pa.x, pa.y, pa.z};
        for (int axis = 0; axis < 3; ++axis) {
            float up = vp[axis];
            float cp0 = cp[3 * axis], cp1 = cp[3 * axis + 1], cp2 = cp[3 * axis + 2];
            float off[3] = {0.5f, 0.5f, 0.5f};
            off[axis] = 0.0f;
            float gg[3], fr[3];
            int base[3];
            for (int a = 0; a < 3; ++a) {
                gg[a] = xp[a] * g.inv_h - off[a];
                base[a] = (int)floorf(gg[a]);
                fr[a] = gg[a] - (float)base[a];
            }
            // Per axis a and corner offset o in {0, 1}: the weight factor, whether the node is in the grid, and
            // the affine term cp_a * ((node + off) h - x). A corner combines them as w = (w_x w_y) w_z and
            // dot = (cp0 dx + cp1 dy) + cp2 dz, the same products and sums, in the same order, as per corner.
            float w1[3][2], cd[3][2];
            bool inb[3][2];
            const float cpa[3] = {cp0, cp1, cp2};
            for (int a = 0; a < 3; ++a)
                for (int o = 0; o < 2; ++o) {
                    const int node = base[a] + o;
                    inb[a][o] = node >= 0 && node < g.n[a];
                    w1[a][o] = o ? fr[a] : 1.0f - fr[a];
                    cd[a][o] = cpa[a] * (((float)node + off[a]) * g.h - xp[a]);
                }
            float wxy[2][2], cxy[2][2];
            for (int oy = 0; oy < 2; ++oy)
                for (int ox = 0; ox < 2; ++ox) {
                    wxy[oy][ox] = w1[0][ox] * w1[1][oy];
                    cxy[oy][ox] = cd[0][ox] + cd[1][oy];
                }
            // the particle lies in brick b, so every corner node lies in its tile (node - tile origin in [0, 9])
            const uint32_t li0 = (uint32_t)((base[0] - ox) + kTileDim * ((base[1] - oy) + kTileDim * (base[2] - oz)));
            int32_t *tm = tile + axis * kTileNodes + li0, *tw = tile + (3 + axis) * kTileNodes + li0;
            // All eight nodes in the grid and no fractional part zero: every weight is positive (a nonzero
            // fraction is at least 2^-15 here, so the product of three factors stays a normal float) and every
            // corner adds its sums. Otherwise a node outside the grid or of zero weight adds nothing.
            const bool all = inb[0][0] & inb[0][1] & inb[1][0] & inb[1][1] & inb[2][0] & inb[2][1] & (fr[0] != 0.0f) &
                             (fr[1] != 0.0f) & (fr[2] != 0.0f);
            if (all) {
                for (uint32_t oi = 0; oi < 8u; ++oi) {
                    const int cx = (int)(oi & 1u), cy = (int)((oi >> 1) & 1u), cz = (int)(oi >> 2);
                    const float w = wxy[cy][cx] * w1[2][cz];
                    const float dot = cxy[cy][cx] + cd[2][cz];
                    const uint32_t lo = (uint32_t)(cx + kTileDim * (cy + kTileDim * cz));
                    atomicAdd(&tm[lo], (int32_t)rintf(w * (up + dot) * kFixedScale));
                    atomicAdd(&tw[lo], (int32_t)rintf(w * kFixedScale));
                }
            } else {
                for (uint32_t oi = 0; oi < 8u; ++oi) {
                    const int cx = (int)(oi & 1u), cy = (int)((oi >> 1) & 1u), cz = (int)(oi >> 2);
                    const float w = wxy[cy][cx] * w1[2][cz];
                    if (!(inb[0][cx] & inb[1][cy] & inb[2][cz]) || w <= 0.0f) continue;
                    const float dot = cxy[cy][cx] + cd[2][cz];
                    const uint32_t lo = (uint32_t)(cx + kTileDim * (cy + kTileDim * cz));
                    atomicAdd(&tm[lo], (int32_t)rintf(w * (up + dot) * kFixedScale));
                    atomicAdd(&tw[lo], (int32_t)rintf(w * kFixedScale));
                }
            }
        }
        int c[3];
        pos_cell(g, xp[0], xp[1], xp[2], c);
        atomicAdd(&cnt[local_of_cell(c[0], c[1], c[2])], 1);
    }
    __syncthreads();
    const bool any = i1 > i0;
    if (threadIdx.x == 0) tile_valid[slot] = any ? 1u : 0u;
    if (any) {
        int32_t *dst = tiles + (size_t)slot * kTileInts;
        for (uint32_t t = threadIdx.x; t < (uint32_t)kTileInts; t += blockDim.x) dst[t] = tile[t];
    }
    uint32_t fluid = 0;
    for (uint32_t l = threadIdx.x; l < kBrickCells; l += blockDim.x) {
        int x = bc[0] * 8 + (int)(l & 7u), y = bc[1] * 8 + (int)((l >> 3) & 7u), z = bc[2] * 8 + (int)(l >> 6);
        if (!in_grid(g, x, y, z)) continue;
        uint32_t packed = kCellAir;
        if (x == 0 || y == 0 || z == 0 || x == g.n[0] - 1 || y == g.n[1] - 1 || z == g.n[2] - 1) {
            packed = kCellSolid;
        } else {
            float pc[3] = {((float)x + 0.5f) * g.h, ((float)y + 0.5f) * g.h, ((float)z + 0.5f) * g.h};
            float best = 0.0f;
            for (uint32_t k = 0; k < ncol; ++k) {
                float dd = collider_sdf(col[k], pc);
                if (dd < best) {
                    best = dd;
                    packed = kCellSolid | ((k + 1u) << kColliderShift);
                }
            }
        }
        if ((packed & kCellTypeMask) != kCellSolid) {
            int32_t count = cnt[l];
            packed = count > 0 ? kCellFluid : kCellAir;
            if ((float)count >= 0.5f * kTargetPpc) packed |= kCellDense;
        }
        uint32_t idx = slot * kBrickCells + l;
        const uint32_t old_type = ws.prev ? ctype[idx] : 0u; // last step's type (a new slot has no history)
        ctype[idx] = packed;
        if ((packed & kCellTypeMask) == kCellFluid) {
            const double old = fresh ? 0.0 : pres[idx];
            double p0 = old * (double)warm;
            if (ws.prev) {
                double q1 = __longlong_as_double(0x7ff8000000000000ll);
                if (!fresh && (old_type & kCellTypeMask) == kCellFluid && ws.dt_prev > 0.0) {
                    q1 = old / ws.dt_prev;
                    const double q2 = ws.prev[idx];
                    if (!isnan(q2)) p0 = ws.dt * (q1 + (q1 - q2) * (ws.dt / ws.dt_prev));
                }
                ws.prev[idx] = q1;
            }
            pres[idx] = p0;
            fluid++;
        } else {
            pres[idx] = 0.0;
            if (ws.prev) ws.prev[idx] = __longlong_as_double(0x7ff8000000000000ll);
        }
    }
    if (fluid) atomicAdd(&fluid_sum, fluid);
    __syncthreads();
    if (threadIdx.x == 0) {
        bp_fluid[e] = fluid_sum;
        if (fresh) new_slot[slot] = 0u;
    }
}

// ------------------------------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t face_index(const Geo &g, int axis, int x, int y, int z)
{
    return (uint32_t)axis * g.pool_cells + cell_index(g, x, y, z);
}

// Face velocities for G2P: the 12^3 nodes [brick_min - 2, brick_min + 10) of each component staged in LDS by the
// brick's workgroup, with a global read (through the brick table) for a node outside that box.
constexpr int kG2PBox = 12;
constexpr int kG2PBoxNodes = kG2PBox * kG2PBox * kG2PBox;

struct FaceSource {
    const float *vel;
    const float *box; // [3][kG2PBoxNodes], the workgroup's LDS box
    bool staged;      // the box holds this brick's neighbourhood
    int org[3];       // global node of box entry 0
    // True when the 2x2x2 nodes from `base` (before clamping to the grid) all lie in the box; clamping only moves a
    // node toward the grid interior, which the box of an in-grid brick covers.
    __device__ bool covers(const int base[3]) const
    {
        bool in = staged;
        for (int a = 0; a < 3; ++a) in = in && base[a] >= org[a] && base[a] + 1 < org[a] + kG2PBox;
        return in;
    }
    __device__ float lds(int axis, int x, int y, int z) const
    {
        return box[axis * kG2PBoxNodes + (x - org[0]) + kG2PBox * ((y - org[1]) + kG2PBox * (z - org[2]))];
    }
};

// Staged fast path of the two samplers below, for a stencil whose nodes need no clamping and lie in the box. The
// 2x2x2 weights are formed from the per-axis factors with the reference's association, w = (wx wy) wz and
// grad_a = (sign_a (w_b w_c)) / h, so every product and sum rounds exactly as in the general form.
__device__ __forceinline__ bool stencil_base(const Geo &g, const FaceSource &src, int axis, const float xp[3],
                                             int base[3], float fr[3], uint32_t &lds0)
{
    bool fast = src.staged;
    for (int a = 0; a < 3; ++a) {
        const float gg = xp[a] * g.inv_h - (a == axis ? 0.0f : 0.5f);
        base[a] = (int)floorf(gg);
        fr[a] = gg - (float)base[a];
        fast = fast && base[a] >= 0 && base[a] <= g.n[a] - 2 && base[a] >= src.org[a] &&
               base[a] + 1 < src.org[a] + kG2PBox;
    }
    lds0 = (uint32_t)(axis * kG2PBoxNodes + (base[0] - src.org[0]) +
                      kG2PBox * ((base[1] - src.org[1]) + kG2PBox * (base[2] - src.org[2])));
    return fast;
}

constexpr uint32_t kBoxCorner[8] = {0u, 1u, (uint32_t)kG2PBox, (uint32_t)kG2PBox + 1u, (uint32_t)(kG2PBox * kG2PBox),
                                    (uint32_t)(kG2PBox * kG2PBox) + 1u, (uint32_t)(kG2PBox * kG2PBox + kG2PBox),
                                    (uint32_t)(kG2PBox * kG2PBox + kG2PBox) + 1u};

__device__ __forceinline__ void gather_lds(const float *box, uint32_t lds0, const float fr[3], float inv_h,
                                           float &unew, float grad[3])
{
    const float wx[2] = {1.0f - fr[0], fr[0]}, wy[2] = {1.0f - fr[1], fr[1]}, wz[2] = {1.0f - fr[2], fr[2]};
    float u_acc = 0.0f, g0 = 0.0f, g1 = 0.0f, g2 = 0.0f;
    for (uint32_t oi = 0; oi < 8u; ++oi) {
        const uint32_t ox = oi & 1u, oy = (oi >> 1) & 1u, oz = oi >> 2;
        const float u = box[lds0 + kBoxCorner[oi]];
        const float w = wx[ox] * wy[oy] * wz[oz];
        const float a0 = (wy[oy] * wz[oz]) * inv_h, a1 = (wx[ox] * wz[oz]) * inv_h, a2 = (wx[ox] * wy[oy]) * inv_h;
        u_acc += w * u;
        g0 += (ox ? a0 : -a0) * u;
        g1 += (oy ? a1 : -a1) * u;
        g2 += (oz ? a2 : -a2) * u;
    }
    unew = u_acc;
    grad[0] = g0;
    grad[1] = g1;
    grad[2] = g2;
}

__device__ __forceinline__ float sample_lds(const float *box, uint32_t lds0, const float fr[3])
{
    const float wx[2] = {1.0f - fr[0], fr[0]}, wy[2] = {1.0f - fr[1], fr[1]}, wz[2] = {1.0f - fr[2], fr[2]};
    float u = 0.0f;
    for (uint32_t oi = 0; oi < 8u; ++oi) u += wx[oi & 1u] * wy[(oi >> 1) & 1u] * wz[oi >> 2] * box[lds0 + kBoxCorner[oi]];
    return u;
}

// Slosh.usf:1376-1410 (GatherComponent, APIC).
__device__ __forceinline__ void gather_component(const Geo &g, const FaceSource &src, int axis, const float xp[3],
                                                 float &unew, float grad[3])
{
    float off[3] = {0.5f, 0.5f, 0.5f};
    off[axis] = 0.0f;
    float gg[3], fr[3];
    int base[3];
    for (int a = 0; a < 3; ++a) {
        gg[a] = xp[a] * g.inv_h - off[a];
        base[a] = (int)floorf(gg[a]);
        fr[a] = gg[a] - (float)base[a];
    }
    float uv[8];
    if (src.covers(base)) {
        for (uint32_t oi = 0; oi < 8u; ++oi)
            uv[oi] = src.lds(axis, clampi(base[0] + (int)(oi & 1u), 0, g.n[0] - 1),
                             clampi(base[1] + (int)((oi >> 1) & 1u), 0, g.n[1] - 1),
                             clampi(base[2] + (int)((oi >> 2) & 1u), 0, g.n[2] - 1));
    } else {
        for (uint32_t oi = 0; oi < 8u; ++oi)
            uv[oi] = src.vel[face_index(g, axis, clampi(base[0] + (int)(oi & 1u), 0, g.n[0] - 1),
                                        clampi(base[1] + (int)((oi >> 1) & 1u), 0, g.n[1] - 1),
                                        clampi(base[2] + (int)((oi >> 2) & 1u), 0, g.n[2] - 1))];
    }
    float u_acc = 0.0f, g0 = 0.0f, g1 = 0.0f, g2 = 0.0f;
    for (uint32_t oi = 0; oi < 8u; ++oi) {
        int od[3] = {(int)(oi & 1u), (int)((oi >> 1) & 1u), (int)((oi >> 2) & 1u)};
        float w3[3], sg[3];
        for (int a = 0; a < 3; ++a) {
            w3[a] = od[a] ? fr[a] : 1.0f - fr[a];
            sg[a] = od[a] ? 1.0f : -1.0f;
        }
        float w = w3[0] * w3[1] * w3[2];
        float gw0 = sg[0] * (w3[1] * w3[2]) * g.inv_h;
        float gw1 = sg[1] * (w3[0] * w3[2]) * g.inv_h;
        float gw2 = sg[2] * (w3[0] * w3[1]) * g.inv_h;
        float u = uv[oi];
        u_acc += w * u;
        g0 += gw0 * u;
        g1 += gw1 * u;
        g2 += gw2 * u;
    }
    unew = u_acc;
    grad[0] = g0;
    grad[1] = g1;
    grad[2] = g2;
}

// Slosh.usf:1417-1443 (SampleVelocityComponent).
__device__ __forceinline__ float sample_component(const Geo &g, const FaceSource &src, const uint8_t *fflags, int axis,
                                                  const float xp[3], float fallback, bool check)
{
    const uint32_t known = kFaceValid | (0xFFFFFFFFu << kFaceExtrapShift);
    float off[3] = {0.5f, 0.5f, 0.5f};
    off[axis] = 0.0f;
    float gg[3], fr[3];
    int base[3];
    for (int a = 0; a < 3; ++a) {
        gg[a] = xp[a] * g.inv_h - off[a];
        base[a] = (int)floorf(gg[a]);
        fr[a] = gg[a] - (float)base[a];
    }
    float uv[8];
    if (src.covers(base) && !check) {
        for (uint32_t oi = 0; oi < 8u; ++oi)
            uv[oi] = src.lds(axis, clampi(base[0] + (int)(oi & 1u), 0, g.n[0] - 1),
                             clampi(base[1] + (int)((oi >> 1) & 1u), 0, g.n[1] - 1),
                             clampi(base[2] + (int)((oi >> 2) & 1u), 0, g.n[2] - 1));
    } else {
        for (uint32_t oi = 0; oi < 8u; ++oi) {
            const uint32_t f = face_index(g, axis, clampi(base[0] + (int)(oi & 1u), 0, g.n[0] - 1),
                                          clampi(base[1] + (int)((oi >> 1) & 1u), 0, g.n[1] - 1),
                                          clampi(base[2] + (int)((oi >> 2) & 1u), 0, g.n[2] - 1));
            float uf = src.vel[f];
            if (check && ((uint32_t)fflags[f] & known) == 0u) uf = fallback;
            uv[oi] = uf;
        }
    }
    float u = 0.0f;
    for (uint32_t oi = 0; oi < 8u; ++oi) {
        int od[3] = {(int)(oi & 1u), (int)((oi >> 1) & 1u), (int)((oi >> 2) & 1u)};
        float w3[3];
        for (int a = 0; a < 3; ++a) w3[a] = od[a] ? fr[a] : 1.0f - fr[a];
        u += w3[0] * w3[1] * w3[2] * uv[oi];
    }
    return u;
}

// Where the second G2P pass writes: each particle's new record goes straight to its place in the order sorted by
// the new brick keys (the first pass's keys, sorted by the same passes as at the start of a step).
struct G2PMove {
    Parts dst;
    const uint32_t *packed;   // per particle: in-tile run index << 8 | rank within the run
    const uint32_t *run_base; // per run: its first destination
    uint32_t *misplaced;      // count of particles whose recomputed key differs from the first pass's
};

// Slosh.usf:1451-1678 (AdvectParticle + G2PAdvectCS), APIC; manipulators, adhesion and drains are off. One
// workgroup per brick key of the sorted particle array: the brick's particles all lie inside it, so their gathers
// read the staged box (an unallocated brick, whose particles are ballistic, reads through the table).
// The step runs it twice on the same inputs. The first pass (kMove false) writes each particle's new brick key, the
// destination-brick flags and the speed statistics; the second recomputes the same values (the same operations on
// the same operands) and writes the 64-byte record to its sorted place, so the records are written once and never
// moved by a separate scatter.
// Slosh.usf:830-853 (SampleDensityDisplacement): the face displacements of the density projection interpolated
// trilinearly per axis at xp (stencil nodes clamped into the grid), shortened to half a cell, in metres.
__device__ __forceinline__ void sample_displacement(const Geo &g, const float *disp, const float xp[3], float out[3])
{
    float dv[3];
    for (int axis = 0; axis < 3; ++axis) {
        float off[3] = {0.5f, 0.5f, 0.5f};
        off[axis] = 0.0f;
        int base[3];
        float fr[3];
        for (int a = 0; a < 3; ++a) {
            const float gg = xp[a] * g.inv_h - off[a];
            base[a] = (int)floorf(gg);
            fr[a] = gg - (float)base[a];
        }
        float value = 0.0f;
        for (uint32_t oi = 0; oi < 8u; ++oi) {
            const int od[3] = {(int)(oi & 1u), (int)((oi >> 1) & 1u), (int)((oi >> 2) & 1u)};
            int node[3];
            float w3[3];
            for (int a = 0; a < 3; ++a) {
                const int v = base[a] + od[a];
                node[a] = v < 0 ? 0 : (v > g.n[a] - 1 ? g.n[a] - 1 : v);
                w3[a] = od[a] ? fr[a] : 1.0f - fr[a];
            }
            const uint32_t b = (uint32_t)((node[0] >> 3) + g.bd[0] * ((node[1] >> 3) + g.bd[1] * (node[2] >> 3)));
            const uint32_t f = (uint32_t)axis * g.pool_cells + g.table[b] * kBrickCells +
                               local_of_cell(node[0], node[1], node[2]);
            value += w3[0] * w3[1] * w3[2] * disp[f];
        }
        dv[axis] = value;
    }
    const float len = sqrtf(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
    const float s = len > 0.5f ? 0.5f / len : 1.0f;
    for (int a = 0; a < 3; ++a) out[a] = (len > 0.5f ? dv[a] * s : dv[a]) * g.h;
}

// kExtras: the density projection's position shift, force fields and drains (compiled out of the common step)
template <bool kMove, bool kExtras>
__device__ __forceinline__ void g2p_body(Geo g, Parts p, const uint32_t *key_off, const float *vel,
                                         const uint8_t *fflags, float dt, float gx, float gy, float gz, float friction,
                                         float radius, float max_speed, const slosh_collider *col, uint32_t ncol,
                                         uint8_t *touch, uint32_t *bp_speed, uint32_t *bp_clamp, uint32_t *keys,
                                         uint32_t dead_key, const G2PMove &mv, const float *disp,
                                         const slosh_manipulator *man, uint32_t nman, uint32_t frame_seed,
                                         uint32_t *bp_drain)
{
    __shared__ float s_box[3 * kG2PBoxNodes];
    __shared__ uint32_t s_nb[27]; // slots of the 3x3x3 brick neighbourhood, (dx+1) + 3(dy+1) + 9(dz+1)
    __shared__ uint32_t smax, sclamp, sdrain;
    const uint32_t key = blockIdx.x;
    const uint32_t beg = key_off[key], end = key_off[key + 1];
    if (threadIdx.x == 0) {
        smax = 0u;
        sclamp = 0u;
        sdrain = 0u;
    }
    if (beg == end) {
        if (!kMove && threadIdx.x == 0) {
            bp_speed[key] = 0u;
            bp_clamp[key] = 0u;
            bp_drain[key] = 0u;
        }
        return;
    }
    int bc[3];
    brick_coord(g, key, bc);
    if (threadIdx.x < 27u) {
        const int x = bc[0] + (int)(threadIdx.x % 3u) - 1, y = bc[1] + (int)(threadIdx.x / 3u % 3u) - 1,
                  z = bc[2] + (int)(threadIdx.x / 9u) - 1;
        s_nb[threadIdx.x] = x >= 0 && y >= 0 && z >= 0 && x < g.bd[0] && y < g.bd[1] && z < g.bd[2]
                                ? g.table[(uint32_t)(x + g.bd[0] * (y + g.bd[1] * z))]
                                : 0u;
    }
    // the first position records, loaded while the box is staged
    const uint32_t i0 = beg + threadIdx.x;
    float4 pa_next = i0 < end ? p.pa[i0] : make_float4(0.0f, 0.0f, 0.0f, -1.0f);
    __syncthreads();
    const bool staged = s_nb[13] != 0u;
    FaceSource src;
    src.vel = vel;
    src.box = s_box;
    src.staged = staged;
    for (int a = 0; a < 3; ++a) src.org[a] = bc[a] * 8 - 2;
    if (staged) {
        // the box's nodes lie in the 3x3x3 neighbourhood; U loads in flight per thread before their stores
        constexpr uint32_t U = 7u, N = 3u * (uint32_t)kG2PBoxNodes;
        for (uint32_t k0 = threadIdx.x; k0 < N; k0 += 256u * U) {
            float v[U];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const uint32_t k = k0 + 256u * u;
                v[u] = 0.0f;
                if (k < N) {
                    const int axis = (int)(k / kG2PBoxNodes), q = (int)(k % kG2PBoxNodes);
                    const int x = src.org[0] + q % kG2PBox, y = src.org[1] + (q / kG2PBox) % kG2PBox,
                              z = src.org[2] + q / (kG2PBox * kG2PBox);
                    if (in_grid(g, x, y, z)) {
                        const uint32_t slot =
                            s_nb[((x >> 3) - bc[0] + 1) + 3 * ((y >> 3) - bc[1] + 1) + 9 * ((z >> 3) - bc[2] + 1)];
                        v[u] = vel[(uint32_t)axis * g.pool_cells + slot * kBrickCells + local_of_cell(x, y, z)];
                    }
                }
            }
#pragma unroll
            for (uint32_t u = 0; u < U; ++u)
                if (k0 + 256u * u < N) s_box[k0 + 256u * u] = v[u];
        }
    }
    __syncthreads();
    uint32_t vmax = 0u, nclamp = 0u, ndrain = 0u; // per thread; speeds are non-negative, so their bits order like
                                                  // the values
    for (uint32_t base_i = beg; base_i < end; base_i += 256u) {
        const uint32_t i = base_i + threadIdx.x;
        uint32_t brick = kNoBrick;
        const float4 pa = pa_next;
        if (i + 256u < end) pa_next = p.pa[i + 256u];
        // second pass: the record's destination and its first-pass key, loaded before the gathers so that their
        // latency overlaps the arithmetic
        uint32_t dst_j = 0u, key_first = 0u;
        if (kMove && i < end) {
            const uint32_t pk = mv.packed[i];
            dst_j = mv.run_base[(i & ~(kTile - 1u)) + (pk >> 8)] + (pk & 255u);
            key_first = keys[i];
        }
        if (i < end) {
            if (!kMove && !(pa.w >= 0.0f)) keys[i] = dead_key; // a dead record is dropped by the sort
            if (pa.w >= 0.0f) {
                const float grav[3] = {gx, gy, gz};
                float pos[3] = {pa.x, pa.y, pa.z};
                float pic[3], cm[9];
                for (int a = 0; a < 3; ++a) {
                    int bs[3];
                    float fr[3];
                    uint32_t l0;
                    if (stencil_base(g, src, a, pos, bs, fr, l0)) gather_lds(s_box, l0, fr, g.inv_h, pic[a], &cm[3 * a]);
                    else gather_component(g, src, a, pos, pic[a], &cm[3 * a]);
                }
                float nv[3] = {pic[0], pic[1], pic[2]};
                float lo = g.h * 1.01f;
                float hi[3];
                for (int a = 0; a < 3; ++a) hi[a] = ((float)g.n[a] - 1.01f) * g.h;
                float hdt = 0.5f * dt;
                float mid[3];
                for (int a = 0; a < 3; ++a) mid[a] = clampf(pos[a] + hdt * pic[a], lo, hi[a]);
                float midcells = hdt * len3(pic[0], pic[1], pic[2]) * g.inv_h;
                float v2[3] = {pic[0], pic[1], pic[2]};
                if (midcells < 2.0f) {
                    bool check = midcells > 1.0f;
                    float t = clampf(midcells - 1.0f, 0.0f, 1.0f);
                    for (int a = 0; a < 3; ++a) {
                        int bs[3];
                        float fr[3];
                        uint32_t l0;
                        float sv;
                        if (!check && stencil_base(g, src, a, mid, bs, fr, l0)) sv = sample_lds(s_box, l0, fr);
                        else sv = sample_component(g, src, fflags, a, mid, pic[a], check);
                        v2[a] = sv + t * (pic[a] - sv);
                    }
                }
                float x[3];
                for (int a = 0; a < 3; ++a) x[a] = pos[a] + dt * v2[a];
                if (kExtras && disp) {
                    // position-only shift of the density projection, sampled where the density was measured
                    float pc[3], dx[3];
                    for (int a = 0; a < 3; ++a) pc[a] = clampf(pos[a], lo, hi[a]);
                    sample_displacement(g, disp, pc, dx);
                    for (int a = 0; a < 3; ++a) x[a] = x[a] + dx[a];
                }
                int c[3];
                // Every live particle of this key lies in brick `key` (its sort key is the brick of its clamped
                // cell), so its cell is in an unallocated brick exactly when the brick is not staged.
                if (!staged) {
                    for (int a = 0; a < 3; ++a) {
                        nv[a] = p.v[a][i] + grav[a] * dt;
                        x[a] = pos[a] + dt * nv[a];
                    }
                    for (int k = 0; k < 9; ++k) cm[k] = 0.0f;
                }
                if (kExtras && nman) apply_force_fields(man, nman, x, nv, dt);
                for (uint32_t k = 0; k < ncol; ++k) {
                    const slosh_collider cc = col[k];
                    float dd = collider_sdf(cc, x);
                    if (dd < radius) {
                        float nn[3];
                        collider_normal(cc, x, 0.25f * g.h, nn);
                        float push = radius - dd;
                        for (int a = 0; a < 3; ++a) x[a] = x[a] + nn[a] * push;
                        float vs[3] = {collider_velocity_axis(cc, x, 0), collider_velocity_axis(cc, x, 1),
                                       collider_velocity_axis(cc, x, 2)};
                        float rel[3] = {nv[0] - vs[0], nv[1] - vs[1], nv[2] - vs[2]};
                        float vn = rel[0] * nn[0] + rel[1] * nn[1] + rel[2] * nn[2];
                        if (vn < 0.0f) {
                            for (int a = 0; a < 3; ++a) rel[a] = rel[a] - vn * nn[a];
                            float rn = rel[0] * nn[0] + rel[1] * nn[1] + rel[2] * nn[2];
                            float tg[3];
                            for (int a = 0; a < 3; ++a) tg[a] = rel[a] - rn * nn[a];
                            for (int a = 0; a < 3; ++a) rel[a] = rel[a] - friction * tg[a];
                        }
                        for (int a = 0; a < 3; ++a) nv[a] = vs[a] + rel[a];
                    }
                }
                uint32_t contact = 0;
                for (int a = 0; a < 3; ++a) {
                    if (x[a] < lo) {
                        x[a] = lo;
                        nv[a] = fmaxf(nv[a], 0.0f);
                        contact |= 1u << a;
                    } else if (x[a] > hi[a]) {
                        x[a] = hi[a];
                        nv[a] = fminf(nv[a], 0.0f);
                        contact |= 1u << a;
                    }
                }
                if (contact != 0u && friction > 0.0f)
                    for (int a = 0; a < 3; ++a)
                        if ((contact & (1u << a)) == 0u) nv[a] *= 1.0f - friction;
                float speed = len3(nv[0], nv[1], nv[2]);
                if (kExtras && nman && drained_at(man, nman, x, dt, i, frame_seed)) {
                    // removed: the sort drops the record (its key is the dead key), its brick is not touched (the
                    // lane keeps brick = kNoBrick and still takes part in the wavefront's touch marking below)
                    if (!kMove) {
                        if (speed > max_speed) nclamp++;
                        ndrain++;
                        keys[i] = dead_key;
                    }
                } else {
                    pos_cell(g, x[0], x[1], x[2], c);
                    brick = brick_of_cell(g, c[0], c[1], c[2]);
                    if (kMove) {
                        const uint32_t j = dst_j;
                        mv.dst.pa[j] = make_float4(x[0], x[1], x[2], pa.w + dt);
                        for (int a = 0; a < 3; ++a) mv.dst.v[a][j] = nv[a];
                        for (int k = 0; k < 9; ++k) mv.dst.c[k][j] = cm[k];
                        if (brick != key_first) atomicAdd(mv.misplaced, 1u);
                    } else {
                        if (speed > max_speed) nclamp++;
                        vmax = max(vmax, __float_as_uint(speed));
                        keys[i] = brick;
                    }
                }
            }
        }
        if (!kMove) {
            // mark the destination bricks, once per run of equal bricks within the wavefront (flags are a set)
            const uint32_t prev = (uint32_t)__shfl_up((int)brick, 1u, 64);
            if (brick != kNoBrick && ((threadIdx.x & 63u) == 0u || prev != brick)) touch[brick] = 1u;
        }
    }
    if (kMove) return;
    for (int w = 32; w > 0; w >>= 1) {
        vmax = max(vmax, (uint32_t)__shfl_xor((int)vmax, w));
        nclamp += (uint32_t)__shfl_xor((int)nclamp, w);
        ndrain += (uint32_t)__shfl_xor((int)ndrain, w);
    }
    if ((threadIdx.x & 63u) == 0u) {
        atomicMax(&smax, vmax);
        atomicAdd(&sclamp, nclamp);
        atomicAdd(&sdrain, ndrain);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        bp_speed[key] = smax;
        bp_clamp[key] = sclamp;
        bp_drain[key] = sdrain;
    }
}

template <bool kExtras>
__global__ __launch_bounds__(256) void k_g2p_keys(Geo g, Parts p, const uint32_t *key_off, const float *vel,
                                                  const uint8_t *fflags, float dt, float gx, float gy, float gz,
                                                  float friction, float radius, float max_speed,
                                                  const slosh_collider *col, uint32_t ncol, uint8_t *touch,
                                                  uint32_t *bp_speed, uint32_t *bp_clamp, uint32_t *keys,
                                                  uint32_t dead_key, const float *disp, const slosh_manipulator *man,
                                                  uint32_t nman, uint32_t frame_seed, uint32_t *bp_drain)
{
    g2p_body<false, kExtras>(g, p, key_off, vel, fflags, dt, gx, gy, gz, friction, radius, max_speed, col, ncol, touch,
                             bp_speed,
                    bp_clamp, keys, dead_key, G2PMove{}, disp, man, nman, frame_seed, bp_drain);
}

template <bool kExtras>
__global__ __launch_bounds__(256) void k_g2p_move(Geo g, Parts p, const uint32_t *key_off, const float *vel,
                                                  const uint8_t *fflags, float dt, float gx, float gy, float gz,
                                                  float friction, float radius, float max_speed,
                                                  const slosh_collider *col, uint32_t ncol, uint32_t *keys, G2PMove mv,
                                                  const float *disp, const slosh_manipulator *man, uint32_t nman,
                                                  uint32_t frame_seed)
{
    g2p_body<true, kExtras>(g, p, key_off, vel, fflags, dt, gx, gy, gz, friction, radius, max_speed, col, ncol, nullptr,
                            nullptr,
                   nullptr, keys, 0u, mv, disp, man, nman, frame_seed, nullptr);
}

inline uint32_t blocks_for(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

} // namespace

void launch_emit(slosh_ctx *c, const slosh_emit_request *r, uint32_t count)
{
    if (!count) return;
    {
        launch_timed(c, K_EMIT, k_emit, dim3(blocks_for(count, 256)), dim3(256), 0, c->geo, c->parts[c->cur],
                     c->np, count, *r, c->touch);
    }
    SLOSH_HIP(hipGetLastError());
}

void launch_unpack_particles(slosh_ctx *c, const slosh_particle *d_src, uint32_t n)
{
    if (!n) return;
    hipLaunchKernelGGL(k_unpack, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, c->geo, c->parts[c->cur], d_src, n,
                       c->touch);
    SLOSH_HIP(hipGetLastError());
}

void launch_pack_particles(slosh_ctx *c, void *d_dst, int layout)
{
    if (!c->np) return;
    hipLaunchKernelGGL(k_pack, dim3(blocks_for(c->np, 256)), dim3(256), 0, c->stream, c->parts[c->cur], c->np, d_dst,
                       layout);
    SLOSH_HIP(hipGetLastError());
}

// The sort's key passes: stable in-tile ranks and runs, per-key counts and offsets, and each run's first
// destination. Keys come from the positions, or from `keys` when given. The offsets go to the spare buffer
// (key_off_next); the caller swaps it in once nothing reads the current array's offsets any more.
static void sort_keys(slosh_ctx *c, const uint32_t *keys)
{
    const uint32_t np = c->np;
    const uint32_t dead = c->nkeys - 1u;
    launch_zero3_u32(c->stream, c->key_count, c->key_runs, c->run_cursor, c->nkeys);
    if (np == 0) {
        scan_exclusive(c, c->key_count, c->key_off_next, c->nkeys);
        return;
    }
    const uint32_t ntiles = blocks_for(np, kTile);
    launch_timed(c, K_SORT_TILES, k_sort_tiles, dim3(ntiles), dim3(kTile), 0, c->geo, c->parts[c->cur].pa, keys, np,
                 dead, c->packed, c->tile_nruns, c->run_key, c->run_len, c->key_count, c->key_runs);
    SLOSH_HIP(hipGetLastError());
    scan_exclusive2(c, c->key_count, c->key_off_next, c->nkeys, c->key_runs, c->run_off, c->nkeys);
    launch_timed(c, K_SORT_RUNS, k_sort_runs_scatter, dim3(ntiles), dim3(kTile), 0, c->tile_nruns, c->run_key,
                 c->run_off, c->run_cursor, c->run_sorted);
    SLOSH_HIP(hipGetLastError());
    launch_timed(c, K_SORT_KEY_RUNS, k_sort_key_runs, dim3(c->nkeys - 1u), dim3(256), 0, c->run_off, c->key_off_next,
                 c->run_len, c->run_sorted, c->run_base);
    SLOSH_HIP(hipGetLastError());
}

void run_sort(slosh_ctx *c)
{
    sort_keys(c, nullptr);
    std::swap(c->key_off, c->key_off_next);
    if (c->np == 0) return;
    launch_timed(c, K_SORT_SCATTER, k_sort_scatter, dim3(blocks_for(c->np, kTile)), dim3(kTile), 0, c->parts[c->cur],
                 c->parts[1 - c->cur], c->np, c->packed, c->run_key, c->run_base, c->nkeys - 1u);
    SLOSH_HIP(hipGetLastError());
    c->cur = 1 - c->cur;
    c->sorted = true;
}

void launch_p2g(slosh_ctx *c, float warm_scale, float dt, float dt_prev, uint32_t first, uint32_t count)
{
    const WarmStart ws{c->pres_prev, (double)dt, (double)dt_prev};
    if (!count) return;
    launch_timed(c, K_P2G, k_p2g, dim3(count), dim3(256), 0, c->geo, c->parts[c->cur], c->key_off, c->tiles,
                 c->tile_valid, c->ctype, c->pres, c->new_slot, warm_scale, ws, c->d_col, c->ncol, c->bp_fluid, first,
                 device_nactive(c));
    SLOSH_HIP(hipGetLastError());
}

// G2P and the sort by the new brick keys in one: the first pass writes each particle's new key, the key passes sort
// them, and the second pass writes every new record straight to its sorted place in the other buffer. The next step
// then starts from a sorted array and skips its own sort (until particles are emitted or written).
void launch_g2p(slosh_ctx *c, float dt, const float g[3])
{
    if (!c->np) return;
    const Geo &geo = c->geo;
    const float radius = c->s.particle_radius_cells * geo.h;
    const bool extras = c->dens_applied || c->nman > 0u;
    launch_timed(c, K_G2P, extras ? k_g2p_keys<true> : k_g2p_keys<false>, dim3(geo.nbricks), dim3(256), 0, geo, c->parts[c->cur], c->key_off,
                 c->vel[c->vcur], c->fflags, dt, g[0], g[1], g[2], c->s.boundary_friction, radius, c->s.max_speed,
                 c->d_col, c->ncol, c->touch, c->bp_speed, c->bp_clamp, c->g2p_keys, c->nkeys - 1u,
                 c->dens_applied ? (const float *)c->dens_disp : nullptr, c->d_man, c->nman,
                 (uint32_t)(c->frame * 16u), c->bp_drain);
    c->n_g2p_blocks = geo.nbricks;
    SLOSH_HIP(hipGetLastError());
    // the new offsets go to the spare buffer: the second pass walks the current array by its own offsets
    sort_keys(c, c->g2p_keys);
    G2PMove mv{c->parts[1 - c->cur], c->packed, c->run_base, &c->stats->misplaced};
    launch_timed(c, K_G2P_MOVE, extras ? k_g2p_move<true> : k_g2p_move<false>, dim3(geo.nbricks), dim3(256), 0, geo, c->parts[c->cur], c->key_off,
                 c->vel[c->vcur], c->fflags, dt, g[0], g[1], g[2], c->s.boundary_friction, radius, c->s.max_speed,
                 c->d_col, c->ncol, c->g2p_keys, mv, c->dens_applied ? (const float *)c->dens_disp : nullptr,
                 c->d_man, c->nman, (uint32_t)(c->frame * 16u));
    SLOSH_HIP(hipGetLastError());
    std::swap(c->key_off, c->key_off_next);
    c->cur = 1 - c->cur;
    c->sorted = true;
}

} // namespace slosh
