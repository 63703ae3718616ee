// pooled.hip -- multigrid levels 1 and 2 stored in the brick pool, on gfx950.
//
// The reference keeps every coarse level as a dense grid (SloshProxy.cpp:269-285), so level 1 of a 256^3 domain is
// a 2M-cell array swept in full although the liquid occupies a thin layer. Here a level-l cell (l = 1, 2) lives
// with the 8^3 brick that contains its children: (8 >> l)^3 cells per allocated brick, at slot * cells_per_brick +
// local. Level 3 has one cell per brick and is stored densely by brick index. Unallocated bricks read as the zero null slot, i.e. AIR, which is what the reference's coarsening rule
// gives for an all-air block. Cells beyond the level's extent (partial bricks) are typed SOLID and never touched.
//
// Operations are those of SloshPressureMG.usf (MGBuildTypeCS :256-278, MGSmoothCS :311-338, MGRestrictCS
// :343-381, MGProlongCS :418-433) on the counted stencil with the 1e-4 diagonal term. A level's V-cycle work is
// two passes: down = black sweep + residual + restriction, up = prolongation + black sweep + red sweep. A pass
// needs the other colour's new value at a neighbour that another workgroup is producing, so it recomputes that
// value from inputs no workgroup writes in the pass (the down-sweep iterate x and the up-sweep result x1 are
// separate buffers); each value is the reference's expression in the reference's operand order.
#include "slosh_ctx.h"

namespace slosh {

namespace {

inline uint32_t blocks_for(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

struct PCell {
    int x, y, z;       // level coordinates
    uint32_t idx;      // pooled index
    uint32_t slot;
    int lx, ly, lz;    // coordinates within the brick's block
    bool inside;       // within the level's extent
};

// Thread i of a pass over the allocated cells of pooled level L.
__device__ __forceinline__ bool pcell(const Geo &g, const PLevel &L, uint32_t nactive, uint32_t i, PCell &c)
{
    const uint32_t e = i >> (3 * L.ls);
    if (e >= nactive) return false;
    const uint32_t local = i & (L.cpb - 1u);
    const uint32_t b = g.active[e];
    int bc[3];
    brick_coord(g, b, bc);
    const int m = L.s - 1;
    c.lx = (int)(local & (uint32_t)m);
    c.ly = (int)((local >> L.ls) & (uint32_t)m);
    c.lz = (int)(local >> (2 * L.ls));
    c.x = bc[0] * L.s + c.lx;
    c.y = bc[1] * L.s + c.ly;
    c.z = bc[2] * L.s + c.lz;
    c.slot = g.table[b];
    c.idx = c.slot * L.cpb + local;
    c.inside = c.x < L.dim[0] && c.y < L.dim[1] && c.z < L.dim[2];
    return true;
}

// Coarse types of pooled level 1 from level 0 (children in the same brick).
__global__ void k_ptype1(Geo g, uint32_t nactive, const uint32_t *ctype, PLevel L, int fluid_first)
{
    PCell c;
    if (!pcell(g, L, nactive, blockIdx.x * blockDim.x + threadIdx.x, c)) return;
    if (!c.inside) {
        L.type[c.idx] = (uint8_t)kCellSolid;
        return;
    }
    bool air = false, fluid = false;
    for (uint32_t oi = 0; oi < 8u; ++oi) {
        int fx = 2 * c.x + (int)(oi & 1u), fy = 2 * c.y + (int)((oi >> 1) & 1u), fz = 2 * c.z + (int)(oi >> 2);
        if (!in_grid(g, fx, fy, fz)) continue;
        uint32_t t = ctype[c.slot * kBrickCells + local_of_cell(fx, fy, fz)] & kCellTypeMask;
        air = air || t == kCellAir;
        fluid = fluid || t == kCellFluid;
    }
    L.type[c.idx] = coarse_type(air, fluid, fluid_first);
}

// Coarse types of pooled level C from pooled level F = C - 1 (children in the same brick).
__global__ void k_ptype(Geo g, uint32_t nactive, PLevel F, PLevel C, int fluid_first)
{
    PCell c;
    if (!pcell(g, C, nactive, blockIdx.x * blockDim.x + threadIdx.x, c)) return;
    if (!c.inside) {
        C.type[c.idx] = (uint8_t)kCellSolid;
        return;
    }
    bool air = false, fluid = false;
    for (uint32_t oi = 0; oi < 8u; ++oi) {
        int ox = (int)(oi & 1u), oy = (int)((oi >> 1) & 1u), oz = (int)(oi >> 2);
        if (!pinside(F, 2 * c.x + ox, 2 * c.y + oy, 2 * c.z + oz)) continue;
        uint32_t fl = (uint32_t)((2 * c.lx + ox) + F.s * ((2 * c.ly + oy) + F.s * (2 * c.lz + oz)));
        uint32_t t = F.type[c.slot * F.cpb + fl];
        air = air || t == kCellAir;
        fluid = fluid || t == kCellFluid;
    }
    C.type[c.idx] = coarse_type(air, fluid, fluid_first);
}

// Stencil word of every allocated cell of pooled level L (after its types).
__global__ void k_pinfo(Geo g, uint32_t nactive, PLevel L)
{
    PCell c;
    if (!pcell(g, L, nactive, blockIdx.x * blockDim.x + threadIdx.x, c)) return;
    uint32_t info = 0u;
    if (c.inside && L.type[c.idx] == kCellFluid) {
        info = kInfoSelf | (is_red(c.x, c.y, c.z) ? kInfoRed : 0u);
        uint32_t cnt = 0u;
        for (int n = 0; n < 6; ++n) {
            int d[3];
            nb_offset(n, d);
            const int x = c.x + d[0], y = c.y + d[1], z = c.z + d[2];
            if (!pinside(L, x, y, z)) continue;
            const uint32_t t = L.type[pindex(g, L, x, y, z)];
            if (t != kCellSolid) cnt++;
            if (t == kCellFluid) info |= 1u << n;
        }
        info |= cnt << 8;
    }
    L.info[c.idx] = (uint16_t)info;
}

// Dense level C (level 3, one cell per brick, and coarser) types from pooled level F.
__global__ void k_dtype_from_pooled(Geo g, PLevel F, DenseLevel C, int fluid_first)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C.cells) return;
    const uint32_t pc = C.coord[i];
    const int x = (int)(pc & 1023u), y = (int)((pc >> 10) & 1023u), z = (int)(pc >> 20);
    bool air = false, fluid = false;
    for (uint32_t oi = 0; oi < 8u; ++oi) {
        int fx = 2 * x + (int)(oi & 1u), fy = 2 * y + (int)((oi >> 1) & 1u), fz = 2 * z + (int)(oi >> 2);
        if (!pinside(F, fx, fy, fz)) continue;
        uint32_t t = F.type[pindex(g, F, fx, fy, fz)];
        air = air || t == kCellAir;
        fluid = fluid || t == kCellFluid;
    }
    C.type[i] = coarse_type(air, fluid, fluid_first);
}

// ---- levels 1 and 2: one workgroup per brick, the brick's block and a two-cell halo staged in LDS ----------
// The box spans block-relative coordinates [-2, S + 2)^3 (S = cells per brick axis); its cells come from the
// brick's 3x3x3 neighbourhood. A black value needed on the one-cell halo is computed there from the staged red
// values, so the block's residual (down) or red sweep (up) needs nothing another workgroup writes.
template <int S> struct Box {
    static constexpr int BD = S + 4;
    static constexpr int NB = BD * BD * BD;
    static constexpr int CPB = S * S * S;
    static constexpr int IN = S + 2; // the block plus a one-cell halo
    __device__ static int off(int n)
    {
        const int s = n < 2 ? 1 : (n < 4 ? BD : BD * BD);
        return (n & 1) ? s : -s;
    }
    __device__ static uint32_t index(const uint32_t *nb27, int u, int v, int w)
    {
        const int ox = u < 0 ? -1 : (u >= S ? 1 : 0), oy = v < 0 ? -1 : (v >= S ? 1 : 0),
                  oz = w < 0 ? -1 : (w >= S ? 1 : 0);
        const uint32_t slot = nb27[(ox + 1) + 3 * (oy + 1) + 9 * (oz + 1)];
        return slot * (uint32_t)CPB + (uint32_t)((u - ox * S) + S * ((v - oy * S) + S * (w - oz * S)));
    }
    __device__ static uint32_t box_of(int u, int v, int w) { return (uint32_t)((u + 2) + BD * ((v + 2) + BD * (w + 2))); }
};

template <int S> __device__ __forceinline__ float box_sum(const float *v, uint32_t k, uint32_t info)
{
    float s = 0.0f;
    for (int n = 0; n < 6; ++n)
        if (info >> n & 1u) s += v[(int)k + Box<S>::off(n)];
    return s;
}

// Workgroup size of the pooled passes. Level 2 (S = 2, a 6^3 box per brick) takes one wavefront per brick, so that
// every brick's workgroup is resident at once: the pass moves little data and its time is the latency of a few
// dependent loads, paid once instead of once per round of workgroups. Level 1 (an 8^3 box) has enough work per brick
// that four wavefronts finish it sooner.
template <int S> constexpr uint32_t pool_threads() { return S == 4 ? 256u : 64u; }

// Down pass of pooled level L (S = 8 >> l): black sweep, residual, restriction into pooled level C.
template <int S>
__device__ __forceinline__ void pdown_body(const Blk &G, const uint32_t *binfo, PLevel L, PLevel C, DenseLevel D,
                                           const SolverState *st)
{
    using Bx = Box<S>;
    constexpr int S2 = S / 2, CPB2 = S2 * S2 * S2;
    constexpr uint32_t kPoolThreads = pool_threads<S>();
    constexpr int kPer = (Bx::NB + (int)kPoolThreads - 1) / (int)kPoolThreads;
    __shared__ uint32_t s_nb[28];
    __shared__ float s_x[Bx::NB], s_b[Bx::NB];
    __shared__ uint16_t s_info[Bx::NB];
    __shared__ float s_res[Bx::CPB];
    const uint32_t t = threadIdx.x;
    const uint32_t e = xcd_remap(G.b, G.nb);
    if (t < 28u) s_nb[t] = binfo[(size_t)e * kBinfoWords + t];
    const uint32_t done = st->done;
    // the level-2 pass restricts into level 3 (dense by brick index): its target word is known from the brick index
    const uint32_t b3 = binfo[(size_t)e * kBinfoWords + 27];
    __syncthreads();
    if (done) return;
    uint32_t vi[kPer];
    float vx[kPer], vb[kPer];
    for (int q = 0; q < kPer; ++q) {
        const uint32_t k = t + kPoolThreads * (uint32_t)q;
        vi[q] = 0u;
        vx[q] = vb[q] = 0.0f;
        if (k >= (uint32_t)Bx::NB) continue;
        const int u = (int)(k % Bx::BD) - 2, v = (int)((k / Bx::BD) % Bx::BD) - 2, w = (int)(k / (Bx::BD * Bx::BD)) - 2;
        const uint32_t gi = Bx::index(s_nb, u, v, w);
        vi[q] = L.info[gi];
        vx[q] = L.x[gi];
        vb[q] = L.b[gi];
    }
    const uint32_t slot = s_nb[13];
    const uint32_t ci = S == 2 ? b3 : slot * CPB2 + (t & (uint32_t)(CPB2 - 1));
    const uint32_t cinfo = S == 2 ? D.info[ci] : C.info[ci];
    for (int q = 0; q < kPer; ++q) {
        const uint32_t k = t + kPoolThreads * (uint32_t)q;
        if (k >= (uint32_t)Bx::NB) continue;
        const bool f = (vi[q] & kInfoSelf) != 0u;
        s_info[k] = (uint16_t)vi[q];
        s_x[k] = f ? vx[q] : 0.0f;
        s_b[k] = f ? vb[q] : 0.0f;
    }
    __syncthreads();
    for (uint32_t k = t; k < (uint32_t)(Bx::IN * Bx::IN * Bx::IN); k += kPoolThreads) {
        const int u = (int)(k % Bx::IN) - 1, v = (int)((k / Bx::IN) % Bx::IN) - 1, w = (int)(k / (Bx::IN * Bx::IN)) - 1;
        const uint32_t kb = Bx::box_of(u, v, w), info = s_info[kb];
        if (!(info & kInfoSelf) || (info & kInfoRed)) continue;
        s_x[kb] = (s_b[kb] + box_sum<S>(s_x, kb, info)) / info_diag(info);
    }
    __syncthreads();
    for (uint32_t k = t; k < (uint32_t)Bx::CPB; k += kPoolThreads) {
        const int u = (int)(k % S), v = (int)((k / S) % S), w = (int)(k / (S * S));
        const uint32_t kb = Bx::box_of(u, v, w), info = s_info[kb];
        float r = 0.0f;
        if (info & kInfoSelf) {
            if (!(info & kInfoRed)) L.x[slot * Bx::CPB + k] = s_x[kb];
            r = s_b[kb] - (info_diag(info) * s_x[kb] - box_sum<S>(s_x, kb, info));
        }
        s_res[k] = r;
    }
    __syncthreads();
    if (t >= (uint32_t)CPB2) return;
    {
        const uint32_t k = t;
        const int cx = (int)(k % S2), cy = (int)((k / S2) % S2), cz = (int)(k / (S2 * S2));
        const LVec cb = S == 2 ? D.b : C.b, cxv = S == 2 ? D.x : C.x;
        if (!(cinfo & kInfoSelf)) {
            cb[ci] = 0.0f;
            cxv[ci] = 0.0f;
            return;
        }
        float sum = 0.0f;
        for (uint32_t oi = 0; oi < 8u; ++oi) {
            const int fx = 2 * cx + (int)(oi & 1u), fy = 2 * cy + (int)((oi >> 1) & 1u), fz = 2 * cz + (int)(oi >> 2);
            if (s_info[Bx::box_of(fx, fy, fz)] & kInfoSelf) sum += s_res[fx + S * (fy + S * fz)];
        }
        const float bv = 0.25f * sum;
        cb[ci] = bv;
        cxv[ci] = (cinfo & kInfoRed) ? bv / info_diag(cinfo) : 0.0f;
    }
}

// Up pass of pooled level L (S = 8 >> l) from pooled parent level C: prolongation, black sweep, red sweep.
template <int S>
__device__ __forceinline__ void pup_body(const Blk &G, const uint32_t *binfo, PLevel L, PLevel C, DenseLevel D,
                                         const SolverState *st)
{
    using Bx = Box<S>;
    constexpr int S2 = S / 2, CPB2 = S2 * S2 * S2;
    constexpr uint32_t kPoolThreads = pool_threads<S>();
    constexpr int kPer = (Bx::NB + (int)kPoolThreads - 1) / (int)kPoolThreads;
    __shared__ uint32_t s_nb[28];
    __shared__ float s_v[Bx::NB], s_b[Bx::NB];
    __shared__ uint16_t s_info[Bx::NB];
    const uint32_t t = threadIdx.x;
    const uint32_t e = xcd_remap(G.b, G.nb);
    if (t < 28u) s_nb[t] = binfo[(size_t)e * kBinfoWords + t];
    const uint32_t done = st->done;
    __syncthreads();
    if (done) return;
    // every box cell's word, iterate, right-hand side and its parent's word and result, all issued at once (a
    // cell outside the grid reads the null slot, and for level 2 its own brick's level-3 cell)
    uint32_t vi[kPer], pinf[kPer];
    float vx[kPer], vb[kPer], px[kPer];
    for (int q = 0; q < kPer; ++q) {
        const uint32_t k = t + kPoolThreads * (uint32_t)q;
        vi[q] = pinf[q] = 0u;
        vx[q] = vb[q] = px[q] = 0.0f;
        if (k >= (uint32_t)Bx::NB) continue;
        const int u = (int)(k % Bx::BD) - 2, v = (int)((k / Bx::BD) % Bx::BD) - 2, w = (int)(k / (Bx::BD * Bx::BD)) - 2;
        const int ox = u < 0 ? -1 : (u >= S ? 1 : 0), oy = v < 0 ? -1 : (v >= S ? 1 : 0),
                  oz = w < 0 ? -1 : (w >= S ? 1 : 0);
        const int lu = u - ox * S, lv = v - oy * S, lw = w - oz * S;
        const uint32_t nslot = s_nb[(ox + 1) + 3 * (oy + 1) + 9 * (oz + 1)];
        const uint32_t gi = nslot * (uint32_t)Bx::CPB + (uint32_t)(lu + S * (lv + S * lw));
        vi[q] = L.info[gi];
        vx[q] = L.x[gi];
        vb[q] = L.b[gi];
        if (S == 2) {
            // parent: the level-3 cell of the (allocated, in-grid) brick holding this cell
            const uint32_t pi =
                nslot ? (uint32_t)((int)s_nb[27] + ox + D.dim[0] * (oy + D.dim[1] * oz)) : s_nb[27];
            pinf[q] = D.info[pi];
            px[q] = D.x1[pi];
        } else {
            const uint32_t pi = nslot * (uint32_t)CPB2 + (uint32_t)((lu >> 1) + S2 * ((lv >> 1) + S2 * (lw >> 1)));
            pinf[q] = C.info[pi];
            px[q] = C.x1[pi];
        }
    }
    for (int q = 0; q < kPer; ++q) {
        const uint32_t k = t + kPoolThreads * (uint32_t)q;
        if (k >= (uint32_t)Bx::NB) continue;
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
    for (uint32_t k = t; k < (uint32_t)(Bx::IN * Bx::IN * Bx::IN); k += kPoolThreads) {
        const int u = (int)(k % Bx::IN) - 1, v = (int)((k / Bx::IN) % Bx::IN) - 1, w = (int)(k / (Bx::IN * Bx::IN)) - 1;
        const uint32_t kb = Bx::box_of(u, v, w), info = s_info[kb];
        if (!(info & kInfoSelf) || (info & kInfoRed)) continue;
        s_v[kb] = (s_b[kb] + box_sum<S>(s_v, kb, info)) / info_diag(info);
    }
    __syncthreads();
    const uint32_t slot = s_nb[13];
    for (uint32_t k = t; k < (uint32_t)Bx::CPB; k += kPoolThreads) {
        const int u = (int)(k % S), v = (int)((k / S) % S), w = (int)(k / (S * S));
        const uint32_t kb = Bx::box_of(u, v, w), info = s_info[kb];
        if (!(info & kInfoSelf)) continue;
        L.x1[slot * Bx::CPB + k] =
            (info & kInfoRed) ? (s_b[kb] + box_sum<S>(s_v, kb, info)) / info_diag(info) : s_v[kb];
    }
}

template <int S>
__global__ __launch_bounds__(pool_threads<S>()) void k_pdown_lds(const uint32_t *binfo, PLevel L, PLevel C, DenseLevel D,
                                                                 const SolverState *st)
{
    pdown_body<S>(this_blk(), binfo, L, C, D, st);
}

template <int S>
__global__ __launch_bounds__(pool_threads<S>()) void k_pup_lds(const uint32_t *binfo, PLevel L, PLevel C, DenseLevel D,
                                                               const SolverState *st)
{
    pup_body<S>(this_blk(), binfo, L, C, D, st);
}

// The level-1 passes over a batch of tanks (one brick per workgroup, every tank in one launch).
__global__ __launch_bounds__(pool_threads<4>()) void kb_pdown1(const BatchTank *T, const uint32_t *range, uint32_t nt)
{
    Blk G;
    const BatchTank &B = T[batch_tank(range, nt, blockIdx.x, G)];
    pdown_body<4>(G, B.binfo_f, B.L1, B.L2, DenseLevel{}, B.state);
}

__global__ __launch_bounds__(pool_threads<4>()) void kb_pup1(const BatchTank *T, const uint32_t *range, uint32_t nt)
{
    Blk G;
    const BatchTank &B = T[batch_tank(range, nt, blockIdx.x, G)];
    pup_body<4>(G, B.binfo_f, B.L1, B.L2, DenseLevel{}, B.state);
}

uint32_t pass_blocks(const slosh_ctx *c, const PLevel &L) { return blocks_for((uint64_t)c->nactive * L.cpb, 256); }

} // namespace

PLevel plevel_of(const slosh_ctx *c, int l)
{
    PLevel P;
    P.dim[0] = c->lv[l].dim[0];
    P.dim[1] = c->lv[l].dim[1];
    P.dim[2] = c->lv[l].dim[2];
    P.ls = 3 - l;
    P.s = 1 << P.ls;
    P.cpb = (uint32_t)(P.s * P.s * P.s);
    P.type = c->lv[l].type;
    P.info = c->lv[l].info;
    P.x = LVec{c->lv[l].x, c->lv[l].half};
    P.x1 = LVec{c->lv[l].x1, c->lv[l].half};
    P.b = LVec{c->lv[l].b, c->lv[l].half};
    return P;
}

DenseLevel dlevel_of(const slosh_ctx *c, int l)
{
    DenseLevel D;
    D.dim[0] = c->lv[l].dim[0];
    D.dim[1] = c->lv[l].dim[1];
    D.dim[2] = c->lv[l].dim[2];
    D.cells = c->lv[l].cells;
    D.type = c->lv[l].type;
    D.info = c->lv[l].info;
    D.coord = c->lv[l].coord;
    D.x = LVec{c->lv[l].x, c->lv[l].half};
    D.x1 = LVec{c->lv[l].x1, c->lv[l].half};
    D.b = LVec{c->lv[l].b, c->lv[l].half};
    return D;
}

void pooled_build_type(slosh_ctx *c, int l)
{
    PLevel L = plevel_of(c, l);
    SLOSH_TIMED(c, K_MG_SETUP);
    if (l == 1)
        hipLaunchKernelGGL(k_ptype1, dim3(pass_blocks(c, L)), dim3(256), 0, c->stream, c->geo, c->nactive, c->ctype, L,
                           fluid_coarsening(c));
    else
        hipLaunchKernelGGL(k_ptype, dim3(pass_blocks(c, L)), dim3(256), 0, c->stream, c->geo, c->nactive,
                           plevel_of(c, l - 1), L, fluid_coarsening(c));
}

void pooled_build_info(slosh_ctx *c, int l)
{
    PLevel L = plevel_of(c, l);
    launch_timed(c, K_MG_SETUP, k_pinfo, dim3(pass_blocks(c, L)), dim3(256), 0, c->geo, c->nactive, L);
}

void dense_build_type_from_pooled(slosh_ctx *c, int l)
{
    DenseLevel C = dlevel_of(c, l);
    launch_timed(c, K_MG_SETUP, k_dtype_from_pooled, dim3(blocks_for(C.cells, 256)), dim3(256), 0, c->geo,
                 plevel_of(c, l - 1), C, fluid_coarsening(c));
}

void pooled_down(slosh_ctx *c, int l)
{
    PLevel L = plevel_of(c, l);
    PLevel C = l == 1 ? plevel_of(c, 2) : PLevel{};
    DenseLevel D = l == 2 ? dlevel_of(c, 3) : DenseLevel{};
    KTimer timer(c, l == 1 ? K_PDOWN1 : K_PDOWN2);
    if (l == 1)
        hipLaunchKernelGGL(k_pdown_lds<4>, dim3(c->nsolve), dim3(pool_threads<4>()), 0, c->stream, c->binfo_f, L, C, D,
                           c->state);
    else
        hipLaunchKernelGGL(k_pdown_lds<2>, dim3(c->nsolve), dim3(pool_threads<2>()), 0, c->stream, c->binfo_f, L, C, D,
                           c->state);
}

void pooled_up(slosh_ctx *c, int l)
{
    PLevel L = plevel_of(c, l);
    PLevel C = l == 1 ? plevel_of(c, 2) : PLevel{};
    DenseLevel D = l == 2 ? dlevel_of(c, 3) : DenseLevel{};
    KTimer timer(c, l == 1 ? K_PUP1 : K_PUP2);
    if (l == 1)
        hipLaunchKernelGGL(k_pup_lds<4>, dim3(c->nsolve), dim3(pool_threads<4>()), 0, c->stream, c->binfo_f, L, C, D,
                           c->state);
    else
        hipLaunchKernelGGL(k_pup_lds<2>, dim3(c->nsolve), dim3(pool_threads<2>()), 0, c->stream, c->binfo_f, L, C, D,
                           c->state);
}

void b_pooled_down1(const BatchRun &B)
{
    hipLaunchKernelGGL(kb_pdown1, dim3(B.total[kRangeBrick]), dim3(pool_threads<4>()), 0u, B.stream, B.d_tanks,
                       B.range(kRangeBrick), B.ntanks);
}

void b_pooled_up1(const BatchRun &B)
{
    hipLaunchKernelGGL(kb_pup1, dim3(B.total[kRangeBrick]), dim3(pool_threads<4>()), 0u, B.stream, B.d_tanks,
                       B.range(kRangeBrick), B.ntanks);
}

} // namespace slosh
