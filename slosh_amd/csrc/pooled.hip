// pooled.hip -- multigrid levels 1..3 stored in the brick pool, on gfx950.
//
// The reference keeps every coarse level as a dense grid (SloshProxy.cpp:269-285), so level 1 of a 256^3 domain is
// a 2M-cell array swept in full although the liquid occupies a thin layer. Here a level-l cell (l = 1, 2, 3) lives
// with the 8^3 brick that contains its children: (8 >> l)^3 cells per allocated brick, at slot * cells_per_brick +
// local. Unallocated bricks read as the zero null slot, i.e. AIR, which is what the reference's coarsening rule
// gives for an all-air block. Cells beyond the level's extent (partial bricks) are typed SOLID and never touched.
// Operations are those of SloshPressureMG.usf (MGBuildTypeCS :256-278, MGSmoothCS :311-338, MGRestrictCS
// :343-381, MGProlongCS :418-433) on the counted stencil with the 1e-4 diagonal term.
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

__device__ __forceinline__ uint32_t ptype_at(const Geo &g, const PLevel &L, int x, int y, int z)
{
    return pinside(L, x, y, z) ? L.type[pindex(g, L, x, y, z)] : kCellSolid;
}

// Counted stencil at (x, y, z): diagonal 1e-4 + non-solid neighbours, sum of v over fluid neighbours.
__device__ __forceinline__ void pstencil(const Geo &g, const PLevel &L, const float *v, const PCell &c, float &dg,
                                         float &sum)
{
    dg = kDiagEps;
    sum = 0.0f;
    for (int nb = 0; nb < 6; ++nb) {
        int d[3];
        nb_offset(nb, d);
        int nx = c.x + d[0], ny = c.y + d[1], nz = c.z + d[2];
        if (!pinside(L, nx, ny, nz)) continue;
        int lx = c.lx + d[0], ly = c.ly + d[1], lz = c.lz + d[2];
        uint32_t ni;
        if (lx >= 0 && ly >= 0 && lz >= 0 && lx < L.s && ly < L.s && lz < L.s)
            ni = c.slot * L.cpb + (uint32_t)(lx + L.s * (ly + L.s * lz));
        else
            ni = pindex(g, L, nx, ny, nz);
        uint32_t t = L.type[ni];
        if (t != kCellSolid) dg += 1.0f;
        if (t == kCellFluid) sum += v[ni];
    }
}

// Coarse types of pooled level 1 from level 0 (children in the same brick).
__global__ void k_ptype1(Geo g, uint32_t nactive, const uint32_t *ctype, PLevel L)
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
    L.type[c.idx] = (uint8_t)(air ? kCellAir : (fluid ? kCellFluid : kCellSolid));
}

// Coarse types of pooled level C from pooled level F = C - 1 (children in the same brick).
__global__ void k_ptype(Geo g, uint32_t nactive, PLevel F, PLevel C)
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
    C.type[c.idx] = (uint8_t)(air ? kCellAir : (fluid ? kCellFluid : kCellSolid));
}

__global__ void k_psmooth(Geo g, uint32_t nactive, PLevel L, uint32_t colour, const SolverState *st)
{
    if (st->done) return;
    PCell c;
    if (!pcell(g, L, nactive, blockIdx.x * blockDim.x + threadIdx.x, c)) return;
    if (!c.inside || L.type[c.idx] != kCellFluid || (is_red(c.x, c.y, c.z) ? 0u : 1u) != colour) return;
    float dg, s;
    pstencil(g, L, L.x, c, dg, s);
    L.x[c.idx] = (L.b[c.idx] + s) / dg;
}

// Restriction pooled F -> pooled C (children in the same brick), with the fused red sweep x_c = b_c / diag_c.
__global__ void k_prestrict(Geo g, uint32_t nactive, PLevel F, PLevel C, const SolverState *st)
{
    if (st->done) return;
    PCell c;
    if (!pcell(g, C, nactive, blockIdx.x * blockDim.x + threadIdx.x, c)) return;
    if (!c.inside) return;
    if (C.type[c.idx] != kCellFluid) {
        C.b[c.idx] = 0.0f;
        C.x[c.idx] = 0.0f;
        return;
    }
    float sum = 0.0f;
    for (uint32_t oi = 0; oi < 8u; ++oi) {
        int ox = (int)(oi & 1u), oy = (int)((oi >> 1) & 1u), oz = (int)(oi >> 2);
        PCell f;
        f.x = 2 * c.x + ox;
        f.y = 2 * c.y + oy;
        f.z = 2 * c.z + oz;
        if (!pinside(F, f.x, f.y, f.z)) continue;
        f.lx = 2 * c.lx + ox;
        f.ly = 2 * c.ly + oy;
        f.lz = 2 * c.lz + oz;
        f.slot = c.slot;
        f.idx = c.slot * F.cpb + (uint32_t)(f.lx + F.s * (f.ly + F.s * f.lz));
        if (F.type[f.idx] != kCellFluid) continue;
        float dg, s;
        pstencil(g, F, F.x, f, dg, s);
        sum += F.b[f.idx] - (dg * F.x[f.idx] - s);
    }
    float bv = 0.25f * sum;
    C.b[c.idx] = bv;
    float cdg = kDiagEps;
    for (int nb = 0; nb < 6; ++nb) {
        int d[3];
        nb_offset(nb, d);
        cdg += ptype_at(g, C, c.x + d[0], c.y + d[1], c.z + d[2]) != kCellSolid ? 1.0f : 0.0f;
    }
    C.x[c.idx] = is_red(c.x, c.y, c.z) ? bv / cdg : 0.0f;
}

// Prolongation pooled C -> pooled F: x_f += x_c(parent) (parent in the same brick).
__global__ void k_pprolong(Geo g, uint32_t nactive, PLevel F, PLevel C, const SolverState *st)
{
    if (st->done) return;
    PCell f;
    if (!pcell(g, F, nactive, blockIdx.x * blockDim.x + threadIdx.x, f)) return;
    if (!f.inside || F.type[f.idx] != kCellFluid) return;
    uint32_t pl = (uint32_t)((f.lx >> 1) + C.s * ((f.ly >> 1) + C.s * (f.lz >> 1)));
    uint32_t pi = f.slot * C.cpb + pl;
    if (C.type[pi] == kCellFluid) F.x[f.idx] += C.x[pi];
}

// Dense level C (16^3 and coarser) from pooled level F: coarse types, restriction and prolongation.
__global__ void k_dtype_from_pooled(Geo g, PLevel F, DenseLevel C)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C.cells) return;
    int x = (int)(i % (uint32_t)C.dim[0]), y = (int)((i / (uint32_t)C.dim[0]) % (uint32_t)C.dim[1]),
        z = (int)(i / (uint32_t)(C.dim[0] * C.dim[1]));
    bool air = false, fluid = false;
    for (uint32_t oi = 0; oi < 8u; ++oi) {
        int fx = 2 * x + (int)(oi & 1u), fy = 2 * y + (int)((oi >> 1) & 1u), fz = 2 * z + (int)(oi >> 2);
        if (!pinside(F, fx, fy, fz)) continue;
        uint32_t t = F.type[pindex(g, F, fx, fy, fz)];
        air = air || t == kCellAir;
        fluid = fluid || t == kCellFluid;
    }
    C.type[i] = (uint8_t)(air ? kCellAir : (fluid ? kCellFluid : kCellSolid));
}

__global__ void k_restrict_pd(Geo g, PLevel F, DenseLevel C, const SolverState *st)
{
    if (st->done) return;
    uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x;
    if (ci >= C.cells) return;
    int cx = (int)(ci % (uint32_t)C.dim[0]), cy = (int)((ci / (uint32_t)C.dim[0]) % (uint32_t)C.dim[1]),
        cz = (int)(ci / (uint32_t)(C.dim[0] * C.dim[1]));
    if (C.type[ci] != kCellFluid) {
        C.b[ci] = 0.0f;
        C.x[ci] = 0.0f;
        return;
    }
    float sum = 0.0f;
    for (uint32_t oi = 0; oi < 8u; ++oi) {
        PCell f;
        f.x = 2 * cx + (int)(oi & 1u);
        f.y = 2 * cy + (int)((oi >> 1) & 1u);
        f.z = 2 * cz + (int)(oi >> 2);
        if (!pinside(F, f.x, f.y, f.z)) continue;
        const int m = F.s - 1;
        f.lx = f.x & m;
        f.ly = f.y & m;
        f.lz = f.z & m;
        f.idx = pindex(g, F, f.x, f.y, f.z);
        f.slot = f.idx / F.cpb;
        if (F.type[f.idx] != kCellFluid) continue;
        float dg, s;
        pstencil(g, F, F.x, f, dg, s);
        sum += F.b[f.idx] - (dg * F.x[f.idx] - s);
    }
    float bv = 0.25f * sum;
    C.b[ci] = bv;
    float cdg = kDiagEps;
    for (int nb = 0; nb < 6; ++nb) {
        int d[3];
        nb_offset(nb, d);
        int x = cx + d[0], y = cy + d[1], z = cz + d[2];
        uint32_t t = (x < 0 || y < 0 || z < 0 || x >= C.dim[0] || y >= C.dim[1] || z >= C.dim[2])
                         ? kCellSolid
                         : C.type[(uint32_t)(x + C.dim[0] * (y + C.dim[1] * z))];
        cdg += t != kCellSolid ? 1.0f : 0.0f;
    }
    C.x[ci] = is_red(cx, cy, cz) ? bv / cdg : 0.0f;
}

__global__ void k_prolong_dp(Geo g, uint32_t nactive, PLevel F, DenseLevel C, const SolverState *st)
{
    if (st->done) return;
    PCell f;
    if (!pcell(g, F, nactive, blockIdx.x * blockDim.x + threadIdx.x, f)) return;
    if (!f.inside || F.type[f.idx] != kCellFluid) return;
    uint32_t pi = (uint32_t)((f.x >> 1) + C.dim[0] * ((f.y >> 1) + C.dim[1] * (f.z >> 1)));
    if (C.type[pi] == kCellFluid) F.x[f.idx] += C.x[pi];
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
    P.x = c->lv[l].x;
    P.b = c->lv[l].b;
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
    D.x = c->lv[l].x;
    D.b = c->lv[l].b;
    return D;
}

void pooled_build_type(slosh_ctx *c, int l)
{
    PLevel L = plevel_of(c, l);
    SLOSH_TIMED(c, K_MG_TYPES);
    if (l == 1)
        hipLaunchKernelGGL(k_ptype1, dim3(pass_blocks(c, L)), dim3(256), 0, c->stream, c->geo, c->nactive, c->ctype, L);
    else
        hipLaunchKernelGGL(k_ptype, dim3(pass_blocks(c, L)), dim3(256), 0, c->stream, c->geo, c->nactive,
                           plevel_of(c, l - 1), L);
}

void pooled_smooth(slosh_ctx *c, int l, uint32_t colour)
{
    PLevel L = plevel_of(c, l);
    SLOSH_TIMED(c, K_SMOOTH);
    hipLaunchKernelGGL(k_psmooth, dim3(pass_blocks(c, L)), dim3(256), 0, c->stream, c->geo, c->nactive, L, colour,
                       c->state);
}

void pooled_restrict(slosh_ctx *c, int l)
{
    PLevel C = plevel_of(c, l + 1);
    SLOSH_TIMED(c, K_RESTRICT);
    hipLaunchKernelGGL(k_prestrict, dim3(pass_blocks(c, C)), dim3(256), 0, c->stream, c->geo, c->nactive,
                       plevel_of(c, l), C, c->state);
}

void pooled_prolong(slosh_ctx *c, int l)
{
    PLevel F = plevel_of(c, l);
    SLOSH_TIMED(c, K_PROLONG);
    hipLaunchKernelGGL(k_pprolong, dim3(pass_blocks(c, F)), dim3(256), 0, c->stream, c->geo, c->nactive, F,
                       plevel_of(c, l + 1), c->state);
}

void dense_build_type_from_pooled(slosh_ctx *c, int l)
{
    DenseLevel C = dlevel_of(c, l);
    SLOSH_TIMED(c, K_MG_TYPES);
    hipLaunchKernelGGL(k_dtype_from_pooled, dim3(blocks_for(C.cells, 256)), dim3(256), 0, c->stream, c->geo,
                       plevel_of(c, l - 1), C);
}

void restrict_pooled_to_dense(slosh_ctx *c, int l)
{
    DenseLevel C = dlevel_of(c, l + 1);
    SLOSH_TIMED(c, K_RESTRICT);
    hipLaunchKernelGGL(k_restrict_pd, dim3(blocks_for(C.cells, 256)), dim3(256), 0, c->stream, c->geo,
                       plevel_of(c, l), C, c->state);
}

void prolong_dense_to_pooled(slosh_ctx *c, int l)
{
    PLevel F = plevel_of(c, l);
    SLOSH_TIMED(c, K_PROLONG);
    hipLaunchKernelGGL(k_prolong_dp, dim3(pass_blocks(c, F)), dim3(256), 0, c->stream, c->geo, c->nactive, F,
                       dlevel_of(c, l + 1), c->state);
}

} // namespace slosh
