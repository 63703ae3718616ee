// level0.hip -- the finest-level (brick-pool) passes of the pressure solvers on gfx950.
//
// SloshPressureMG.usf's level-0 kernels (PCGInitCS :284-305, MGSmoothCS :311-338, MGRestrictCS :343-381,
// MGProlongCS :418-433, PCGApplyACS :438-454, PCGUpdatePRCS :495-524, PCGUpdateDCS :527-543) and
// Slosh.usf:856-892 (PressureRBGSCS), restated for the MI355X:
//   - a capped grid of 256-thread workgroups strides over the allocated bricks, so a reduction leaves at most
//     kMaxGroups partials for its finalize kernel;
//   - the stencil reads a per-cell mask of fluid neighbours (written by the divergence pass) instead of six
//     neighbour cell types, and neighbours in another brick are addressed through the six face-neighbour slots of
//     the brick, which are wave-uniform (scalar loads), instead of one brick-table lookup per neighbour;
//   - a red-black sweep maps the 256 threads onto the 256 cells of one colour, so no lane idles.
#include "slosh_ctx.h"

namespace slosh {

namespace {

constexpr uint32_t kThreads = 256;
constexpr uint32_t kMaxGroups = 2048;
constexpr uint32_t kSelfFluid = 1u << 6;

// Wave-uniform description of one allocated brick: its slot base and the slot bases of its six face neighbours
// (0 = null brick, which only masked-out neighbours can point into).
struct Brick {
    int bc[3];
    uint32_t base;
    uint32_t nb[6];
};

__device__ __forceinline__ Brick brick_at(const Geo &g, uint32_t e)
{
    Brick B;
    const uint32_t b = g.active[e];
    brick_coord(g, b, B.bc);
    B.base = g.table[b] * kBrickCells;
    for (int n = 0; n < 6; ++n) {
        int d[3];
        nb_offset(n, d);
        int x = B.bc[0] + d[0], y = B.bc[1] + d[1], z = B.bc[2] + d[2];
        uint32_t s = 0u;
        if (x >= 0 && y >= 0 && z >= 0 && x < g.bd[0] && y < g.bd[1] && z < g.bd[2])
            s = g.table[(uint32_t)(x + g.bd[0] * (y + g.bd[1] * z))];
        B.nb[n] = s * kBrickCells;
    }
    return B;
}

// Pool index of neighbour n (order -x +x -y +y -z +z) of local cell l.
__device__ __forceinline__ uint32_t nbr(const Brick &B, uint32_t l, int n)
{
    switch (n) {
    case 0: return (l & 7u) ? B.base + l - 1u : B.nb[0] + l + 7u;
    case 1: return ((l & 7u) != 7u) ? B.base + l + 1u : B.nb[1] + l - 7u;
    case 2: return (l & 56u) ? B.base + l - 8u : B.nb[2] + l + 56u;
    case 3: return ((l & 56u) != 56u) ? B.base + l + 8u : B.nb[3] + l - 56u;
    case 4: return (l & 448u) ? B.base + l - 64u : B.nb[4] + l + 448u;
    default: return ((l & 448u) != 448u) ? B.base + l + 64u : B.nb[5] + l - 448u;
    }
}

template <typename T, typename S> __device__ __forceinline__ T nbsum(const Brick &B, uint32_t l, uint32_t m, const S *v)
{
    T s = (T)0;
    for (int n = 0; n < 6; ++n)
        if (m >> n & 1u) s += (T)v[nbr(B, l, n)];
    return s;
}

// Local cell of colour `colour` (0 = red = even x+y+z; bricks start at even coordinates) handled by thread t.
__device__ __forceinline__ uint32_t colour_cell(uint32_t t, uint32_t colour)
{
    uint32_t k = t & 3u, y = (t >> 2) & 7u, z = t >> 5;
    uint32_t x = 2u * k + ((y + z + colour) & 1u);
    return x + 8u * y + 64u * z;
}

__device__ __forceinline__ bool red_local(uint32_t l) { return (((l & 7u) + ((l >> 3) & 7u) + (l >> 6)) & 1u) == 0u; }

template <int NV> __device__ void block_partial(const double (&v)[NV], double *partials)
{
    __shared__ double sh[NV][kThreads];
    for (int k = 0; k < NV; ++k) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (uint32_t s = kThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int k = 0; k < NV; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < NV; ++k) partials[(size_t)blockIdx.x * NV + k] = sh[k][0];
}

// r = b - A p (FP64 p), z_red = r / diag, z_black = 0; partials ||b||^2, ||r||^2.
__global__ __launch_bounds__(kThreads) void k_pcg_init(Geo g, uint32_t nactive, const uint8_t *fmask, const float *div,
                                                       const float *diag, const double *pres, float *r, float *z,
                                                       float *d, double *partials)
{
    double v[2] = {0.0, 0.0};
    for (uint32_t e = blockIdx.x; e < nactive; e += gridDim.x) {
        const Brick B = brick_at(g, e);
        for (uint32_t l = threadIdx.x; l < kBrickCells; l += kThreads) {
            const uint32_t i = B.base + l, m = fmask[i];
            if (!(m & kSelfFluid)) {
                r[i] = 0.0f;
                z[i] = 0.0f;
                d[i] = 0.0f;
                continue;
            }
            double s = nbsum<double>(B, l, m, pres);
            float dg = diag[i] + kDiagEps;
            float ri = (float)(-(double)div[i] - ((double)dg * pres[i] - s));
            r[i] = ri;
            z[i] = red_local(l) ? ri / dg : 0.0f;
            v[0] += (double)div[i] * (double)div[i];
            v[1] += (double)ri * (double)ri;
        }
    }
    block_partial<2>(v, partials);
}

// One Gauss-Seidel colour sweep x = (b + sum x_n) / diag (x = z, b = r); with want_dot also the r.z partials.
__global__ __launch_bounds__(kThreads) void k_smooth0(Geo g, uint32_t nactive, const uint8_t *fmask,
                                                      const float *diag, float *x, const float *bvec, uint32_t colour,
                                                      int want_dot, double *partials, const SolverState *st)
{
    if (st->done) return;
    double v[1] = {0.0};
    const uint32_t l = colour_cell(threadIdx.x, colour);
    for (uint32_t e = blockIdx.x; e < nactive; e += gridDim.x) {
        const Brick B = brick_at(g, e);
        const uint32_t i = B.base + l, m = fmask[i];
        if (m & kSelfFluid) {
            float s = nbsum<float>(B, l, m, x);
            float dg = diag[i] + kDiagEps;
            float b = bvec[i];
            float xv = (b + s) / dg;
            x[i] = xv;
            v[0] += (double)b * (double)xv;
        }
        if (want_dot) {
            const uint32_t j = B.base + (l ^ 1u);
            if (fmask[j] & kSelfFluid) v[0] += (double)bvec[j] * (double)x[j];
        }
    }
    if (want_dot) block_partial<1>(v, partials);
}

// Type of level-1 cell (local lx, ly, lz of brick B, possibly one step outside the brick) in the pooled level.
__device__ __forceinline__ uint32_t type1_at(const Brick &B, const PLevel &L1, int lx, int ly, int lz)
{
    const int x = B.bc[0] * 4 + lx, y = B.bc[1] * 4 + ly, z = B.bc[2] * 4 + lz;
    if (x < 0 || y < 0 || z < 0 || x >= L1.dim[0] || y >= L1.dim[1] || z >= L1.dim[2]) return kCellSolid;
    uint32_t base = B.base >> 3; // slot * 64
    if (lx < 0) base = B.nb[0] >> 3, lx += 4;
    else if (lx > 3) base = B.nb[1] >> 3, lx -= 4;
    else if (ly < 0) base = B.nb[2] >> 3, ly += 4;
    else if (ly > 3) base = B.nb[3] >> 3, ly -= 4;
    else if (lz < 0) base = B.nb[4] >> 3, lz += 4;
    else if (lz > 3) base = B.nb[5] >> 3, lz -= 4;
    return L1.type[base + (uint32_t)(lx + 4 * ly + 16 * lz)];
}

// Restriction into pooled level 1: the residuals of the brick's 512 cells in LDS, then b_c = 1/4 of the fluid
// children's sum for the brick's 4^3 level-1 cells, and x_c = b_c / diag_c on red cells (fused red sweep).
__global__ __launch_bounds__(kThreads) void k_restrict0(Geo g, uint32_t nactive, const uint8_t *fmask,
                                                        const float *diag, const float *xf, const float *bf, PLevel L1,
                                                        const SolverState *st)
{
    if (st->done) return;
    __shared__ float res[kBrickCells];
    __shared__ uint8_t fl[kBrickCells];
    for (uint32_t e = blockIdx.x; e < nactive; e += gridDim.x) {
        const Brick B = brick_at(g, e);
        for (uint32_t l = threadIdx.x; l < kBrickCells; l += kThreads) {
            const uint32_t i = B.base + l, m = fmask[i];
            fl[l] = (uint8_t)((m & kSelfFluid) ? 1u : 0u);
            if (m & kSelfFluid) {
                float s = nbsum<float>(B, l, m, xf);
                float dg = diag[i] + kDiagEps;
                res[l] = bf[i] - (dg * xf[i] - s);
            }
        }
        __syncthreads();
        if (threadIdx.x < 64u) {
            const uint32_t t = threadIdx.x;
            const int lx = (int)(t & 3u), ly = (int)((t >> 2) & 3u), lz = (int)(t >> 4);
            const int cx = B.bc[0] * 4 + lx, cy = B.bc[1] * 4 + ly, cz = B.bc[2] * 4 + lz;
            if (cx < L1.dim[0] && cy < L1.dim[1] && cz < L1.dim[2]) {
                const uint32_t ci = (B.base >> 3) + t;
                float *xc = L1.x, *bc = L1.b;
                if (L1.type[ci] != kCellFluid) {
                    bc[ci] = 0.0f;
                    xc[ci] = 0.0f;
                } else {
                    float sum = 0.0f;
                    for (uint32_t oi = 0; oi < 8u; ++oi) {
                        uint32_t fx = 2u * lx + (oi & 1u), fy = 2u * ly + ((oi >> 1) & 1u), fz = 2u * lz + (oi >> 2);
                        uint32_t fl_i = fx + 8u * fy + 64u * fz;
                        if (fl[fl_i]) sum += res[fl_i];
                    }
                    float bv = 0.25f * sum;
                    bc[ci] = bv;
                    float cdg = kDiagEps;
                    for (int nb = 0; nb < 6; ++nb) {
                        int d[3];
                        nb_offset(nb, d);
                        uint32_t ty = type1_at(B, L1, lx + d[0], ly + d[1], lz + d[2]);
                        cdg += ty != kCellSolid ? 1.0f : 0.0f;
                    }
                    xc[ci] = is_red(cx, cy, cz) ? bv / cdg : 0.0f;
                }
            }
        }
        __syncthreads();
    }
}

// Prolongation from pooled level 1 (the parent lives in the same brick): x_f += x_c(parent) on fluid cells with
// a fluid parent.
__global__ __launch_bounds__(kThreads) void k_prolong0(Geo g, uint32_t nactive, const uint8_t *fmask, float *xf,
                                                       PLevel L1, const SolverState *st)
{
    if (st->done) return;
    for (uint32_t e = blockIdx.x; e < nactive; e += gridDim.x) {
        const Brick B = brick_at(g, e);
        for (uint32_t l = threadIdx.x; l < kBrickCells; l += kThreads) {
            const uint32_t i = B.base + l;
            if (!(fmask[i] & kSelfFluid)) continue;
            const uint32_t pi = (B.base >> 3) + ((l & 7u) >> 1) + 4u * (((l >> 3) & 7u) >> 1) + 16u * ((l >> 6) >> 1);
            if (L1.type[pi] == kCellFluid) xf[i] += L1.x[pi];
        }
    }
}

// q = A d with an FP64 stencil sum (for smooth d the diagonal and neighbour terms cancel, and an FP32 sum would
// put a large relative error into q that the recurrence residual never sees); partials d.q.
__global__ __launch_bounds__(kThreads) void k_apply_a(Geo g, uint32_t nactive, const uint8_t *fmask, const float *diag,
                                                      const float *d, float *q, double *partials,
                                                      const SolverState *st)
{
    if (st->done) return;
    double v[1] = {0.0};
    for (uint32_t e = blockIdx.x; e < nactive; e += gridDim.x) {
        const Brick B = brick_at(g, e);
        for (uint32_t l = threadIdx.x; l < kBrickCells; l += kThreads) {
            const uint32_t i = B.base + l, m = fmask[i];
            if (!(m & kSelfFluid)) continue;
            double s = nbsum<double>(B, l, m, d);
            float dg = diag[i] + kDiagEps;
            float dv = d[i];
            double qd = (double)dg * (double)dv - s;
            q[i] = (float)qd;
            v[0] += (double)dv * qd;
        }
    }
    block_partial<1>(v, partials);
}

// p += alpha d (FP64), r -= alpha q, z_red = r / diag (the next V-cycle's fused red sweep); partials ||r||^2.
__global__ __launch_bounds__(kThreads) void k_update_pr(Geo g, uint32_t nactive, const uint8_t *fmask,
                                                        const float *diag, double *pres, float *r, float *z,
                                                        const float *d, const float *q, double *partials,
                                                        const SolverState *st)
{
    if (st->done) return;
    const double alpha = st->alpha;
    double v[1] = {0.0};
    for (uint32_t e = blockIdx.x; e < nactive; e += gridDim.x) {
        const Brick B = brick_at(g, e);
        for (uint32_t l = threadIdx.x; l < kBrickCells; l += kThreads) {
            const uint32_t i = B.base + l;
            if (!(fmask[i] & kSelfFluid)) continue;
            pres[i] = pres[i] + alpha * (double)d[i];
            float ri = (float)((double)r[i] - alpha * (double)q[i]);
            r[i] = ri;
            float dg = diag[i] + kDiagEps;
            z[i] = red_local(l) ? ri / dg : 0.0f;
            v[0] += (double)ri * (double)ri;
        }
    }
    block_partial<1>(v, partials);
}

// d = z + beta d (beta = 0 for the first direction).
__global__ __launch_bounds__(kThreads) void k_update_d(Geo g, uint32_t nactive, const uint8_t *fmask, const float *z,
                                                       float *d, int first, const SolverState *st)
{
    if (st->done) return;
    const double beta = first ? 0.0 : st->beta;
    for (uint32_t e = blockIdx.x; e < nactive; e += gridDim.x) {
        const Brick B = brick_at(g, e);
        for (uint32_t l = threadIdx.x; l < kBrickCells; l += kThreads) {
            const uint32_t i = B.base + l;
            if (!(fmask[i] & kSelfFluid)) continue;
            d[i] = (float)((double)z[i] + beta * (double)d[i]);
        }
    }
}

// Red-black SOR sweep on the FP64 pressure: p += w ((sum p_n - div) / diag - p).
__global__ __launch_bounds__(kThreads) void k_sor_sweep(Geo g, uint32_t nactive, const uint8_t *fmask,
                                                        const float *div, const float *diag, double *pres,
                                                        uint32_t colour, double w, const SolverState *st)
{
    if (st->done) return;
    const uint32_t l = colour_cell(threadIdx.x, colour);
    for (uint32_t e = blockIdx.x; e < nactive; e += gridDim.x) {
        const Brick B = brick_at(g, e);
        const uint32_t i = B.base + l, m = fmask[i];
        if (!(m & kSelfFluid)) continue;
        double s = nbsum<double>(B, l, m, pres);
        double dg = (double)(diag[i] + kDiagEps);
        double gs = (s - (double)div[i]) / dg;
        double old = pres[i];
        pres[i] = old + w * (gs - old);
    }
}

// Partials ||b||^2 and ||b - A p||^2 for the SOR convergence check.
__global__ __launch_bounds__(kThreads) void k_sor_residual(Geo g, uint32_t nactive, const uint8_t *fmask,
                                                           const float *div, const float *diag, const double *pres,
                                                           double *partials, const SolverState *st)
{
    if (st->done) return;
    double v[2] = {0.0, 0.0};
    for (uint32_t e = blockIdx.x; e < nactive; e += gridDim.x) {
        const Brick B = brick_at(g, e);
        for (uint32_t l = threadIdx.x; l < kBrickCells; l += kThreads) {
            const uint32_t i = B.base + l, m = fmask[i];
            if (!(m & kSelfFluid)) continue;
            double s = nbsum<double>(B, l, m, pres);
            double dg = (double)(diag[i] + kDiagEps);
            double ri = -(double)div[i] - (dg * pres[i] - s);
            v[0] += (double)div[i] * (double)div[i];
            v[1] += ri * ri;
        }
    }
    block_partial<2>(v, partials);
}

} // namespace

uint32_t l0_groups(const slosh_ctx *c) { return c->nactive < kMaxGroups ? c->nactive : kMaxGroups; }

#define L0_LAUNCH(kid, kernel, ...)                                                                              \
    do {                                                                                                         \
        SLOSH_TIMED(c, kid);                                                                                     \
        hipLaunchKernelGGL(kernel, dim3(l0_groups(c)), dim3(kThreads), 0, c->stream, c->geo, c->nactive, c->fmask, \
                           __VA_ARGS__);                                                                         \
    } while (0)

void l0_pcg_init(slosh_ctx *c) { L0_LAUNCH(K_PCG_INIT, k_pcg_init, c->div, c->diag, c->pres, c->r, c->z, c->d, c->partials); }

void l0_smooth(slosh_ctx *c, uint32_t colour, int want_dot)
{
    L0_LAUNCH(K_SMOOTH0, k_smooth0, c->diag, c->z, c->r, colour, want_dot, c->partials, c->state);
}

void l0_restrict(slosh_ctx *c) { L0_LAUNCH(K_RESTRICT0, k_restrict0, c->diag, c->z, c->r, plevel_of(c, 1), c->state); }

void l0_prolong(slosh_ctx *c) { L0_LAUNCH(K_PROLONG0, k_prolong0, c->z, plevel_of(c, 1), c->state); }

void l0_apply_a(slosh_ctx *c) { L0_LAUNCH(K_APPLY_A, k_apply_a, c->diag, c->d, c->q, c->partials, c->state); }

void l0_update_pr(slosh_ctx *c)
{
    L0_LAUNCH(K_UPDATE_PR, k_update_pr, c->diag, c->pres, c->r, c->z, c->d, c->q, c->partials, c->state);
}

void l0_update_d(slosh_ctx *c, int first) { L0_LAUNCH(K_UPDATE_D, k_update_d, c->z, c->d, first, c->state); }

void l0_sor_sweep(slosh_ctx *c, uint32_t colour, double w)
{
    L0_LAUNCH(K_SOR_SWEEP, k_sor_sweep, c->div, c->diag, c->pres, colour, w, c->state);
}

void l0_sor_residual(slosh_ctx *c)
{
    L0_LAUNCH(K_SOR_RESIDUAL, k_sor_residual, c->div, c->diag, c->pres, c->partials, c->state);
}

} // namespace slosh
