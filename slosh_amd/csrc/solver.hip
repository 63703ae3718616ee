// solver.hip -- pressure solvers on gfx950: MGPCG (SloshPressureMG.usf) and red-black SOR (Slosh.usf:856-892).
//
// Both solve (CellDiag_c + 1e-4) p_c - sum_{fluid n} p_n = -Divergence_c on the fluid cells of the brick pool.
// MGPCG keeps FP32 vectors, an FP64 pressure (the alpha-update accumulator), FP64 dot products and alpha/beta, and
// solves the coarsest level (<= 64 cells) exactly in FP64. All scalars live in a device SolverState; every solver
// kernel returns at once when `done` is set, so the host enqueues iterations in chunks and reads the state back
// once per chunk. Reductions are deterministic: per-workgroup partials summed in a fixed order by the last
// workgroup to finish.
#include "slosh_ctx.h"

namespace slosh {

namespace {

inline uint32_t blocks_for(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

struct L0 {
    int x, y, z;
    uint32_t idx;
    bool ok;
};

__device__ __forceinline__ L0 l0_cell(const Geo &g)
{
    L0 c;
    const uint32_t b = g.active[blockIdx.x];
    const uint32_t l = threadIdx.x;
    int bc[3];
    brick_coord(g, b, bc);
    c.x = bc[0] * 8 + (int)(l & 7u);
    c.y = bc[1] * 8 + (int)((l >> 3) & 7u);
    c.z = bc[2] * 8 + (int)(l >> 6);
    c.idx = g.table[b] * kBrickCells + l;
    c.ok = in_grid(g, c.x, c.y, c.z);
    return c;
}

__device__ __forceinline__ uint32_t type0(const Geo &g, const uint32_t *ctype, int x, int y, int z)
{
    return in_grid(g, x, y, z) ? (ctype[cell_index(g, x, y, z)] & kCellTypeMask) : kCellSolid;
}

// SloshPressureMG.usf:142-157 (FINE_STENCIL) on level 0 with the ghost-fluid diagonal.
template <typename T>
__device__ __forceinline__ T nbsum0(const Geo &g, const uint32_t *ctype, const T *buf, int x, int y, int z)
{
    T s = (T)0;
    for (int nb = 0; nb < 6; ++nb) {
        int d[3];
        nb_offset(nb, d);
        int nx = x + d[0], ny = y + d[1], nz = z + d[2];
        if (!in_grid(g, nx, ny, nz)) continue;
        uint32_t ni = cell_index(g, nx, ny, nz);
        if ((ctype[ni] & kCellTypeMask) == kCellFluid) s += buf[ni];
    }
    return s;
}

struct Dense {
    int dim[3];
    uint32_t cells;
    const uint8_t *type;
};

__device__ __forceinline__ uint32_t dtype_at(const Dense &L, int x, int y, int z)
{
    if (x < 0 || y < 0 || z < 0 || x >= L.dim[0] || y >= L.dim[1] || z >= L.dim[2]) return kCellSolid;
    return L.type[(uint32_t)(x + L.dim[0] * (y + L.dim[1] * z))];
}

// Stencil of a dense coarse level: diagonal = 1e-4 + non-solid neighbours, sum over fluid neighbours.
__device__ __forceinline__ void dstencil(const Dense &L, const float *buf, int x, int y, int z, float &dg, float &sum)
{
    dg = kDiagEps;
    sum = 0.0f;
    for (int nb = 0; nb < 6; ++nb) {
        int d[3];
        nb_offset(nb, d);
        int nx = x + d[0], ny = y + d[1], nz = z + d[2];
        if (nx < 0 || ny < 0 || nz < 0 || nx >= L.dim[0] || ny >= L.dim[1] || nz >= L.dim[2]) continue;
        uint32_t li = (uint32_t)(nx + L.dim[0] * (ny + L.dim[1] * nz));
        uint32_t t = L.type[li];
        if (t != kCellSolid) dg += 1.0f;
        if (t == kCellFluid) sum += buf[li];
    }
}

// Deterministic grid-wide sum of NV doubles per thread: per-workgroup tree sums into partials, then the last
// workgroup (atomic ticket) sums the partials in a fixed order. Returns true in thread 0 of the last workgroup,
// with the totals in tot.
template <int NV, int BS>
__device__ bool grid_sum(const double (&v)[NV], double *partials, uint32_t *counter, double (&tot)[NV])
{
    __shared__ double sh[NV][BS];
    __shared__ bool last;
    for (int k = 0; k < NV; ++k) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (uint32_t s = BS / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int k = 0; k < NV; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int k = 0; k < NV; ++k) partials[(size_t)blockIdx.x * NV + k] = sh[k][0];
        __threadfence();
        last = atomicAdd(counter, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    double acc[NV];
    for (int k = 0; k < NV; ++k) acc[k] = 0.0;
    for (uint32_t b = threadIdx.x; b < gridDim.x; b += BS)
        for (int k = 0; k < NV; ++k) acc[k] += partials[(size_t)b * NV + k];
    for (int k = 0; k < NV; ++k) sh[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (uint32_t s = BS / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int k = 0; k < NV; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x != 0) return false;
    for (int k = 0; k < NV; ++k) tot[k] = sh[k][0];
    *counter = 0u;
    return true;
}

// ---- level-0 kernels (one 512-thread workgroup per allocated brick) ------------------------------------------

// SloshPressureMG.usf:284-305 (PCGInitCS): r = b - A p (FP64 p), z_red = r / diag, z_black = 0; ||b||^2, ||r||^2.
__global__ __launch_bounds__(512) void k_pcg_init(Geo g, const uint32_t *ctype, const float *div, const float *diag,
                                                  const double *pres, float *r, float *z, float *d, double *partials,
                                                  uint32_t *counter, SolverState *st)
{
    L0 c = l0_cell(g);
    double v[2] = {0.0, 0.0};
    if (c.ok) {
        if ((ctype[c.idx] & kCellTypeMask) != kCellFluid) {
            r[c.idx] = 0.0f;
            z[c.idx] = 0.0f;
            d[c.idx] = 0.0f;
        } else {
            double s = nbsum0<double>(g, ctype, pres, c.x, c.y, c.z);
            float dg = diag[c.idx] + kDiagEps;
            float ri = (float)(-(double)div[c.idx] - ((double)dg * pres[c.idx] - s));
            r[c.idx] = ri;
            z[c.idx] = is_red(c.x, c.y, c.z) ? ri / dg : 0.0f;
            v[0] = (double)div[c.idx] * (double)div[c.idx];
            v[1] = (double)ri * (double)ri;
        }
    }
    double tot[2];
    if (grid_sum<2, 512>(v, partials, counter, tot)) {
        st->bb = tot[0];
        st->rr = tot[1];
        double bn = sqrt(tot[0]);
        st->rel = bn > 0.0 ? sqrt(tot[1]) / bn : 0.0;
        st->iters = 0;
        st->rho = 0.0;
        st->beta = 0.0;
        st->alpha = 0.0;
        st->converged = (bn == 0.0 || st->rel <= st->tol) ? 1 : 0;
        st->done = st->converged;
    }
}

// SloshPressureMG.usf:311-338 (MGSmoothCS) on level 0 (x = z, b = r). With want_dot, also r.z -> rho and beta.
__global__ __launch_bounds__(512) void k_smooth0(Geo g, const uint32_t *ctype, const float *diag, float *x,
                                                 const float *bvec, uint32_t colour, int want_dot, double *partials,
                                                 uint32_t *counter, SolverState *st)
{
    if (st->done) return;
    L0 c = l0_cell(g);
    double v[1] = {0.0};
    if (c.ok && (ctype[c.idx] & kCellTypeMask) == kCellFluid) {
        if ((is_red(c.x, c.y, c.z) ? 0u : 1u) == colour) {
            float s = nbsum0<float>(g, ctype, x, c.x, c.y, c.z);
            float dg = diag[c.idx] + kDiagEps;
            float b = bvec[c.idx];
            float xv = (b + s) / dg;
            x[c.idx] = xv;
            v[0] = (double)b * (double)xv;
        } else if (want_dot) {
            v[0] = (double)bvec[c.idx] * (double)x[c.idx];
        }
    }
    if (!want_dot) return;
    double tot[1];
    if (grid_sum<1, 512>(v, partials, counter, tot)) {
        double rho_new = tot[0];
        st->beta = fabs(st->rho) > 1e-30 ? rho_new / st->rho : 0.0;
        st->rho = rho_new;
    }
}

// SloshPressureMG.usf:343-381 (MGRestrictCS) from level 0 into dense level 1: the 4^3 level-1 cells of each brick.
__global__ __launch_bounds__(64) void k_restrict0(Geo g, const uint32_t *ctype, const float *diag, const float *xf,
                                                  const float *bf, Dense C, float *xc, float *bc, const SolverState *st)
{
    if (st->done) return;
    const uint32_t b = g.active[blockIdx.x];
    int bcd[3];
    brick_coord(g, b, bcd);
    const uint32_t l = threadIdx.x;
    int cx = bcd[0] * 4 + (int)(l & 3u), cy = bcd[1] * 4 + (int)((l >> 2) & 3u), cz = bcd[2] * 4 + (int)(l >> 4);
    if (cx >= C.dim[0] || cy >= C.dim[1] || cz >= C.dim[2]) return;
    uint32_t ci = (uint32_t)(cx + C.dim[0] * (cy + C.dim[1] * cz));
    if (C.type[ci] != kCellFluid) {
        bc[ci] = 0.0f;
        xc[ci] = 0.0f;
        return;
    }
    float sum = 0.0f;
    for (uint32_t oi = 0; oi < 8u; ++oi) {
        int fx = 2 * cx + (int)(oi & 1u), fy = 2 * cy + (int)((oi >> 1) & 1u), fz = 2 * cz + (int)(oi >> 2);
        if (type0(g, ctype, fx, fy, fz) != kCellFluid) continue;
        uint32_t fi = cell_index(g, fx, fy, fz);
        float s = nbsum0<float>(g, ctype, xf, fx, fy, fz);
        float dg = diag[fi] + kDiagEps;
        sum += bf[fi] - (dg * xf[fi] - s);
    }
    float bv = 0.25f * sum;
    bc[ci] = bv;
    float cdg = kDiagEps;
    for (int nb = 0; nb < 6; ++nb) {
        int d[3];
        nb_offset(nb, d);
        cdg += dtype_at(C, cx + d[0], cy + d[1], cz + d[2]) != kCellSolid ? 1.0f : 0.0f;
    }
    xc[ci] = is_red(cx, cy, cz) ? bv / cdg : 0.0f;
}

// SloshPressureMG.usf:418-433 (MGProlongCS) from dense level 1 into level 0.
__global__ __launch_bounds__(512) void k_prolong0(Geo g, const uint32_t *ctype, float *xf, Dense C, const float *xc,
                                                  const SolverState *st)
{
    if (st->done) return;
    L0 c = l0_cell(g);
    if (!c.ok || (ctype[c.idx] & kCellTypeMask) != kCellFluid) return;
    uint32_t pi = (uint32_t)((c.x >> 1) + C.dim[0] * ((c.y >> 1) + C.dim[1] * (c.z >> 1)));
    if (C.type[pi] == kCellFluid) xf[c.idx] += xc[pi];
}

// SloshPressureMG.usf:438-454 (PCGApplyACS): q = A d and d.q -> alpha = rho / d.q.
__global__ __launch_bounds__(512) void k_apply_a(Geo g, const uint32_t *ctype, const float *diag, const float *d,
                                                 float *q, double *partials, uint32_t *counter, SolverState *st)
{
    if (st->done) return;
    L0 c = l0_cell(g);
    double v[1] = {0.0};
    if (c.ok && (ctype[c.idx] & kCellTypeMask) == kCellFluid) {
        // FP64 accumulation: for smooth d the diagonal and neighbour terms cancel, and an FP32 sum would put a
        // large relative error into q that the recurrence residual r -= alpha q never sees.
        double s = 0.0;
        for (int nb = 0; nb < 6; ++nb) {
            int dd[3];
            nb_offset(nb, dd);
            int nx = c.x + dd[0], ny = c.y + dd[1], nz = c.z + dd[2];
            if (!in_grid(g, nx, ny, nz)) continue;
            uint32_t ni = cell_index(g, nx, ny, nz);
            if ((ctype[ni] & kCellTypeMask) == kCellFluid) s += (double)d[ni];
        }
        float dg = diag[c.idx] + kDiagEps;
        float dv = d[c.idx];
        double qd = (double)dg * (double)dv - s;
        q[c.idx] = (float)qd;
        v[0] = (double)dv * qd;
    }
    double tot[1];
    if (grid_sum<1, 512>(v, partials, counter, tot)) {
        st->dq = tot[0];
        st->alpha = fabs(tot[0]) > 1e-30 ? st->rho / tot[0] : 0.0;
    }
}

// SloshPressureMG.usf:495-524 (PCGUpdatePRCS): p += alpha d (FP64), r -= alpha q, ||r||^2, z_red = r / diag.
__global__ __launch_bounds__(512) void k_update_pr(Geo g, const uint32_t *ctype, const float *diag, double *pres,
                                                   float *r, float *z, const float *d, const float *q,
                                                   double *partials, uint32_t *counter, SolverState *st)
{
    if (st->done) return;
    L0 c = l0_cell(g);
    const double alpha = st->alpha;
    double v[1] = {0.0};
    if (c.ok && (ctype[c.idx] & kCellTypeMask) == kCellFluid) {
        pres[c.idx] = pres[c.idx] + alpha * (double)d[c.idx];
        float ri = (float)((double)r[c.idx] - alpha * (double)q[c.idx]);
        r[c.idx] = ri;
        float dg = diag[c.idx] + kDiagEps;
        z[c.idx] = is_red(c.x, c.y, c.z) ? ri / dg : 0.0f;
        v[0] = (double)ri * (double)ri;
    }
    double tot[1];
    if (grid_sum<1, 512>(v, partials, counter, tot)) {
        st->rr = tot[0];
        st->rel = sqrt(tot[0]) / sqrt(st->bb);
        st->iters += 1;
        if (st->rel <= st->tol) {
            st->converged = 1;
            st->done = 1;
        } else if (st->iters >= st->maxit) {
            st->done = 1;
        }
    }
}

// SloshPressureMG.usf:527-543 (PCGUpdateDCS): d = z + beta d (beta = 0 on the first direction).
__global__ __launch_bounds__(512) void k_update_d(Geo g, const uint32_t *ctype, const float *z, float *d, int first,
                                                  const SolverState *st)
{
    if (st->done) return;
    L0 c = l0_cell(g);
    if (!c.ok || (ctype[c.idx] & kCellTypeMask) != kCellFluid) return;
    double beta = first ? 0.0 : st->beta;
    d[c.idx] = (float)((double)z[c.idx] + beta * (double)d[c.idx]);
}

// ---- dense coarse levels ------------------------------------------------------------------------------------
// SloshPressureMG.usf:256-278 (MGBuildTypeCS): AIR if any child is air, else FLUID if any child is fluid.
__global__ void k_build_type1(Geo g, const uint32_t *ctype, Dense C, uint8_t *type)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C.cells) return;
    int x = (int)(i % (uint32_t)C.dim[0]), y = (int)((i / (uint32_t)C.dim[0]) % (uint32_t)C.dim[1]),
        z = (int)(i / (uint32_t)(C.dim[0] * C.dim[1]));
    bool air = false, fluid = false;
    for (uint32_t oi = 0; oi < 8u; ++oi) {
        int fx = 2 * x + (int)(oi & 1u), fy = 2 * y + (int)((oi >> 1) & 1u), fz = 2 * z + (int)(oi >> 2);
        if (!in_grid(g, fx, fy, fz)) continue;
        uint32_t t = ctype[cell_index(g, fx, fy, fz)] & kCellTypeMask;
        air = air || t == kCellAir;
        fluid = fluid || t == kCellFluid;
    }
    type[i] = (uint8_t)(air ? kCellAir : (fluid ? kCellFluid : kCellSolid));
}

__global__ void k_build_type(Dense F, Dense C, uint8_t *type)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C.cells) return;
    int x = (int)(i % (uint32_t)C.dim[0]), y = (int)((i / (uint32_t)C.dim[0]) % (uint32_t)C.dim[1]),
        z = (int)(i / (uint32_t)(C.dim[0] * C.dim[1]));
    bool air = false, fluid = false;
    for (uint32_t oi = 0; oi < 8u; ++oi) {
        int fx = 2 * x + (int)(oi & 1u), fy = 2 * y + (int)((oi >> 1) & 1u), fz = 2 * z + (int)(oi >> 2);
        if (fx >= F.dim[0] || fy >= F.dim[1] || fz >= F.dim[2]) continue;
        uint32_t t = F.type[(uint32_t)(fx + F.dim[0] * (fy + F.dim[1] * fz))];
        air = air || t == kCellAir;
        fluid = fluid || t == kCellFluid;
    }
    type[i] = (uint8_t)(air ? kCellAir : (fluid ? kCellFluid : kCellSolid));
}

__global__ void k_smooth(Dense L, float *x, const float *b, uint32_t colour, const SolverState *st)
{
    if (st->done) return;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L.cells || L.type[i] != kCellFluid) return;
    int cx = (int)(i % (uint32_t)L.dim[0]), cy = (int)((i / (uint32_t)L.dim[0]) % (uint32_t)L.dim[1]),
        cz = (int)(i / (uint32_t)(L.dim[0] * L.dim[1]));
    if ((is_red(cx, cy, cz) ? 0u : 1u) != colour) return;
    float dg, s;
    dstencil(L, x, cx, cy, cz, dg, s);
    x[i] = (b[i] + s) / dg;
}

__global__ void k_restrict(Dense F, const float *xf, const float *bf, Dense C, float *xc, float *bc,
                           const SolverState *st)
{
    if (st->done) return;
    uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x;
    if (ci >= C.cells) return;
    int cx = (int)(ci % (uint32_t)C.dim[0]), cy = (int)((ci / (uint32_t)C.dim[0]) % (uint32_t)C.dim[1]),
        cz = (int)(ci / (uint32_t)(C.dim[0] * C.dim[1]));
    if (C.type[ci] != kCellFluid) {
        bc[ci] = 0.0f;
        xc[ci] = 0.0f;
        return;
    }
    float sum = 0.0f;
    for (uint32_t oi = 0; oi < 8u; ++oi) {
        int fx = 2 * cx + (int)(oi & 1u), fy = 2 * cy + (int)((oi >> 1) & 1u), fz = 2 * cz + (int)(oi >> 2);
        if (dtype_at(F, fx, fy, fz) != kCellFluid) continue;
        uint32_t fi = (uint32_t)(fx + F.dim[0] * (fy + F.dim[1] * fz));
        float dg, s;
        dstencil(F, xf, fx, fy, fz, dg, s);
        sum += bf[fi] - (dg * xf[fi] - s);
    }
    float bv = 0.25f * sum;
    bc[ci] = bv;
    float cdg = kDiagEps;
    for (int nb = 0; nb < 6; ++nb) {
        int d[3];
        nb_offset(nb, d);
        cdg += dtype_at(C, cx + d[0], cy + d[1], cz + d[2]) != kCellSolid ? 1.0f : 0.0f;
    }
    xc[ci] = is_red(cx, cy, cz) ? bv / cdg : 0.0f;
}

__global__ void k_prolong(Dense F, float *xf, Dense C, const float *xc, const SolverState *st)
{
    if (st->done) return;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= F.cells || F.type[i] != kCellFluid) return;
    int x = (int)(i % (uint32_t)F.dim[0]), y = (int)((i / (uint32_t)F.dim[0]) % (uint32_t)F.dim[1]),
        z = (int)(i / (uint32_t)(F.dim[0] * F.dim[1]));
    uint32_t pi = (uint32_t)((x >> 1) + C.dim[0] * ((y >> 1) + C.dim[1] * (z >> 1)));
    if (C.type[pi] == kCellFluid) xf[i] += xc[pi];
}

// ---- coarsest level: FP64 Cholesky of the <= 64 fluid unknowns and its explicit inverse (once per solve) ----
// Level 0 is the coarsest only for grids of <= 64 cells (one brick); then the pool diagonal is used.
__global__ __launch_bounds__(256) void k_coarse_factor(Geo g, const uint32_t *ctype, const float *diag, Dense C,
                                                       int level0, uint32_t *cidx, uint32_t *cn, double *inv)
{
    constexpr int N = (int)kCoarsestCells;
    __shared__ double a[N][N + 1];
    __shared__ int px[N], py[N], pz[N];
    __shared__ uint32_t pidx[N];
    __shared__ uint32_t n_s;
    const int t = threadIdx.x;
    const int dims[3] = {level0 ? g.n[0] : C.dim[0], level0 ? g.n[1] : C.dim[1], level0 ? g.n[2] : C.dim[2]};
    const uint32_t cells = (uint32_t)(dims[0] * dims[1] * dims[2]);
    auto type_of = [&](int x, int y, int z) -> uint32_t {
        if (x < 0 || y < 0 || z < 0 || x >= dims[0] || y >= dims[1] || z >= dims[2]) return kCellSolid;
        if (level0) return ctype[cell_index(g, x, y, z)] & kCellTypeMask;
        return C.type[(uint32_t)(x + dims[0] * (y + dims[1] * z))];
    };
    if (t == 0) {
        uint32_t n = 0;
        for (uint32_t i = 0; i < cells && n < (uint32_t)N; ++i) {
            int x = (int)(i % (uint32_t)dims[0]), y = (int)((i / (uint32_t)dims[0]) % (uint32_t)dims[1]),
                z = (int)(i / (uint32_t)(dims[0] * dims[1]));
            if (type_of(x, y, z) != kCellFluid) continue;
            px[n] = x;
            py[n] = y;
            pz[n] = z;
            pidx[n] = level0 ? cell_index(g, x, y, z) : i;
            n++;
        }
        n_s = n;
        *cn = n;
    }
    __syncthreads();
    const int n = (int)n_s;
    for (int e = t; e < N * N; e += blockDim.x) a[e / N][e % N] = 0.0;
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
        if (level0) dg = diag[pidx[t]] + kDiagEps;
        a[t][t] = (double)dg;
    }
    __syncthreads();
    // left-looking Cholesky, lower factor in place
    for (int j = 0; j < n; ++j) {
        if (t == 0) {
            double s = a[j][j];
            for (int k = 0; k < j; ++k) s -= a[j][k] * a[j][k];
            a[j][j] = sqrt(s > 0.0 ? s : 1e-300);
        }
        __syncthreads();
        if (t > j && t < n) {
            double s = a[t][j];
            for (int k = 0; k < j; ++k) s -= a[t][k] * a[j][k];
            a[t][j] = s / a[j][j];
        }
        __syncthreads();
    }
    // column t of the inverse: solve L L^T x = e_t
    if (t < n) {
        double y[N];
        for (int i = 0; i < n; ++i) {
            double s = i == t ? 1.0 : 0.0;
            for (int k = 0; k < i; ++k) s -= a[i][k] * y[k];
            y[i] = s / a[i][i];
        }
        for (int i = n - 1; i >= 0; --i) {
            double s = y[i];
            for (int k = i + 1; k < n; ++k) s -= a[k][i] * y[k];
            y[i] = s / a[i][i];
        }
        for (int i = 0; i < n; ++i) inv[(size_t)i * N + t] = y[i];
    }
}

// Exact coarsest solve x = A^-1 b in FP64 (replaces MGCoarseSolveCS, SloshPressureMG.usf:386-413).
__global__ __launch_bounds__(64) void k_coarse_solve(const uint32_t *cidx, const uint32_t *cn, const double *inv,
                                                     const float *b, float *x, const SolverState *st)
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

// ---- red-black SOR on the FP64 pressure -------------------------------------------------------------------
__global__ __launch_bounds__(512) void k_sor_sweep(Geo g, const uint32_t *ctype, const float *div, const float *diag,
                                                   double *pres, uint32_t colour, double w, const SolverState *st)
{
    if (st->done) return;
    L0 c = l0_cell(g);
    if (!c.ok || (ctype[c.idx] & kCellTypeMask) != kCellFluid) return;
    if ((((unsigned)(c.x + c.y + c.z)) & 1u) != colour) return;
    double s = nbsum0<double>(g, ctype, pres, c.x, c.y, c.z);
    double dg = (double)(diag[c.idx] + kDiagEps);
    double gs = (s - (double)div[c.idx]) / dg;
    double old = pres[c.idx];
    pres[c.idx] = old + w * (gs - old);
}

__global__ __launch_bounds__(512) void k_sor_residual(Geo g, const uint32_t *ctype, const float *div,
                                                      const float *diag, const double *pres, int iters,
                                                      double *partials, uint32_t *counter, SolverState *st)
{
    if (st->done) return;
    L0 c = l0_cell(g);
    double v[2] = {0.0, 0.0};
    if (c.ok && (ctype[c.idx] & kCellTypeMask) == kCellFluid) {
        double s = nbsum0<double>(g, ctype, pres, c.x, c.y, c.z);
        double dg = (double)(diag[c.idx] + kDiagEps);
        double ri = -(double)div[c.idx] - (dg * pres[c.idx] - s);
        v[0] = (double)div[c.idx] * (double)div[c.idx];
        v[1] = ri * ri;
    }
    double tot[2];
    if (grid_sum<2, 512>(v, partials, counter, tot)) {
        st->bb = tot[0];
        st->rr = tot[1];
        double bn = sqrt(tot[0]);
        st->rel = bn > 0.0 ? sqrt(tot[1]) / bn : 0.0;
        st->iters = iters;
        if (bn == 0.0 || st->rel <= st->tol) {
            st->converged = 1;
            st->done = 1;
        } else if (iters >= st->maxit) {
            st->done = 1;
        }
    }
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

Dense dense_of(const slosh_ctx *c, int l)
{
    Dense D;
    D.dim[0] = c->lv[l].dim[0];
    D.dim[1] = c->lv[l].dim[1];
    D.dim[2] = c->lv[l].dim[2];
    D.cells = c->lv[l].cells;
    D.type = c->lv[l].type;
    return D;
}

void smooth_level(slosh_ctx *c, int l, uint32_t colour, int want_dot)
{
    if (l == 0) {
        {
            SLOSH_TIMED(c, K_SMOOTH0);
            hipLaunchKernelGGL(k_smooth0, dim3(c->nactive), dim3(512), 0, c->stream, c->geo, c->ctype, c->diag, c->z, c->r,
                               colour, want_dot, c->partials, c->red_counter, c->state);
        }
    } else {
        Dense L = dense_of(c, l);
        {
            SLOSH_TIMED(c, K_SMOOTH);
            hipLaunchKernelGGL(k_smooth, dim3(blocks_for(L.cells, 256)), dim3(256), 0, c->stream, L, c->lv[l].x,
                               c->lv[l].b, colour, c->state);
        }
    }
}

void coarse_solve(slosh_ctx *c, int l)
{
    float *x = l == 0 ? c->z : c->lv[l].x;
    const float *b = l == 0 ? c->r : c->lv[l].b;
    {
        SLOSH_TIMED(c, K_COARSE_SOLVE);
        hipLaunchKernelGGL(k_coarse_solve, dim3(1), dim3(64), 0, c->stream, c->coarse_idx, c->coarse_n, c->coarse_inv, b, x,
                           c->state);
    }
}

// SloshProxy.cpp:904-959 (VCycle) with the exact coarsest solve. The level-0 red pre-sweep is already in z.
void vcycle(slosh_ctx *c)
{
    const int m = c->nlevels;
    if (m == 1) {
        coarse_solve(c, 0);
        smooth_level(c, 0, 0u, 1);
        return;
    }
    smooth_level(c, 0, 1u, 0);
    for (int l = 0; l < m - 1; ++l) {
        Dense C = dense_of(c, l + 1);
        if (l == 0) {
            {
                SLOSH_TIMED(c, K_RESTRICT0);
                hipLaunchKernelGGL(k_restrict0, dim3(c->nactive), dim3(64), 0, c->stream, c->geo, c->ctype, c->diag, c->z,
                                   c->r, C, c->lv[1].x, c->lv[1].b, c->state);
            }
        } else {
            Dense F = dense_of(c, l);
            {
                SLOSH_TIMED(c, K_RESTRICT);
                hipLaunchKernelGGL(k_restrict, dim3(blocks_for(C.cells, 256)), dim3(256), 0, c->stream, F, c->lv[l].x,
                                   c->lv[l].b, C, c->lv[l + 1].x, c->lv[l + 1].b, c->state);
            }
        }
        if (l + 1 < m - 1) smooth_level(c, l + 1, 1u, 0);
    }
    coarse_solve(c, m - 1);
    for (int l = m - 2; l >= 0; --l) {
        Dense C = dense_of(c, l + 1);
        if (l == 0) {
            {
                SLOSH_TIMED(c, K_PROLONG0);
                hipLaunchKernelGGL(k_prolong0, dim3(c->nactive), dim3(512), 0, c->stream, c->geo, c->ctype, c->z, C,
                                   c->lv[1].x, c->state);
            }
        } else {
            Dense F = dense_of(c, l);
            {
                SLOSH_TIMED(c, K_PROLONG);
                hipLaunchKernelGGL(k_prolong, dim3(blocks_for(F.cells, 256)), dim3(256), 0, c->stream, F, c->lv[l].x, C,
                                   c->lv[l + 1].x, c->state);
            }
        }
        smooth_level(c, l, 1u, 0);
        smooth_level(c, l, 0u, l == 0 ? 1 : 0);
    }
}

} // namespace

// SloshProxy.cpp:807-1020 (AddPressureSolveMGPCG), to ||r|| <= tol ||b||.
void run_mgpcg(slosh_ctx *c)
{
    if (!c->nactive) return;
    const int maxit = c->s.max_iterations > 0 ? c->s.max_iterations : 200;
    hipLaunchKernelGGL(k_state_reset, dim3(1), dim3(1), 0, c->stream, c->state, (double)c->s.tolerance, maxit);
    const int m = c->nlevels;
    for (int l = 1; l < m; ++l) {
        Dense C = dense_of(c, l);
        if (l == 1) {
            {
                SLOSH_TIMED(c, K_MG_TYPES);
                hipLaunchKernelGGL(k_build_type1, dim3(blocks_for(C.cells, 256)), dim3(256), 0, c->stream, c->geo, c->ctype,
                                   C, c->lv[1].type);
            }
        } else {
            Dense F = dense_of(c, l - 1);
            {
                SLOSH_TIMED(c, K_MG_TYPES);
                hipLaunchKernelGGL(k_build_type, dim3(blocks_for(C.cells, 256)), dim3(256), 0, c->stream, F, C,
                                   c->lv[l].type);
            }
        }
    }
    {
        Dense C = dense_of(c, m - 1);
        {
            SLOSH_TIMED(c, K_COARSE_FACTOR);
            hipLaunchKernelGGL(k_coarse_factor, dim3(1), dim3(256), 0, c->stream, c->geo, c->ctype, c->diag, C,
                               m == 1 ? 1 : 0, c->coarse_idx, c->coarse_n, c->coarse_inv);
        }
    }
    {
        SLOSH_TIMED(c, K_PCG_INIT);
        hipLaunchKernelGGL(k_pcg_init, dim3(c->nactive), dim3(512), 0, c->stream, c->geo, c->ctype, c->div, c->diag,
                           c->pres, c->r, c->z, c->d, c->partials, c->red_counter, c->state);
    }
    vcycle(c);
    {
        SLOSH_TIMED(c, K_UPDATE_D);
        hipLaunchKernelGGL(k_update_d, dim3(c->nactive), dim3(512), 0, c->stream, c->geo, c->ctype, c->z, c->d, 1, c->state);
    }
    SLOSH_HIP(hipGetLastError());
    const int chunk = 4;
    for (int it = 0; it <= maxit; it += chunk) {
        for (int k = 0; k < chunk; ++k) {
            {
                SLOSH_TIMED(c, K_APPLY_A);
                hipLaunchKernelGGL(k_apply_a, dim3(c->nactive), dim3(512), 0, c->stream, c->geo, c->ctype, c->diag, c->d,
                                   c->q, c->partials, c->red_counter, c->state);
            }
            {
                SLOSH_TIMED(c, K_UPDATE_PR);
                hipLaunchKernelGGL(k_update_pr, dim3(c->nactive), dim3(512), 0, c->stream, c->geo, c->ctype, c->diag,
                                   c->pres, c->r, c->z, c->d, c->q, c->partials, c->red_counter, c->state);
            }
            vcycle(c);
            {
                SLOSH_TIMED(c, K_UPDATE_D);
                hipLaunchKernelGGL(k_update_d, dim3(c->nactive), dim3(512), 0, c->stream, c->geo, c->ctype, c->z, c->d, 0,
                                   c->state);
            }
        }
        SLOSH_HIP(hipGetLastError());
        SLOSH_HIP(hipMemcpyAsync(c->h_state, c->state, sizeof(SolverState), hipMemcpyDeviceToHost, c->stream));
        SLOSH_HIP(hipStreamSynchronize(c->stream));
        if (c->h_state->done) break;
    }
}

// Slosh.usf:856-892 (PressureRBGSCS) iterated to the tolerance, residual checked every 8 iterations.
void run_rbsor(slosh_ctx *c)
{
    if (!c->nactive) return;
    const int maxit = c->s.max_iterations > 0 ? c->s.max_iterations : 1000;
    hipLaunchKernelGGL(k_state_reset, dim3(1), dim3(1), 0, c->stream, c->state, (double)c->s.tolerance, maxit);
    const double w = (double)c->s.sor_weight;
    int it = 0;
    {
        SLOSH_TIMED(c, K_SOR_RESIDUAL);
        hipLaunchKernelGGL(k_sor_residual, dim3(c->nactive), dim3(512), 0, c->stream, c->geo, c->ctype, c->div, c->diag,
                           c->pres, it, c->partials, c->red_counter, c->state);
    }
    while (it < maxit) {
        for (int chunk = 0; chunk < 16 && it < maxit; ++chunk) {
            for (int k = 0; k < 8 && it < maxit; ++k, ++it) {
                for (uint32_t colour = 0; colour < 2u; ++colour)
                    {
                        SLOSH_TIMED(c, K_SOR_SWEEP);
                        hipLaunchKernelGGL(k_sor_sweep, dim3(c->nactive), dim3(512), 0, c->stream, c->geo, c->ctype, c->div,
                                           c->diag, c->pres, colour, w, c->state);
                    }
            }
            {
                SLOSH_TIMED(c, K_SOR_RESIDUAL);
                hipLaunchKernelGGL(k_sor_residual, dim3(c->nactive), dim3(512), 0, c->stream, c->geo, c->ctype, c->div,
                                   c->diag, c->pres, it, c->partials, c->red_counter, c->state);
            }
        }
        SLOSH_HIP(hipGetLastError());
        SLOSH_HIP(hipMemcpyAsync(c->h_state, c->state, sizeof(SolverState), hipMemcpyDeviceToHost, c->stream));
        SLOSH_HIP(hipStreamSynchronize(c->stream));
        if (c->h_state->done) break;
    }
}

} // namespace slosh
