// solver.hip -- pressure solvers on gfx950: MGPCG (SloshPressureMG.usf) and red-black SOR (Slosh.usf:856-892).
//
// Both solve (CellDiag_c + 1e-4) p_c - sum_{fluid n} p_n = -Divergence_c on the fluid cells of the brick pool.
// MGPCG keeps FP32 vectors, an FP64 pressure (the alpha-update accumulator), FP64 dot products and alpha/beta, and
// solves the coarsest level (<= 64 cells) exactly in FP64. All scalars live in a device SolverState; every solver
// kernel returns at once when `done` is set, so the host enqueues iterations in chunks and reads the state back
// once per chunk.
//
// Level layout: level 0 is the brick pool (level0.hip), levels 1..3 are pooled per brick (pooled.hip), levels 4
// and coarser are dense and run inside one workgroup, in LDS (k_coarse_tail below).
//
// Reductions are deterministic and free of cross-workgroup synchronisation: each workgroup writes its FP64
// partial sum, and a one-workgroup finalize kernel sums the partials in a fixed order and updates the state. (A
// grid-wide "last workgroup" ticket needs a device-scope fence per workgroup, which on gfx950's separate per-XCD
// L2 caches costs more than the whole pass.)
#include "slosh_ctx.h"

namespace slosh {

namespace {

inline uint32_t blocks_for(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

constexpr int kFirstDense = 4;        // first dense MG level
constexpr uint32_t kTailCells = 4800; // LDS capacity of the tail (16^3 + 8^3 + 4^3 + ...)

__device__ __forceinline__ void dcoord(const DenseLevel &L, uint32_t i, int &x, int &y, int &z)
{
    x = (int)(i % (uint32_t)L.dim[0]);
    y = (int)((i / (uint32_t)L.dim[0]) % (uint32_t)L.dim[1]);
    z = (int)(i / (uint32_t)(L.dim[0] * L.dim[1]));
}

__device__ __forceinline__ uint32_t dtype_at(const DenseLevel &L, int x, int y, int z)
{
    if (x < 0 || y < 0 || z < 0 || x >= L.dim[0] || y >= L.dim[1] || z >= L.dim[2]) return kCellSolid;
    return L.type[(uint32_t)(x + L.dim[0] * (y + L.dim[1] * z))];
}

// Stencil of a dense coarse level (SloshPressureMG.usf:142-157 without the ghost diagonal): diagonal = 1e-4 +
// non-solid neighbours, sum over fluid neighbours, neighbours in the order -x +x -y +y -z +z.
__device__ __forceinline__ void dstencil(const DenseLevel &L, const float *buf, int x, int y, int z, float &dg,
                                         float &sum)
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

enum FinKind { kFinInit = 0, kFinRho = 1, kFinAlpha = 2, kFinRr = 3, kFinSor = 4 };

// Sums the partials of the preceding reduction in a fixed order and applies them to the solver state.
__global__ __launch_bounds__(256) void k_finalize(int kind, const double *partials, uint32_t nblocks, int iters,
                                                  SolverState *st)
{
    if (kind != kFinInit && st->done) return;
    const int nv = (kind == kFinInit || kind == kFinSor) ? 2 : 1;
    __shared__ double sh[2][256];
    double a0 = 0.0, a1 = 0.0;
    for (uint32_t b = threadIdx.x; b < nblocks; b += 256) {
        a0 += partials[(size_t)b * nv];
        if (nv == 2) a1 += partials[(size_t)b * nv + 1];
    }
    sh[0][threadIdx.x] = a0;
    sh[1][threadIdx.x] = a1;
    __syncthreads();
    for (uint32_t s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double t0 = sh[0][0], t1 = sh[1][0];
    if (kind == kFinInit || kind == kFinSor) {
        // t0 = ||b||^2, t1 = ||r||^2
        st->bb = t0;
        st->rr = t1;
        double bn = sqrt(t0);
        st->rel = bn > 0.0 ? sqrt(t1) / bn : 0.0;
        if (kind == kFinInit) {
            st->iters = 0;
            st->rho = 0.0;
            st->beta = 0.0;
            st->alpha = 0.0;
            st->converged = (bn == 0.0 || st->rel <= st->tol) ? 1 : 0;
            st->done = st->converged;
        } else {
            st->iters = iters;
            if (bn == 0.0 || st->rel <= st->tol) {
                st->converged = 1;
                st->done = 1;
            } else if (iters >= st->maxit) {
                st->done = 1;
            }
        }
    } else if (kind == kFinRho) {
        st->beta = fabs(st->rho) > 1e-30 ? t0 / st->rho : 0.0;
        st->rho = t0;
    } else if (kind == kFinAlpha) {
        st->dq = t0;
        st->alpha = fabs(t0) > 1e-30 ? st->rho / t0 : 0.0;
    } else {
        st->rr = t0;
        st->rel = sqrt(t0) / sqrt(st->bb);
        st->iters += 1;
        if (st->rel <= st->tol) {
            st->converged = 1;
            st->done = 1;
        } else if (st->iters >= st->maxit) {
            st->done = 1;
        }
    }
}

// ---- dense levels (>= 4) ----------------------------------------------------------------------------------
// SloshPressureMG.usf:256-278 (MGBuildTypeCS): AIR if any child is air, else FLUID if any child is fluid.
__global__ void k_build_type(DenseLevel F, DenseLevel C)
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
    C.type[i] = (uint8_t)(air ? kCellAir : (fluid ? kCellFluid : kCellSolid));
}

__device__ __forceinline__ void smooth_cell(const DenseLevel &L, uint32_t colour, uint32_t i)
{
    if (L.type[i] != kCellFluid) return;
    int cx, cy, cz;
    dcoord(L, i, cx, cy, cz);
    if ((is_red(cx, cy, cz) ? 0u : 1u) != colour) return;
    float dg, s;
    dstencil(L, L.x, cx, cy, cz, dg, s);
    L.x[i] = (L.b[i] + s) / dg;
}

// SloshPressureMG.usf:343-381 (MGRestrictCS)
__device__ __forceinline__ void restrict_cell(const DenseLevel &F, const DenseLevel &C, uint32_t ci)
{
    int cx, cy, cz;
    dcoord(C, ci, cx, cy, cz);
    if (C.type[ci] != kCellFluid) {
        C.b[ci] = 0.0f;
        C.x[ci] = 0.0f;
        return;
    }
    float sum = 0.0f;
    for (uint32_t oi = 0; oi < 8u; ++oi) {
        int fx = 2 * cx + (int)(oi & 1u), fy = 2 * cy + (int)((oi >> 1) & 1u), fz = 2 * cz + (int)(oi >> 2);
        if (dtype_at(F, fx, fy, fz) != kCellFluid) continue;
        uint32_t fi = (uint32_t)(fx + F.dim[0] * (fy + F.dim[1] * fz));
        float dg, s;
        dstencil(F, F.x, fx, fy, fz, dg, s);
        sum += F.b[fi] - (dg * F.x[fi] - s);
    }
    float bv = 0.25f * sum;
    C.b[ci] = bv;
    float cdg = kDiagEps;
    for (int nb = 0; nb < 6; ++nb) {
        int d[3];
        nb_offset(nb, d);
        cdg += dtype_at(C, cx + d[0], cy + d[1], cz + d[2]) != kCellSolid ? 1.0f : 0.0f;
    }
    C.x[ci] = is_red(cx, cy, cz) ? bv / cdg : 0.0f;
}

// SloshPressureMG.usf:418-433 (MGProlongCS)
__device__ __forceinline__ void prolong_cell(const DenseLevel &F, const DenseLevel &C, uint32_t i)
{
    if (F.type[i] != kCellFluid) return;
    int x, y, z;
    dcoord(F, i, x, y, z);
    uint32_t pi = (uint32_t)((x >> 1) + C.dim[0] * ((y >> 1) + C.dim[1] * (z >> 1)));
    if (C.type[pi] == kCellFluid) F.x[i] += C.x[pi];
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

__global__ void k_dprolong(DenseLevel F, DenseLevel C, const SolverState *st)
{
    if (st->done) return;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < F.cells) prolong_cell(F, C, i);
}

struct TailLevels {
    int n;                       // levels t .. m-1, the last one the coarsest
    DenseLevel L[kMaxLevels];
};

// The dense levels of the V-cycle (SloshProxy.cpp:904-959) in one 1024-thread workgroup, entirely in LDS: black
// pre-sweep of the first tail level, restrictions with fused red sweeps and black pre-sweeps down to the
// coarsest level, the exact FP64 coarsest solve, then prolongation with black and red post-sweeps back up. Only
// the first level's right-hand side and initial guess come from global memory, and only its result goes back.
__global__ __launch_bounds__(1024) void k_coarse_tail(TailLevels T, const uint32_t *cidx, const uint32_t *cn,
                                                      const double *inv, const SolverState *st)
{
    if (st->done) return;
    constexpr int N = (int)kCoarsestCells;
    __shared__ uint8_t s_type[kTailCells];
    __shared__ float s_x[kTailCells], s_b[kTailCells];
    __shared__ double bs[N];
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    DenseLevel L[kMaxLevels];
    uint32_t off = 0;
    for (int k = 0; k < T.n; ++k) {
        L[k] = T.L[k];
        L[k].type = s_type + off;
        L[k].x = s_x + off;
        L[k].b = s_b + off;
        for (uint32_t i = tid; i < L[k].cells; i += nt) L[k].type[i] = T.L[k].type[i];
        off += L[k].cells;
    }
    for (uint32_t i = tid; i < L[0].cells; i += nt) {
        L[0].x[i] = T.L[0].x[i];
        L[0].b[i] = T.L[0].b[i];
    }
    __syncthreads();
    const int m = T.n;
    if (m > 1) {
        for (uint32_t i = tid; i < L[0].cells; i += nt) smooth_cell(L[0], 1u, i);
        __syncthreads();
    }
    for (int l = 0; l < m - 1; ++l) {
        for (uint32_t ci = tid; ci < L[l + 1].cells; ci += nt) restrict_cell(L[l], L[l + 1], ci);
        __syncthreads();
        if (l + 1 < m - 1) {
            for (uint32_t i = tid; i < L[l + 1].cells; i += nt) smooth_cell(L[l + 1], 1u, i);
            __syncthreads();
        }
    }
    const int n = (int)*cn;
    if ((int)tid < n) bs[tid] = (double)L[m - 1].b[cidx[tid]];
    __syncthreads();
    if ((int)tid < n) {
        double s = 0.0;
        for (int j = 0; j < n; ++j) s += inv[(size_t)tid * N + j] * bs[j];
        L[m - 1].x[cidx[tid]] = (float)s;
    }
    __syncthreads();
    for (int l = m - 2; l >= 0; --l) {
        for (uint32_t i = tid; i < L[l].cells; i += nt) prolong_cell(L[l], L[l + 1], i);
        __syncthreads();
        for (uint32_t i = tid; i < L[l].cells; i += nt) smooth_cell(L[l], 1u, i);
        __syncthreads();
        for (uint32_t i = tid; i < L[l].cells; i += nt) smooth_cell(L[l], 0u, i);
        __syncthreads();
    }
    for (uint32_t i = tid; i < L[0].cells; i += nt) T.L[0].x[i] = L[0].x[i];
}

// ---- coarsest level: FP64 Cholesky of the <= 64 fluid unknowns and its explicit inverse (once per solve) ----
// The coarsest level is level 0 (mode 0, grids of <= 64 cells), a pooled level (mode 1) or a dense level (mode 2).
__global__ __launch_bounds__(256) void k_coarse_factor(Geo g, const uint32_t *ctype, const float *diag, int mode,
                                                       PLevel P, DenseLevel D, uint32_t *cidx, uint32_t *cn,
                                                       double *inv)
{
    constexpr int N = (int)kCoarsestCells;
    __shared__ double a[N][N + 1];
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
    if (t == 0) {
        uint32_t n = 0;
        for (uint32_t i = 0; i < cells && n < (uint32_t)N; ++i) {
            int x = (int)(i % (uint32_t)dims[0]), y = (int)((i / (uint32_t)dims[0]) % (uint32_t)dims[1]),
                z = (int)(i / (uint32_t)(dims[0] * dims[1]));
            if (type_of(x, y, z) != kCellFluid) continue;
            px[n] = x;
            py[n] = y;
            pz[n] = z;
            pidx[n] = index_of(x, y, z);
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
        if (mode == 0) dg = diag[pidx[t]] + kDiagEps;
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

// Exact coarsest solve x = A^-1 b in FP64 (replaces MGCoarseSolveCS, SloshPressureMG.usf:386-413) when the
// coarsest level is level 0 or pooled; b and x are indexed by the factor's cell list.
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

void finalize(slosh_ctx *c, int kind, int iters = 0)
{
    SLOSH_TIMED(c, K_FINALIZE);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, c->stream, kind, c->partials, l0_groups(c), iters, c->state);
}

void smooth_level(slosh_ctx *c, int l, uint32_t colour, int want_dot)
{
    if (l == 0) {
        l0_smooth(c, colour, want_dot);
        if (want_dot) finalize(c, kFinRho);
    } else {
        pooled_smooth(c, l, colour);
    }
}

// SloshProxy.cpp:904-959 (VCycle) with the exact coarsest solve. The level-0 red pre-sweep is already in z.
void vcycle(slosh_ctx *c)
{
    const int m = c->nlevels;
    if (m == 1) {
        {
            SLOSH_TIMED(c, K_COARSE_SOLVE);
            hipLaunchKernelGGL(k_coarse_solve, dim3(1), dim3(64), 0, c->stream, c->coarse_idx, c->coarse_n,
                               c->coarse_inv, c->r, c->z, c->state);
        }
        smooth_level(c, 0, 0u, 1);
        return;
    }
    // multi-workgroup levels: 0 .. top, where top is the coarsest pooled level
    const int top = m - 1 < kFirstDense ? m - 1 : kFirstDense - 1;
    smooth_level(c, 0, 1u, 0);
    for (int l = 0; l < top; ++l) {
        if (l == 0) l0_restrict(c);
        else pooled_restrict(c, l);
        if (l + 1 < m - 1) smooth_level(c, l + 1, 1u, 0);
    }
    if (m - 1 == top) {
        // the coarsest level is pooled (or level 1..3): exact solve on it
        SLOSH_TIMED(c, K_COARSE_SOLVE);
        hipLaunchKernelGGL(k_coarse_solve, dim3(1), dim3(64), 0, c->stream, c->coarse_idx, c->coarse_n, c->coarse_inv,
                           c->lv[top].b, c->lv[top].x, c->state);
    } else {
        // dense levels kFirstDense .. t-1 as grid-wide passes, levels t .. m-1 in the LDS tail
        int t = m - 1;
        uint32_t tail_cells = c->lv[m - 1].cells;
        while (t - 1 >= kFirstDense && tail_cells + c->lv[t - 1].cells <= kTailCells) tail_cells += c->lv[--t].cells;
        restrict_pooled_to_dense(c, top);
        for (int l = kFirstDense; l < t; ++l) {
            DenseLevel F = dlevel_of(c, l), C = dlevel_of(c, l + 1);
            {
                SLOSH_TIMED(c, K_SMOOTH);
                hipLaunchKernelGGL(k_dsmooth, dim3(blocks_for(F.cells, 256)), dim3(256), 0, c->stream, F, 1u, c->state);
            }
            SLOSH_TIMED(c, K_RESTRICT);
            hipLaunchKernelGGL(k_drestrict, dim3(blocks_for(C.cells, 256)), dim3(256), 0, c->stream, F, C, c->state);
        }
        TailLevels T;
        T.n = m - t;
        for (int k = 0; k < kMaxLevels; ++k) T.L[k] = k < T.n ? dlevel_of(c, t + k) : DenseLevel{};
        {
            SLOSH_TIMED(c, K_COARSE_TAIL);
            hipLaunchKernelGGL(k_coarse_tail, dim3(1), dim3(1024), 0, c->stream, T, c->coarse_idx, c->coarse_n,
                               c->coarse_inv, c->state);
        }
        for (int l = t - 1; l >= kFirstDense; --l) {
            DenseLevel F = dlevel_of(c, l), C = dlevel_of(c, l + 1);
            {
                SLOSH_TIMED(c, K_PROLONG);
                hipLaunchKernelGGL(k_dprolong, dim3(blocks_for(F.cells, 256)), dim3(256), 0, c->stream, F, C, c->state);
            }
            for (uint32_t colour : {1u, 0u}) {
                SLOSH_TIMED(c, K_SMOOTH);
                hipLaunchKernelGGL(k_dsmooth, dim3(blocks_for(F.cells, 256)), dim3(256), 0, c->stream, F, colour,
                                   c->state);
            }
        }
        prolong_dense_to_pooled(c, top);
        smooth_level(c, top, 1u, 0);
        smooth_level(c, top, 0u, 0);
    }
    for (int l = top - 1; l >= 0; --l) {
        if (l == 0) l0_prolong(c);
        else pooled_prolong(c, l);
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
        if (c->lv[l].pooled) {
            pooled_build_type(c, l);
        } else if (c->lv[l - 1].pooled) {
            dense_build_type_from_pooled(c, l);
        } else {
            DenseLevel C = dlevel_of(c, l);
            SLOSH_TIMED(c, K_MG_TYPES);
            hipLaunchKernelGGL(k_build_type, dim3(blocks_for(C.cells, 256)), dim3(256), 0, c->stream, dlevel_of(c, l - 1),
                               C);
        }
    }
    {
        const int mode = m == 1 ? 0 : (c->lv[m - 1].pooled ? 1 : 2);
        PLevel P = mode == 1 ? plevel_of(c, m - 1) : PLevel{};
        DenseLevel D = mode == 2 ? dlevel_of(c, m - 1) : DenseLevel{};
        SLOSH_TIMED(c, K_COARSE_FACTOR);
        hipLaunchKernelGGL(k_coarse_factor, dim3(1), dim3(256), 0, c->stream, c->geo, c->ctype, c->diag, mode, P, D,
                           c->coarse_idx, c->coarse_n, c->coarse_inv);
    }
    l0_pcg_init(c);
    finalize(c, kFinInit);
    vcycle(c);
    l0_update_d(c, 1);
    SLOSH_HIP(hipGetLastError());
    const int chunk = 4;
    for (int it = 0; it <= maxit; it += chunk) {
        for (int k = 0; k < chunk; ++k) {
            l0_apply_a(c);
            finalize(c, kFinAlpha);
            l0_update_pr(c);
            finalize(c, kFinRr);
            vcycle(c);
            l0_update_d(c, 0);
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
