// level0.hip -- the finest-level (brick-pool) passes of the pressure solvers on gfx950.
//
// SloshPressureMG.usf's level-0 kernels (PCGInitCS :284-305, MGSmoothCS :311-338, MGRestrictCS :343-381,
// MGProlongCS :418-433, PCGApplyACS :438-454, PCGUpdatePRCS :495-524, PCGUpdateDCS :527-543) and
// Slosh.usf:856-892 (PressureRBGSCS), restated for the MI355X:
//   - one workgroup per allocated brick, 256 threads, thread t owning the red/black pair of cells
//     (colour_cell(t, 0), colour_cell(t, 1)); the brick's slot and its six face neighbours' slots come from a
//     per-step table of every allocated brick's 3x3x3 neighbourhood (scalar loads), so no pass walks the brick
//     table;
//   - the stencil reads a per-cell mask of fluid neighbours (written by the divergence pass) instead of six
//     neighbour cell types;
//   - passes that the reference runs back to back are fused where the second needs only values the first can
//     recompute for a neighbour: the residual update with the red initialisation and the black sweep, the
//     prolongation with the black sweep, and the direction update with the operator application. The residual,
//     the iterate and the direction ping-pong between buffers so that no workgroup reads a value another one is
//     overwriting in the same pass. Every value is computed with the reference's expression and operand order.
#include "slosh_ctx.h"

namespace slosh {

namespace {

constexpr uint32_t kThreads = 256;
constexpr uint32_t kSelfFluid = 1u << 6;

// Wave-uniform description of one allocated brick: its slot base and the slot bases of its six face neighbours
// (0 = null brick, which only masked-out neighbours can point into).
struct Brick {
    uint32_t base;
    uint32_t nb[6];
};

// Entries of the 27-neighbourhood brick table (index (dx+1) + 3(dy+1) + 9(dz+1)) of the brick itself and of its
// face neighbours in the order -x +x -y +y -z +z.
__device__ __forceinline__ Brick load_brick(const uint32_t *binfo, uint32_t e)
{
    constexpr int kFace[6] = {12, 14, 10, 16, 4, 22};
    Brick B;
    const uint32_t *p = binfo + (size_t)e * kBinfoWords;
    B.base = p[13] * kBrickCells;
    for (int n = 0; n < 6; ++n) B.nb[n] = p[kFace[n]] * kBrickCells;
    return B;
}

// Pool index of neighbour n (order -x +x -y +y -z +z) of local cell l, and its local index in its own brick.
__device__ __forceinline__ uint32_t nbr(const Brick &B, uint32_t l, int n, uint32_t &nl)
{
    switch (n) {
    case 0: nl = (l & 7u) ? l - 1u : l + 7u; return ((l & 7u) ? B.base : B.nb[0]) + nl;
    case 1: nl = ((l & 7u) != 7u) ? l + 1u : l - 7u; return (((l & 7u) != 7u) ? B.base : B.nb[1]) + nl;
    case 2: nl = (l & 56u) ? l - 8u : l + 56u; return ((l & 56u) ? B.base : B.nb[2]) + nl;
    case 3: nl = ((l & 56u) != 56u) ? l + 8u : l - 56u; return (((l & 56u) != 56u) ? B.base : B.nb[3]) + nl;
    case 4: nl = (l & 448u) ? l - 64u : l + 448u; return ((l & 448u) ? B.base : B.nb[4]) + nl;
    default: nl = ((l & 448u) != 448u) ? l + 64u : l - 448u; return (((l & 448u) != 448u) ? B.base : B.nb[5]) + nl;
    }
}

__device__ __forceinline__ uint32_t nbr(const Brick &B, uint32_t l, int n)
{
    uint32_t nl;
    return nbr(B, l, n, nl);
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

// Level-1 parent (local index within the brick's 4^3 block) of level-0 local cell l.
__device__ __forceinline__ uint32_t parent1(uint32_t l)
{
    return ((l & 7u) >> 1) + 4u * (((l >> 3) & 7u) >> 1) + 16u * ((l >> 6) >> 1);
}

// Workgroup sum of each thread's values into partials[blockIdx.x * NV + k]: wave_sum within each wavefront, then
// the four wavefront sums added in wavefront order by thread 0 (one barrier). The order is fixed, so repeated runs
// give identical partials.
template <int NV> __device__ void block_partial(const double (&v)[NV], double *partials)
{
    static_assert(NV >= 1 && NV <= (int)kMaxBlockPartials, "partials buffer holds kMaxBlockPartials per brick");
    static_assert(kThreads == 256u, "four wavefronts");
    __shared__ double sh[NV][kThreads / 64u];
    const uint32_t t = threadIdx.x;
    double a[NV];
    for (int k = 0; k < NV; ++k) a[k] = wave_sum(v[k]);
    if ((t & 63u) == 0u)
        for (int k = 0; k < NV; ++k) sh[k][t >> 6] = a[k];
    __syncthreads();
    if (t == 0)
        for (int k = 0; k < NV; ++k)
            partials[(size_t)blockIdx.x * NV + k] = ((sh[k][0] + sh[k][1]) + sh[k][2]) + sh[k][3];
}

// 1 for the allocated bricks holding fluid cells (their count from the P2G classification).
__global__ void k_fluid_flags(uint32_t nactive, const uint32_t *bp_fluid, uint32_t *flag)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nactive) flag[e] = bp_fluid[e] != 0u ? 1u : 0u;
}

// The neighbourhood rows (built once per step after the pool) of the bricks with fluid, in active order, to binfo_f.
__global__ void k_fluid_rows(uint32_t nactive, const uint32_t *binfo, const uint32_t *flag, const uint32_t *rank,
                             uint32_t *binfo_f)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t e = t / kBinfoWords, k = t % kBinfoWords;
    if (e >= nactive || !flag[e]) return;
    binfo_f[(size_t)rank[e] * kBinfoWords + k] = binfo[t];
}

// r = b - A p (FP64 p); partials ||b||^2, ||r||^2 and the fluid cell count.
__global__ __launch_bounds__(kThreads) void k_pcg_init(const uint32_t *binfo, const uint8_t *fmask, const float *div,
                                                       const float *diag, const double *pres, float *r,
                                                       double *partials)
{
    const Brick B = load_brick(binfo, xcd_remap(blockIdx.x, gridDim.x));
    double v[3] = {0.0, 0.0, 0.0};
    for (uint32_t c = 0; c < 2u; ++c) {
        const uint32_t l = colour_cell(threadIdx.x, c), i = B.base + l, m = fmask[i];
        if (!(m & kSelfFluid)) {
            r[i] = 0.0f;
            continue;
        }
        v[2] += 1.0;
        double s = nbsum<double>(B, l, m, pres);
        float dg = diag[i] + kDiagEps;
        float ri = (float)(-(double)div[i] - ((double)dg * pres[i] - s));
        r[i] = ri;
        v[0] += (double)div[i] * (double)div[i];
        v[1] += (double)ri * (double)ri;
    }
    block_partial<3>(v, partials);
}

// The residual update of the iteration fused with the V-cycle's first level-0 sweeps (PCGUpdatePRCS with its
// fused red initialisation, then MGSmoothCS black): red z = r / diag, black z = (r + sum_n z_n) / diag where
// the red neighbours' z is recomputed from r_in and q.
__global__ __launch_bounds__(kThreads) void k_update_black(const uint32_t *binfo, const uint8_t *fmask,
                                                           const float *diag, double *pres, const float *r_in,
                                                           float *r_out, float *z, const float *d, const float *q,
                                                           int first, const SolverState *st)
{
    const Brick B = load_brick(binfo, xcd_remap(blockIdx.x, gridDim.x));
    const uint32_t done = st->done;
    const double alpha = first ? 0.0 : st->alpha;
    issue_point();
    if (done == 1u) return; // done == 2: the solve's last update, made before every other pass stops
    const uint32_t l0 = colour_cell(threadIdx.x, 0u), l1 = colour_cell(threadIdx.x, 1u);
    const uint32_t i0 = B.base + l0, i1 = B.base + l1;
    const uint32_t m0 = fmask[i0], m1 = fmask[i1];
    const float g0 = diag[i0], g1 = diag[i1], ra0 = r_in[i0], ra1 = r_in[i1];
    float q0 = 0.0f, q1 = 0.0f, d0 = 0.0f, d1 = 0.0f;
    double p0 = 0.0, p1 = 0.0;
    float rj[6], qj[6], gj[6];
    for (int n = 0; n < 6; ++n) {
        const uint32_t j = nbr(B, l1, n);
        rj[n] = r_in[j];
        gj[n] = diag[j];
    }
    if (!first) {
        q0 = q[i0];
        q1 = q[i1];
        d0 = d[i0];
        d1 = d[i1];
        p0 = pres[i0];
        p1 = pres[i1];
        for (int n = 0; n < 6; ++n) qj[n] = q[nbr(B, l1, n)];
    }
    if (m0 & kSelfFluid) {
        float ri = ra0;
        if (!first) {
            pres[i0] = p0 + alpha * (double)d0;
            ri = (float)((double)ra0 - alpha * (double)q0);
        }
        r_out[i0] = ri;
        z[i0] = ri / (g0 + kDiagEps);
    }
    if (m1 & kSelfFluid) {
        float ri = ra1;
        if (!first) {
            pres[i1] = p1 + alpha * (double)d1;
            ri = (float)((double)ra1 - alpha * (double)q1);
        }
        r_out[i1] = ri;
        float s = 0.0f;
        for (int n = 0; n < 6; ++n) {
            if (!(m1 >> n & 1u)) continue;
            const float r = first ? rj[n] : (float)((double)rj[n] - alpha * (double)qj[n]);
            s += r / (gj[n] + kDiagEps);
        }
        z[i1] = (ri + s) / (g1 + kDiagEps);
    }
}

// Restriction into pooled level 1: the residuals of the brick's 512 cells in LDS, then b_c = 1/4 of the fluid
// children's sum for the brick's 4^3 level-1 cells, and x_c = b_c / diag_c on red cells (fused red sweep).
__global__ __launch_bounds__(kThreads) void k_restrict0(const uint32_t *binfo, const uint8_t *fmask,
                                                        const float *diag, const float *xf, const float *bf,
                                                        PLevel L1, const SolverState *st)
{
    __shared__ float res[kBrickCells];
    __shared__ uint8_t fl[kBrickCells];
    const Brick B = load_brick(binfo, xcd_remap(blockIdx.x, gridDim.x));
    const uint32_t done = st->done;
    const double mg_scale = st->mg_scale;
    issue_point();
    if (done) return;
    const uint32_t t = threadIdx.x;
    uint32_t l[2], m[2];
    float xv[2], bv[2], gv[2], xn[2][6];
    for (uint32_t c = 0; c < 2u; ++c) {
        l[c] = colour_cell(t, c);
        const uint32_t i = B.base + l[c];
        m[c] = fmask[i];
        xv[c] = xf[i];
        bv[c] = bf[i];
        gv[c] = diag[i];
        for (int n = 0; n < 6; ++n) xn[c][n] = xf[nbr(B, l[c], n)];
    }
    const uint32_t ci = (B.base >> 3) + (t & 63u);
    const uint32_t info = L1.info[ci];
    for (uint32_t c = 0; c < 2u; ++c) {
        fl[l[c]] = (uint8_t)((m[c] & kSelfFluid) ? 1u : 0u);
        if (m[c] & kSelfFluid) {
            float s = 0.0f;
            for (int n = 0; n < 6; ++n)
                if (m[c] >> n & 1u) s += xn[c][n];
            float dg = gv[c] + kDiagEps;
            res[l[c]] = bv[c] - (dg * xv[c] - s);
        }
    }
    __syncthreads();
    if (t >= 64u) return;
    const uint32_t lx = t & 3u, ly = (t >> 2) & 3u, lz = t >> 4;
    if (!(info & kInfoSelf)) {
        L1.b[ci] = 0.0f;
        L1.x[ci] = 0.0f;
        return;
    }
    float sum = 0.0f;
    for (uint32_t oi = 0; oi < 8u; ++oi) {
        uint32_t fx = 2u * lx + (oi & 1u), fy = 2u * ly + ((oi >> 1) & 1u), fz = 2u * lz + (oi >> 2);
        uint32_t f = fx + 8u * fy + 64u * fz;
        if (fl[f]) sum += res[f];
    }
    float b = 0.25f * sum;
    if (L1.b.half) b *= (float)mg_scale; // FP16 levels hold the scaled cycle (the V-cycle is linear)
    L1.b[ci] = b;
    L1.x[ci] = (info & kInfoRed) ? b / info_diag(info) : 0.0f;
}

// Prolongation from pooled level 1 (x1) fused with the black sweep: red z1 = z + x1(parent), black
// z1 = (r + sum_n (z_n + x1(parent_n))) / diag; a parent contributes only when it is fluid.
__global__ __launch_bounds__(kThreads) void k_prolong_black(const uint32_t *binfo, const uint8_t *fmask,
                                                            const float *diag, const float *r, const float *z,
                                                            float *z1, PLevel L1, const SolverState *st)
{
    const Brick B = load_brick(binfo, xcd_remap(blockIdx.x, gridDim.x));
    const uint32_t done = st->done;
    const double mg_scale = st->mg_scale;
    issue_point();
    if (done) return;
    const uint32_t l0 = colour_cell(threadIdx.x, 0u), l1 = colour_cell(threadIdx.x, 1u);
    const uint32_t i0 = B.base + l0, i1 = B.base + l1;
    const uint32_t m0 = fmask[i0], m1 = fmask[i1];
    const uint32_t pi0 = (B.base >> 3) + parent1(l0);
    const float z0 = z[i0], r1 = r[i1], g1 = diag[i1];
    const uint32_t pinfo0 = L1.info[pi0];
    const float px0 = L1.x1[pi0];
    float zn[6], pxn[6];
    uint32_t pinfon[6];
    for (int n = 0; n < 6; ++n) {
        uint32_t nl;
        const uint32_t j = nbr(B, l1, n, nl);
        const uint32_t pj = ((j - nl) >> 3) + parent1(nl);
        zn[n] = z[j];
        pinfon[n] = L1.info[pj];
        pxn[n] = L1.x1[pj];
    }
    const float unscale = L1.x1.half ? (float)(1.0 / mg_scale) : 1.0f;
    if (m0 & kSelfFluid) {
        float v = z0;
        if (pinfo0 & kInfoSelf) v += px0 * unscale;
        z1[i0] = v;
    }
    if (m1 & kSelfFluid) {
        float s = 0.0f;
        for (int n = 0; n < 6; ++n) {
            if (!(m1 >> n & 1u)) continue;
            float v = zn[n];
            if (pinfon[n] & kInfoSelf) v += pxn[n] * unscale;
            s += v;
        }
        z1[i1] = (r1 + s) / (g1 + kDiagEps);
    }
}

// Red sweep z1 = (r + sum_n z1_n) / diag and the partials of r.z1 over all fluid cells.
__global__ __launch_bounds__(kThreads) void k_red_dot(const uint32_t *binfo, const uint8_t *fmask, const float *diag,
                                                      const float *r, float *z1, double *partials,
                                                      const SolverState *st)
{
    const Brick B = load_brick(binfo, xcd_remap(blockIdx.x, gridDim.x));
    const uint32_t done = st->done;
    issue_point();
    if (done) return;
    const uint32_t l0 = colour_cell(threadIdx.x, 0u), i0 = B.base + l0, i1 = B.base + colour_cell(threadIdx.x, 1u);
    const uint32_t m0 = fmask[i0], m1 = fmask[i1];
    const float r0 = r[i0], g0 = diag[i0], r1 = r[i1], zb1 = z1[i1];
    float zn[6];
    for (int n = 0; n < 6; ++n) zn[n] = z1[nbr(B, l0, n)];
    double v[1] = {0.0};
    if (m0 & kSelfFluid) {
        float s = 0.0f;
        for (int n = 0; n < 6; ++n)
            if (m0 >> n & 1u) s += zn[n];
        float xv = (r0 + s) / (g0 + kDiagEps);
        z1[i0] = xv;
        v[0] += (double)r0 * (double)xv;
    }
    if (m1 & kSelfFluid) v[0] += (double)r1 * (double)zb1;
    block_partial<1>(v, partials);
}

// d_out = z1 + beta d_in (beta = 0 for the first direction) fused with q = A d_out: the neighbours' new
// direction is recomputed from z1 and d_in. The stencil sum is FP64 (for smooth d the diagonal and neighbour
// terms cancel, and an FP32 sum would put a large relative error into q that the recurrence residual never
// sees); partials d.q.
__global__ __launch_bounds__(kThreads) void k_direction(const uint32_t *binfo, const uint8_t *fmask,
                                                        const float *diag, const float *z1, const float *d_in,
                                                        float *d_out, float *q, const float *r, int first,
                                                        double *partials, const SolverState *st)
{
    const Brick B = load_brick(binfo, xcd_remap(blockIdx.x, gridDim.x));
    const uint32_t done = st->done;
    const double beta = first ? 0.0 : st->beta;
    issue_point();
    if (done) return;
    uint32_t i[2], m[2];
    float zc[2], dc[2] = {0.0f, 0.0f}, gc[2], rc[2], zn[2][6], dn[2][6];
    for (uint32_t c = 0; c < 2u; ++c) {
        const uint32_t l = colour_cell(threadIdx.x, c);
        i[c] = B.base + l;
        m[c] = fmask[i[c]];
        zc[c] = z1[i[c]];
        gc[c] = diag[i[c]];
        rc[c] = r[i[c]];
        for (int n = 0; n < 6; ++n) zn[c][n] = z1[nbr(B, l, n)];
    }
    if (!first) {
        for (uint32_t c = 0; c < 2u; ++c) {
            const uint32_t l = colour_cell(threadIdx.x, c);
            dc[c] = d_in[i[c]];
            for (int n = 0; n < 6; ++n) dn[c][n] = d_in[nbr(B, l, n)];
        }
    }
    // d.q, and for the stopping test of the coming update r - alpha q: r.q, q.q and r.r (q as stored, in FP32)
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t c = 0; c < 2u; ++c) {
        if (!(m[c] & kSelfFluid)) continue;
        const float dv = first ? zc[c] : (float)((double)zc[c] + beta * (double)dc[c]);
        d_out[i[c]] = dv;
        double s = 0.0;
        for (int n = 0; n < 6; ++n) {
            if (!(m[c] >> n & 1u)) continue;
            const float dj = first ? zn[c][n] : (float)((double)zn[c][n] + beta * (double)dn[c][n]);
            s += (double)dj;
        }
        const float dg = gc[c] + kDiagEps;
        const double qd = (double)dg * (double)dv - s;
        const float qf = (float)qd;
        q[i[c]] = qf;
        v[0] += (double)dv * qd;
        v[1] += (double)rc[c] * (double)qf;
        v[2] += (double)qf * (double)qf;
        v[3] += (double)rc[c] * (double)rc[c];
    }
    block_partial<4>(v, partials);
}

// Red-black SOR sweep on the FP64 pressure: p += w ((sum p_n - div) / diag - p).
__global__ __launch_bounds__(kThreads) void k_sor_sweep(const uint32_t *binfo, const uint8_t *fmask, const float *div,
                                                        const float *diag, double *pres, uint32_t colour, double w,
                                                        const SolverState *st)
{
    if (st->done) return;
    const Brick B = load_brick(binfo, xcd_remap(blockIdx.x, gridDim.x));
    const uint32_t l = colour_cell(threadIdx.x, colour), i = B.base + l, m = fmask[i];
    if (!(m & kSelfFluid)) return;
    double s = nbsum<double>(B, l, m, pres);
    double dg = (double)(diag[i] + kDiagEps);
    double gs = (s - (double)div[i]) / dg;
    double old = pres[i];
    pres[i] = old + w * (gs - old);
}

// Partials ||b||^2 and ||b - A p||^2: mode 0 the SOR convergence check (skipped once the solver is done), 1 always,
// 2 only once the solver is done (the true-residual check of a finished MGPCG solve).
__global__ __launch_bounds__(kThreads) void k_sor_residual(const uint32_t *binfo, const uint8_t *fmask,
                                                           const float *div, const float *diag, const double *pres,
                                                           double *partials, const SolverState *st, int mode)
{
    const int32_t done = st->done;
    if ((mode == 0 && done) || (mode == 2 && !done)) return;
    const Brick B = load_brick(binfo, xcd_remap(blockIdx.x, gridDim.x));
    double v[2] = {0.0, 0.0};
    for (uint32_t c = 0; c < 2u; ++c) {
        const uint32_t l = colour_cell(threadIdx.x, c), i = B.base + l, m = fmask[i];
        if (!(m & kSelfFluid)) continue;
        double s = nbsum<double>(B, l, m, pres);
        double dg = (double)(diag[i] + kDiagEps);
        double ri = -(double)div[i] - (dg * pres[i] - s);
        v[0] += (double)div[i] * (double)div[i];
        v[1] += ri * ri;
    }
    block_partial<2>(v, partials);
}

// ---- flexible CG (Notay's FCG(1)) with a nonlinear preconditioner z = M(r) ----
// Partials of z . q_prev (the orthogonalisation coefficient of the next direction).
__global__ __launch_bounds__(kThreads) void k_fcg_zq(const uint32_t *binfo, const uint8_t *fmask, const float *z,
                                                     const float *q, double *partials, const SolverState *st)
{
    if (st->done) return;
    const Brick B = load_brick(binfo, xcd_remap(blockIdx.x, gridDim.x));
    double v[1] = {0.0};
    for (uint32_t c = 0; c < 2u; ++c) {
        const uint32_t i = B.base + colour_cell(threadIdx.x, c);
        if (fmask[i] & kSelfFluid) v[0] += (double)z[i] * (double)q[i];
    }
    block_partial<1>(v, partials);
}

// d_out = z + beta d_in (first: z), beta = -(z . q_prev) / (d_prev . q_prev), fused with q = A d_out (FP64 stencil
// sum, neighbours' directions recomputed); partials d . q and d . r.
__global__ __launch_bounds__(kThreads) void k_fcg_direction(const uint32_t *binfo, const uint8_t *fmask,
                                                            const float *diag, const float *z, const float *d_in,
                                                            float *d_out, float *q, const float *r, int first,
                                                            double *partials, const SolverState *st)
{
    if (st->done) return;
    const double beta = first ? 0.0 : st->beta;
    const Brick B = load_brick(binfo, xcd_remap(blockIdx.x, gridDim.x));
    double v[2] = {0.0, 0.0};
    for (uint32_t c = 0; c < 2u; ++c) {
        const uint32_t l = colour_cell(threadIdx.x, c), i = B.base + l, m = fmask[i];
        if (!(m & kSelfFluid)) continue;
        const float dv = first ? z[i] : (float)((double)z[i] + beta * (double)d_in[i]);
        d_out[i] = dv;
        double s = 0.0;
        for (int n = 0; n < 6; ++n) {
            if (!(m >> n & 1u)) continue;
            const uint32_t j = nbr(B, l, n);
            s += (double)(first ? z[j] : (float)((double)z[j] + beta * (double)d_in[j]));
        }
        const double qd = (double)(diag[i] + kDiagEps) * (double)dv - s;
        q[i] = (float)qd;
        v[0] += (double)dv * qd;
        v[1] += (double)dv * (double)r[i];
    }
    block_partial<2>(v, partials);
}

// p += alpha d (FP64), r -= alpha q; partials ||r||^2.
__global__ __launch_bounds__(kThreads) void k_fcg_update(const uint32_t *binfo, const uint8_t *fmask, double *pres,
                                                         float *r, const float *d, const float *q, double *partials,
                                                         const SolverState *st)
{
    if (st->done) return;
    const double alpha = st->alpha;
    const Brick B = load_brick(binfo, xcd_remap(blockIdx.x, gridDim.x));
    double v[1] = {0.0};
    for (uint32_t c = 0; c < 2u; ++c) {
        const uint32_t i = B.base + colour_cell(threadIdx.x, c);
        if (!(fmask[i] & kSelfFluid)) continue;
        pres[i] = pres[i] + alpha * (double)d[i];
        const float ri = (float)((double)r[i] - alpha * (double)q[i]);
        r[i] = ri;
        v[0] += (double)ri * (double)ri;
    }
    block_partial<1>(v, partials);
}

// out = r - A z over fluid cells (FP64 stencil sum, FP32 result), 0 elsewhere: the residual the V-cycle's result
// z leaves, which the network of the hybrid preconditioner corrects.
__global__ __launch_bounds__(kThreads) void k_fcg_resid(const uint32_t *binfo, const uint8_t *fmask,
                                                        const float *diag, const float *r, const float *z, float *out,
                                                        const SolverState *st)
{
    if (st->done) return;
    const Brick B = load_brick(binfo, xcd_remap(blockIdx.x, gridDim.x));
    for (uint32_t c = 0; c < 2u; ++c) {
        const uint32_t l = colour_cell(threadIdx.x, c), i = B.base + l, m = fmask[i];
        if (!(m & kSelfFluid)) {
            out[i] = 0.0f;
            continue;
        }
        const double s = nbsum<double>(B, l, m, z);
        const double az = (double)(diag[i] + kDiagEps) * (double)z[i] - s;
        out[i] = (float)((double)r[i] - az);
    }
}

// Partials sum (p_ref - p)^2 and sum p^2 over fluid cells (the error of a kept solution against the current one).
__global__ __launch_bounds__(kThreads) void k_error_partials(const uint32_t *binfo, const uint8_t *fmask,
                                                             const double *pres, const double *p_ref,
                                                             double *partials)
{
    const Brick B = load_brick(binfo, xcd_remap(blockIdx.x, gridDim.x));
    double v[2] = {0.0, 0.0};
    for (uint32_t c = 0; c < 2u; ++c) {
        const uint32_t i = B.base + colour_cell(threadIdx.x, c);
        if (!(fmask[i] & kSelfFluid)) continue;
        const double e = p_ref[i] - pres[i];
        v[0] += e * e;
        v[1] += pres[i] * pres[i];
    }
    block_partial<2>(v, partials);
}

} // namespace

#define L0_LAUNCH(kid, kernel, ...)                                                                              \
    launch_timed(c, slosh::kid, kernel, dim3(c->nsolve), dim3(kThreads), 0u, c->binfo_f, c->fmask, __VA_ARGS__)

void l0_brick_info(slosh_ctx *c, bool readback)
{
    SLOSH_TIMED(c, K_MG_SETUP);
    const uint32_t nb = (c->nactive + 255u) / 256u;
    hipLaunchKernelGGL(k_fluid_flags, dim3(nb), dim3(256), 0, c->stream, c->nactive, c->bp_fluid, c->fluid_flag);
    scan_exclusive(c, c->fluid_flag, c->fluid_rank, c->nactive);
    hipLaunchKernelGGL(k_fluid_rows, dim3(nb * kBinfoWords), dim3(256), 0, c->stream, c->nactive, c->binfo,
                       c->fluid_flag, c->fluid_rank, c->binfo_f);
    // the level-0 and pooled passes launch one workgroup per brick with fluid: their count comes back here (or with
    // the caller's readback)
    SLOSH_HIP(hipMemcpyAsync(c->h_nsolve, c->fluid_rank + c->nactive, sizeof(uint32_t), hipMemcpyDeviceToHost,
                             c->stream));
    if (readback) {
        SLOSH_HIP(hipStreamSynchronize(c->stream));
        c->nsolve = c->h_nsolve[0];
    }
}

// Over every allocated brick (non-fluid cells get r = 0, also in bricks without fluid).
void l0_pcg_init(slosh_ctx *c, float *r_out)
{
    launch_timed(c, K_PCG_INIT, k_pcg_init, dim3(c->nactive), dim3(kThreads), 0, c->binfo, c->fmask, c->div, c->diag,
                 c->pres, r_out, c->partials);
}

void l0_update_black(slosh_ctx *c, const float *r_in, float *r_out, const float *d, int first)
{
    L0_LAUNCH(K_UPDATE_BLACK, k_update_black, c->diag, c->pres, r_in, r_out, c->z, d, c->q, first, c->state);
}

void l0_restrict(slosh_ctx *c, const float *r)
{
    L0_LAUNCH(K_RESTRICT0, k_restrict0, c->diag, c->z, r, plevel_of(c, 1), c->state);
}

void l0_prolong_black(slosh_ctx *c, const float *r)
{
    L0_LAUNCH(K_PROLONG_BLACK, k_prolong_black, c->diag, r, c->z, c->z1, plevel_of(c, 1), c->state);
}

void l0_red_dot(slosh_ctx *c, const float *r) { L0_LAUNCH(K_RED_DOT, k_red_dot, c->diag, r, c->z1, c->partials, c->state); }

void l0_direction(slosh_ctx *c, const float *d_in, float *d_out, const float *r, int first)
{
    L0_LAUNCH(K_DIRECTION, k_direction, c->diag, c->z1, d_in, d_out, c->q, r, first, c->partials, c->state);
}

void l0_fcg_zq(slosh_ctx *c, const float *z) { L0_LAUNCH(K_FCG, k_fcg_zq, z, c->q, c->partials, c->state); }

void l0_fcg_direction(slosh_ctx *c, const float *z, const float *d_in, float *d_out, const float *r, int first)
{
    L0_LAUNCH(K_FCG, k_fcg_direction, c->diag, z, d_in, d_out, c->q, r, first, c->partials, c->state);
}

void l0_fcg_update(slosh_ctx *c, float *r, const float *d)
{
    L0_LAUNCH(K_FCG, k_fcg_update, c->pres, r, d, c->q, c->partials, c->state);
}

void l0_sor_sweep(slosh_ctx *c, uint32_t colour, double w)
{
    L0_LAUNCH(K_SOR_SWEEP, k_sor_sweep, c->div, c->diag, c->pres, colour, w, c->state);
}

void l0_sor_residual(slosh_ctx *c)
{
    L0_LAUNCH(K_SOR_RESIDUAL, k_sor_residual, c->div, c->diag, c->pres, c->partials, c->state, 0);
}

void l0_fcg_resid(slosh_ctx *c, const float *r, const float *z, float *out)
{
    L0_LAUNCH(K_FCG, k_fcg_resid, c->diag, r, z, out, c->state);
}

void l0_error_partials(slosh_ctx *c, const double *p_ref)
{
    L0_LAUNCH(K_SOR_RESIDUAL, k_error_partials, c->pres, p_ref, c->partials);
}

void l0_true_residual(slosh_ctx *c)
{
    L0_LAUNCH(K_SOR_RESIDUAL, k_sor_residual, c->div, c->diag, c->pres, c->partials, c->state, 1);
}

void l0_true_residual_if_done(slosh_ctx *c)
{
    L0_LAUNCH(K_SOR_RESIDUAL, k_sor_residual, c->div, c->diag, c->pres, c->partials, c->state, 2);
}

} // namespace slosh
