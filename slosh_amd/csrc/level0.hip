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

// ---- column form of the per-iteration passes ----
// One wavefront per brick: lane = x + 8y owns the brick's column (x, y, 0..7) and walks its eight planes z = k.
// Every plane is one coalesced 256-byte load per array; a cell's neighbours in z are the planes above and below in
// registers (the planes of the bricks below and above for k = 0 and 7), in x and y the values of the lanes beside
// it (lane shuffles). The 32 cells beside the brick's x and y faces in a plane are loaded by lanes 0-31, one each,
// and handed to the lanes on the faces by two more shuffles. Compared with two cells per thread and every
// neighbour loaded, this moves each array's bytes through the caches once and issues a quarter of the load
// instructions, and the wavefront's FP64 sums need no barrier: each brick's partials come from one wave_sum.
constexpr uint32_t kColMaxBricks = 4; // most bricks (wavefronts) per workgroup (the launch picks 1, 2 or 4)
constexpr uint32_t kColMaxThreads = 64u * kColMaxBricks;

// Brick of this wavefront in a column pass over n bricks: blockDim.x / 64 consecutive bricks per workgroup,
// workgroups spread over the XCDs in contiguous ranges. `live` is false for the last workgroup's waves past n. The
// result does not depend on how many bricks a workgroup takes (each brick writes its own partials).
struct ColBrick {
    uint32_t e;
    bool live;
};
__device__ __forceinline__ ColBrick col_brick(const Blk &G, uint32_t n)
{
    ColBrick b;
    b.e = xcd_remap(G.b, G.nb) * (blockDim.x >> 6) + (threadIdx.x >> 6);
    b.live = b.e < n;
    return b;
}

// Element i of a pool array as the array's (wavefront-uniform) base plus a 32-bit byte offset: the loads and stores
// of a column then share one offset register with the planes as immediate offsets, instead of a 64-bit address
// per access. Pool arrays stay below 4 GiB (the context refuses larger pools).
template <typename T> __device__ __forceinline__ T pld(const T *p, uint32_t i)
{
    return *(const T *)((const char *)p + i * (uint32_t)sizeof(T));
}
template <typename T> __device__ __forceinline__ void pst(T *p, uint32_t i, T v)
{
    *(T *)((char *)p + i * (uint32_t)sizeof(T)) = v;
}


struct Col {
    uint32_t lane, x, y;
    uint32_t at;      // pool index of the lane's cell in plane 0 (plane k: + 64 k)
    uint32_t below;   // the cell under plane 0 (plane 7 of the -z brick) and
    uint32_t above;   // the cell over plane 7 (plane 0 of the +z brick)
    // The face cell this lane loads in plane 0 (+ 64 k for plane k): lanes 0-7 the -x brick's (7, lane), 8-15 the
    // +x brick's (0, lane - 8), 16-23 the -y brick's (lane - 16, 7), 24-31 the +y brick's (lane - 24, 0); lanes
    // 32-63 their own cell (an unused value: every lane issues the same loads, so none waits behind a branch).
    uint32_t h;
    uint32_t hbase, hx, hy; // pool base of that cell's brick and its local x and y
    int sx, sy;             // lanes holding this lane's x-face (x == 0: -x, x == 7: +x) and y-face neighbour values
};

__device__ __forceinline__ Col make_col(const Brick &B, uint32_t thread)
{
    Col C;
    const uint32_t lane = thread & 63u;
    C.lane = lane;
    C.x = lane & 7u;
    C.y = lane >> 3;
    C.at = B.base + lane;
    C.below = B.nb[4] + lane + 448u;
    C.above = B.nb[5] + lane;
    const uint32_t j = lane & 7u;
    switch (lane >> 3) {
    case 0: C.hbase = B.nb[0]; C.hx = 7u; C.hy = j; break;
    case 1: C.hbase = B.nb[1]; C.hx = 0u; C.hy = j; break;
    case 2: C.hbase = B.nb[2]; C.hx = j; C.hy = 7u; break;
    case 3: C.hbase = B.nb[3]; C.hx = j; C.hy = 0u; break;
    default: C.hbase = B.base; C.hx = C.x; C.hy = C.y; break;
    }
    C.h = C.hbase + C.hx + 8u * C.hy;
    C.sx = C.x == 0u ? (int)C.y : (C.x == 7u ? 8 + (int)C.y : (int)lane);
    C.sy = C.y == 0u ? 16 + (int)C.x : (C.y == 7u ? 24 + (int)C.x : (int)lane);
    return C;
}

// The six neighbour values (-x +x -y +y -z +z) of the lane's cell in one plane, whose own value is v; hv is the
// value of the face cell the lane loaded in that plane, lo / hi the values of the planes under and over it. Every
// lane of the wavefront must call it (the shuffles read other lanes).
__device__ __forceinline__ void col_nbrs(const Col &C, float v, float hv, float lo, float hi, float (&nb)[6])
{
    const float xm = __shfl_up(v, 1u, 64), xp = __shfl_down(v, 1u, 64);
    const float ym = __shfl_up(v, 8u, 64), yp = __shfl_down(v, 8u, 64);
    const float hxv = __shfl(hv, C.sx, 64), hyv = __shfl(hv, C.sy, 64);
    nb[0] = C.x == 0u ? hxv : xm;
    nb[1] = C.x == 7u ? hxv : xp;
    nb[2] = C.y == 0u ? hyv : ym;
    nb[3] = C.y == 7u ? hyv : yp;
    nb[4] = lo;
    nb[5] = hi;
}

// c ? a : b, with a and b first passed through an empty asm: a select between two elements of a register array
// would otherwise be folded into one load at a computed index (i + c), which sends the array to scratch memory.
template <typename T> __device__ __forceinline__ T pick(bool c, T a, T b)
{
    __asm__("" : "+v"(a), "+v"(b));
    return c ? a : b;
}

// The four in-plane neighbours (-x +x -y +y) of the lane's cell when every lane offers v and lanes 0-31 offer hv
// (their face cell's value); nb[4] and nb[5] are left to the caller.
template <typename T> __device__ __forceinline__ void plane_nbrs(const Col &C, T v, T hv, T (&nb)[6])
{
    const T xm = __shfl_up(v, 1u, 64), xp = __shfl_down(v, 1u, 64);
    const T ym = __shfl_up(v, 8u, 64), yp = __shfl_down(v, 8u, 64);
    const T hxv = __shfl(hv, C.sx, 64), hyv = __shfl(hv, C.sy, 64);
    nb[0] = C.x == 0u ? hxv : xm;
    nb[1] = C.x == 7u ? hxv : xp;
    nb[2] = C.y == 0u ? hyv : ym;
    nb[3] = C.y == 7u ? hyv : yp;
}

// Red-black passes take a column's planes in pairs (2j, 2j + 1): the lane's red cell of pair j is in plane 2j + par
// and its black cell in plane 2j + 1 - par (par: parity of x + y), so a lane makes one update per pair and colour.
// The lanes beside it have the other parity: the value a red cell needs from each of them is that lane's black cell
// of the same pair, and the other way round, so one shuffle per direction serves both planes of a pair. A face cell
// (lanes 0-31) has parity q = parity of its own x + y: its red cell of pair j is in plane 2j + q.

// Writes the brick's partials (the wavefront sums of v) to partials[e * NV + k].
template <int NV> __device__ __forceinline__ void col_partial(const double (&v)[NV], double *partials, uint32_t e,
                                                              uint32_t lane)
{
    static_assert(NV >= 1 && NV <= (int)kMaxBlockPartials, "partials buffer holds kMaxBlockPartials per brick");
    double a[NV];
    for (int k = 0; k < NV; ++k) a[k] = wave_sum(v[k]);
    if (lane == 0u)
        for (int k = 0; k < NV; ++k) partials[(size_t)e * NV + k] = a[k];
}

// The residual update of the iteration fused with the V-cycle's first level-0 sweeps (PCGUpdatePRCS with its
// fused red initialisation, then MGSmoothCS black): p += alpha d, r_out = r_in - alpha q, red z = r_out / diag,
// black z = (r_out + sum_n z_n) / diag where the red neighbours' z is recomputed from r_in, q and diag.
// kFirst: the solve's first update (no alpha yet: r_out = r_in, p unchanged); a template parameter so that the
// loads of q, d and p are not split into branches of their own.
template <bool kFirst>
__device__ __forceinline__ void update_black_body(const Blk &G, uint32_t n, const uint32_t *binfo, const uint8_t *fmask,
                                                              const float *diag, double *pres, const float *r_in,
                                                              float *r_out, float *z, const float *d, const float *q,
                                                              const SolverState *st)
{
    constexpr bool first = kFirst;
    const ColBrick cb = col_brick(G, n);
    const Brick B = load_brick(binfo, cb.live ? cb.e : 0u);
    const uint32_t done = st->done;
    const double alpha = first ? 0.0 : st->alpha;
    issue_point();
    if (done == 1u) return; // done == 2: the solve's last update, made before every other pass stops
    if (!cb.live) return;
    const Col C = make_col(B, threadIdx.x);
    // the red value r / diag of a neighbour cell after the update
    auto red_of = [&](uint32_t i) {
        const float ra = pld(r_in, i);
        const float g = pld(diag, i);
        const float qv = first ? 0.0f : pld(q, i);
        const float r = first ? ra : (float)((double)ra - alpha * (double)qv);
        return r / (g + kDiagEps);
    };
    // pairs of planes: the red values of the face cells and of the cell under (par) or over the column
    const bool par = ((C.x + C.y) & 1u) != 0u, hq = ((C.hx + C.hy) & 1u) != 0u;
    const uint32_t red0 = C.at + (par ? 64u : 0u), blk0 = C.at + (par ? 0u : 64u), hred0 = C.h + (hq ? 64u : 0u);
    uint32_t m[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) m[k] = pld(fmask, C.at + 64u * k);
    // the pressure update first: its values are dead once stored, so they do not add to the registers the
    // multigrid part holds
    if (!first) {
        double pv[8];
        float dv[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            pv[k] = pld(pres, C.at + 64u * k);
            dv[k] = pld(d, C.at + 64u * k);
        }
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (m[k] & kSelfFluid) pst(pres, C.at + 64u * k, pv[k] + alpha * (double)dv[k]);
    }
    issue_point();
    const float zext = red_of(par ? C.below : C.above);
    float hr[4], ra[8], qv[8], g[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t i = C.at + 64u * k;
        ra[k] = pld(r_in, i);
        g[k] = pld(diag, i);
        qv[k] = first ? 0.0f : pld(q, i);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) hr[j] = red_of(hred0 + 128u * j);
    float rn[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) rn[k] = first ? ra[k] : (float)((double)ra[k] - alpha * (double)qv[k]);
    float zr[4]; // the red cells' z
#pragma unroll
    for (int j = 0; j < 4; ++j) zr[j] = pick(par, rn[2 * j + 1], rn[2 * j]) / (pick(par, g[2 * j + 1], g[2 * j]) + kDiagEps);
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if (m[k] & kSelfFluid) pst(r_out, C.at + 64u * k, rn[k]);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t mr = pick(par, m[2 * j + 1], m[2 * j]), mb = pick(par, m[2 * j], m[2 * j + 1]);
        if (mr & kSelfFluid) pst(z, red0 + 128u * j, zr[j]);
        float nb[6];
        plane_nbrs(C, zr[j], hr[j], nb);
        nb[4] = pick(par, j == 0 ? zext : zr[j - 1], zr[j]);
        nb[5] = pick(par, zr[j], j == 3 ? zext : zr[j + 1]);
        float s = 0.0f;
        for (int t = 0; t < 6; ++t)
            if (mb >> t & 1u) s += nb[t];
        if (mb & kSelfFluid)
            pst(z, blk0 + 128u * j, (pick(par, rn[2 * j], rn[2 * j + 1]) + s) / (pick(par, g[2 * j], g[2 * j + 1]) + kDiagEps));
    }
}

// Restriction into pooled level 1: the residual b - A x of every fluid cell (neighbours' x through the column
// form), then b_c = 1/4 of the fluid children's sum for the brick's 4^3 level-1 cells, in the children's order
// (x, then y, then z fastest to slowest), and x_c = b_c / diag_c on red cells (fused red sweep). The lane with even
// x and y sums its 2x2 group of lanes over the plane pair.
__device__ __forceinline__ void restrict0_body(const Blk &G, uint32_t n, const uint32_t *binfo, const uint8_t *fmask,
                                                           const float *diag, const float *xf, const float *bf,
                                                           PLevel L1, const SolverState *st)
{
    const ColBrick cb = col_brick(G, n);
    const Brick B = load_brick(binfo, cb.live ? cb.e : 0u);
    const uint32_t done = st->done;
    const double mg_scale = st->mg_scale;
    issue_point();
    if (done) return;
    if (!cb.live) return;
    const Col C = make_col(B, threadIdx.x);
    float xv[10], hv[8], bv[8], gv[8];
    uint32_t m[8];
    xv[0] = pld(xf, C.below);
    xv[9] = pld(xf, C.above);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t i = C.at + 64u * k;
        m[k] = pld(fmask, i);
        xv[k + 1] = pld(xf, i);
        bv[k] = pld(bf, i);
        gv[k] = pld(diag, i);
        hv[k] = pld(xf, C.h + 64u * k);
    }
    const bool writer = ((C.x | C.y) & 1u) == 0u;
    const uint32_t c0 = (B.base >> 3) + (C.x >> 1) + 4u * (C.y >> 1);
    uint32_t info[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) info[p] = writer ? L1.info[c0 + 16u * p] : 0u;
    float sum[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        float nb[6];
        col_nbrs(C, xv[k + 1], hv[k], xv[k], xv[k + 2], nb);
        const bool fluid = (m[k] & kSelfFluid) != 0u;
        float res = 0.0f;
        if (fluid) {
            float s = 0.0f;
            for (int n = 0; n < 6; ++n)
                if (m[k] >> n & 1u) s += nb[n];
            const float dg = gv[k] + kDiagEps;
            res = bv[k] - (dg * xv[k + 1] - s);
        }
        // children of this plane in order (x, y), (x+1, y), (x, y+1), (x+1, y+1); only fluid children count
        const uint64_t fl = __ballot(fluid);
        const float r1 = __shfl_down(res, 1u, 64), r8 = __shfl_down(res, 8u, 64), r9 = __shfl_down(res, 9u, 64);
        const uint32_t L = C.lane;
        float &acc = sum[k >> 1];
        if (fl >> L & 1u) acc += res;
        if (writer && (fl >> (L + 1u) & 1u)) acc += r1;
        if (writer && (fl >> (L + 8u) & 1u)) acc += r8;
        if (writer && (fl >> (L + 9u) & 1u)) acc += r9;
    }
    if (!writer) return;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const uint32_t ci = c0 + 16u * p;
        if (!(info[p] & kInfoSelf)) {
            L1.b[ci] = 0.0f;
            L1.x[ci] = 0.0f;
            continue;
        }
        float b = 0.25f * sum[p];
        if (L1.b.half) b *= (float)mg_scale; // FP16 levels hold the scaled cycle (the V-cycle is linear)
        L1.b[ci] = b;
        L1.x[ci] = (info[p] & kInfoRed) ? b / info_diag(info[p]) : 0.0f;
    }
}

// Prolongation from pooled level 1 (x1) fused with the black sweep: every cell's w = z + x1(parent) (the parent
// contributes only when it is fluid); red z1 = w, black z1 = (r + sum_n w_n) / diag.
__device__ __forceinline__ void prolong_black_body(const Blk &G, uint32_t n, const uint32_t *binfo, const uint8_t *fmask,
                                                               const float *diag, const float *r, const float *z,
                                                               float *z1, PLevel L1, const SolverState *st)
{
    const ColBrick cb = col_brick(G, n);
    const Brick B = load_brick(binfo, cb.live ? cb.e : 0u);
    const uint32_t done = st->done;
    const double mg_scale = st->mg_scale;
    issue_point();
    if (done) return;
    if (!cb.live) return;
    const Col C = make_col(B, threadIdx.x);
    const float unscale = L1.x1.half ? (float)(1.0 / mg_scale) : 1.0f;
    // w of the cell at pool index i whose level-1 parent is at pooled index pi
    auto w_of = [&](uint32_t i, uint32_t pi) {
        const float v = pld(z, i), px = L1.x1[pi];
        const uint32_t info = L1.info[pi];
        return (info & kInfoSelf) ? v + px * unscale : v;
    };
    // level-1 parent of local cell (x, y, z) of the brick whose pool base is base
    auto parent = [](uint32_t base, uint32_t x, uint32_t y, uint32_t zz) {
        return (base >> 3) + (x >> 1) + 4u * (y >> 1) + 16u * (zz >> 1);
    };
    // pairs of planes (the two planes of a pair share their level-1 parents): w is needed on the red cells only,
    // r and the diagonal on the black ones
    const bool par = ((C.x + C.y) & 1u) != 0u, q = ((C.hx + C.hy) & 1u) != 0u;
    const uint32_t red0 = C.at + (par ? 64u : 0u), blk0 = C.at + (par ? 0u : 64u), hred0 = C.h + (q ? 64u : 0u);
    // the red cell under the column (par) or over it
    const float wext = par ? w_of(C.below, parent(B.nb[4], C.x, C.y, 7u)) : w_of(C.above, parent(B.nb[5], C.x, C.y, 0u));
    float wr[4], hw[4], rb[4], gb[4];
    uint32_t m[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) m[k] = pld(fmask, C.at + 64u * k);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        wr[j] = w_of(red0 + 128u * j, parent(B.base, C.x, C.y, 2u * j));
        hw[j] = w_of(hred0 + 128u * j, parent(C.hbase, C.hx, C.hy, 2u * j));
        rb[j] = pld(r, blk0 + 128u * j);
        gb[j] = pld(diag, blk0 + 128u * j);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t mr = pick(par, m[2 * j + 1], m[2 * j]), mb = pick(par, m[2 * j], m[2 * j + 1]);
        if (mr & kSelfFluid) pst(z1, red0 + 128u * j, wr[j]);
        float nb[6];
        plane_nbrs(C, wr[j], hw[j], nb);
        nb[4] = pick(par, j == 0 ? wext : wr[j - 1], wr[j]);
        nb[5] = pick(par, wr[j], j == 3 ? wext : wr[j + 1]);
        float s = 0.0f;
        for (int t = 0; t < 6; ++t)
            if (mb >> t & 1u) s += nb[t];
        if (mb & kSelfFluid) pst(z1, blk0 + 128u * j, (rb[j] + s) / (gb[j] + kDiagEps));
    }
}

// Red sweep z1 = (r + sum_n z1_n) / diag and the partials of r.z1 over all fluid cells.
__device__ __forceinline__ void red_dot_body(const Blk &G, uint32_t n, const uint32_t *binfo, const uint8_t *fmask,
                                                         const float *diag, const float *r, float *z1,
                                                         double *partials, const SolverState *st)
{
    const ColBrick cb = col_brick(G, n);
    const Brick B = load_brick(binfo, cb.live ? cb.e : 0u);
    const uint32_t done = st->done;
    issue_point();
    if (done) return;
    if (!cb.live) return;
    const Col C = make_col(B, threadIdx.x);
    // pairs of planes: the red cell of pair j at red0 + 128 j, the black one at blk0 + 128 j, the face cell's black
    // one at hblk0 + 128 j
    const bool par = ((C.x + C.y) & 1u) != 0u, q = ((C.hx + C.hy) & 1u) != 0u;
    const uint32_t red0 = C.at + (par ? 64u : 0u), blk0 = C.at + (par ? 0u : 64u), hblk0 = C.h + (q ? 0u : 64u);
    float rv[8], zb[4], gr[4], hb[4];
    uint32_t m[8];
    const float zext = pld(z1, par ? C.above : C.below); // the black cell over the column (par) or under it
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        m[k] = pld(fmask, C.at + 64u * k);
        rv[k] = pld(r, C.at + 64u * k);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        zb[j] = pld(z1, blk0 + 128u * j);
        gr[j] = pld(diag, red0 + 128u * j);
        hb[j] = pld(z1, hblk0 + 128u * j);
    }
    double v[1] = {0.0};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        float nb[6];
        plane_nbrs(C, zb[j], hb[j], nb);
        nb[4] = pick(par, zb[j], j == 0 ? zext : zb[j - 1]);
        nb[5] = pick(par, j == 3 ? zext : zb[j + 1], zb[j]);
        const uint32_t mr = pick(par, m[2 * j + 1], m[2 * j]), mb = pick(par, m[2 * j], m[2 * j + 1]);
        const float rr = pick(par, rv[2 * j + 1], rv[2 * j]), rb = pick(par, rv[2 * j], rv[2 * j + 1]);
        float s = 0.0f;
        for (int t = 0; t < 6; ++t)
            if (mr >> t & 1u) s += nb[t];
        const float xr = (rr + s) / (gr[j] + kDiagEps);
        if (mr & kSelfFluid) pst(z1, red0 + 128u * j, xr);
        // the dot over the column in plane order (a skipped cell adds +0.0, which leaves the sum's bits unchanged)
        const double c_r = (mr & kSelfFluid) ? (double)rr * (double)xr : 0.0;
        const double c_b = (mb & kSelfFluid) ? (double)rb * (double)zb[j] : 0.0;
        v[0] += pick(par, c_b, c_r);
        v[0] += pick(par, c_r, c_b);
    }
    col_partial<1>(v, partials, cb.e, C.lane);
}

// d_out = z1 + beta d_in (beta = 0 for the first direction) fused with q = A d_out: the neighbours' new
// direction is recomputed from z1 and d_in. The stencil sum is FP64 (for smooth d the diagonal and neighbour
// terms cancel, and an FP32 sum would put a large relative error into q that the recurrence residual never
// sees); partials d.q, and for the stopping test of the coming update r - alpha q: r.q, q.q and r.r (q as stored,
// in FP32). kFirst: the first direction d = z1 (a template parameter so that the loads of d_in are not split into
// branches of their own).
template <bool kFirst>
__device__ __forceinline__ void direction_body(const Blk &G, uint32_t n, const uint32_t *binfo, const uint8_t *fmask,
                                                           const float *diag, const float *z1, const float *d_in,
                                                           float *d_out, float *q, const float *r, double *partials,
                                                           const SolverState *st)
{
    constexpr bool first = kFirst;
    const ColBrick cb = col_brick(G, n);
    const Brick B = load_brick(binfo, cb.live ? cb.e : 0u);
    const uint32_t done = st->done;
    const double beta = first ? 0.0 : st->beta;
    issue_point();
    if (done) return;
    if (!cb.live) return;
    const Col C = make_col(B, threadIdx.x);
    auto dir = [&](float zv, float dv) { return first ? zv : (float)((double)zv + beta * (double)dv); };
    float dd[10], hv[8], gv[8], rv[8];
    uint32_t m[8];
    dd[0] = dir(pld(z1, C.below), first ? 0.0f : pld(d_in, C.below));
    dd[9] = dir(pld(z1, C.above), first ? 0.0f : pld(d_in, C.above));
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t i = C.at + 64u * k;
        m[k] = pld(fmask, i);
        gv[k] = pld(diag, i);
        rv[k] = pld(r, i);
        dd[k + 1] = dir(pld(z1, i), first ? 0.0f : pld(d_in, i));
        hv[k] = dir(pld(z1, C.h + 64u * k), first ? 0.0f : pld(d_in, C.h + 64u * k));
    }
    double v[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        float nb[6];
        col_nbrs(C, dd[k + 1], hv[k], dd[k], dd[k + 2], nb);
        if (!(m[k] & kSelfFluid)) continue;
        const uint32_t i = C.at + 64u * k;
        const float dv = dd[k + 1];
        pst(d_out, i, dv);
        double s = 0.0;
        for (int n = 0; n < 6; ++n)
            if (m[k] >> n & 1u) s += (double)nb[n];
        const float dg = gv[k] + kDiagEps;
        const double qd = (double)dg * (double)dv - s;
        const float qf = (float)qd;
        pst(q, i, qf);
        v[0] += (double)dv * qd;
        v[1] += (double)rv[k] * (double)qf;
        v[2] += (double)qf * (double)qf;
        v[3] += (double)rv[k] * (double)rv[k];
    }
    col_partial<4>(v, partials, cb.e, C.lane);
}

// ---- passes on the FP64 pressure, in column form ----
// The FP64 counterpart of col_nbrs.
__device__ __forceinline__ void col_nbrs64(const Col &C, double v, double hv, double lo, double hi, double (&nb)[6])
{
    const double xm = __shfl_up(v, 1u, 64), xp = __shfl_down(v, 1u, 64);
    const double ym = __shfl_up(v, 8u, 64), yp = __shfl_down(v, 8u, 64);
    const double hxv = __shfl(hv, C.sx, 64), hyv = __shfl(hv, C.sy, 64);
    nb[0] = C.x == 0u ? hxv : xm;
    nb[1] = C.x == 7u ? hxv : xp;
    nb[2] = C.y == 0u ? hyv : ym;
    nb[3] = C.y == 7u ? hyv : yp;
    nb[4] = lo;
    nb[5] = hi;
}

// A column's pressure planes (with the planes under and over it), its face cells' pressure, and its divergence,
// diagonal and stencil masks.
struct PresCol {
    double p[10], h[8];
    float div[8], g[8];
    uint32_t m[8];
};
__device__ __forceinline__ void load_pres_col(const Col &C, const uint8_t *fmask, const float *div, const float *diag,
                                              const double *pres, PresCol &P)
{
    P.p[0] = pld(pres, C.below);
    P.p[9] = pld(pres, C.above);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t i = C.at + 64u * k;
        P.m[k] = pld(fmask, i);
        P.div[k] = pld(div, i);
        P.g[k] = pld(diag, i);
        P.p[k + 1] = pld(pres, i);
        P.h[k] = pld(pres, C.h + 64u * k);
    }
}

// r = b - A p (FP64 p) over every cell of an allocated brick (0 off the fluid cells); partials ||b||^2, ||r||^2 and
// the fluid cell count.
__device__ __forceinline__ void pcg_init_body(const Blk &G, uint32_t n, const uint32_t *binfo, const uint8_t *fmask, const float *div,
                                                          const float *diag, const double *pres, float *r,
                                                          double *partials)
{
    const ColBrick cb = col_brick(G, n);
    if (!cb.live) return;
    const Col C = make_col(load_brick(binfo, cb.e), threadIdx.x);
    PresCol P;
    load_pres_col(C, fmask, div, diag, pres, P);
    double v[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        double nb[6];
        col_nbrs64(C, P.p[k + 1], P.h[k], P.p[k], P.p[k + 2], nb);
        const uint32_t i = C.at + 64u * k;
        if (!(P.m[k] & kSelfFluid)) {
            pst(r, i, 0.0f);
            continue;
        }
        v[2] += 1.0;
        double s = 0.0;
        for (int n = 0; n < 6; ++n)
            if (P.m[k] >> n & 1u) s += nb[n];
        const float dg = P.g[k] + kDiagEps;
        const float ri = (float)(-(double)P.div[k] - ((double)dg * P.p[k + 1] - s));
        pst(r, i, ri);
        v[0] += (double)P.div[k] * (double)P.div[k];
        v[1] += (double)ri * (double)ri;
    }
    col_partial<3>(v, partials, cb.e, C.lane);
}

// One SOR update of a cell's FP64 pressure, p + w ((sum of the fluid neighbours' p - div) / diag - p): bit n of
// the stencil mask m says neighbour n (-x +x -y +y -z +z, summed in that order) is fluid.
__device__ __forceinline__ double sor_update(uint32_t m, const double (&nb)[6], float div, float diag, double old,
                                             double w)
{
    double s = 0.0;
    for (int n = 0; n < 6; ++n)
        if (m >> n & 1u) s += nb[n];
    const double dg = (double)(diag + kDiagEps);
    const double gs = (s - (double)div) / dg;
    return old + w * (gs - old);
}

// A red sweep and then a black sweep of red-black SOR in one pass, bit for bit the two sweeps. A red cell's
// neighbours are black and a black cell's red, so each wavefront, from the old values alone, updates its brick's red
// cells and the red cells just outside the brick that its black cells read (the cells beside its x and y faces and
// the planes under and over it, each from its own old neighbours exactly as its own brick's wavefront updates it),
// then its black cells; only the brick's own fluid cells are stored. The updates are made in place in the
// registers: every red update reads black values only, which the red half does not change, and every black update
// reads red values only. The pass reads one array and writes another (the neighbouring bricks' old values must stay
// readable while they are updated); the cells it does not write (not fluid) are never used.
//
// The planes are taken in pairs (2j, 2j + 1): a lane's column has exactly one red and one black cell in each pair
// (red in plane 2j + par, par the parity of x + y), so a lane makes one update per pair and colour, and the values
// its neighbours need from it are all its cell of the other colour in that pair (the lanes beside it have the other
// parity): one shuffle per direction serves both planes of a pair.
//
// Lane L < 32 holds face cell L (its plane k at C.h + 64 k; red planes k = 2 m + q, q the parity of the cell's
// x + y) and loads, for those red planes, the cell's mask, divergence and diagonal and the two neighbours that no
// other lane holds: the one beyond it (`deep`) and, at the ends of a face (j = 0, 7), the corner cell beside it in
// an edge brick (`end`). Its neighbours along the face are lanes L -+ 1, the brick's own cell beside it is lane
// `own`. Of the planes under and over the brick a lane updates the one whose cell is red (under: x + y odd, over:
// even), and the face cell under or over face cell L that this needs is the one that cell's own red update needs at
// the end of its column (`zf`). The per-lane addresses are formed with selects (a switch over the face would be
// compiled to a table in scratch memory).
__device__ __forceinline__ void sor_pair_body(const Blk &G, uint32_t n, const uint32_t *binfo, const uint8_t *fmask,
                                              const float *div, const float *diag, const double *pres, double *pout,
                                              double w, const SolverState *st)
{
    constexpr int NP = 8;
    const ColBrick cb = col_brick(G, n);
    const uint32_t *row = binfo + (size_t)(cb.live ? cb.e : 0u) * kBinfoWords;
    const Brick B = load_brick(binfo, cb.live ? cb.e : 0u);
    // the corner bricks beside the z edges (word (dx+1) + 3(dy+1) + 9(dz+1)) and the edge bricks under and over the
    // faces (-x +x -y +y)
    const uint32_t e_mm = row[9] * kBrickCells, e_pm = row[11] * kBrickCells, e_mp = row[15] * kBrickCells,
                   e_pp = row[17] * kBrickCells;
    const uint32_t zl0 = row[3] * kBrickCells, zl1 = row[5] * kBrickCells, zl2 = row[1] * kBrickCells,
                   zl3 = row[7] * kBrickCells;
    const uint32_t zh0 = row[21] * kBrickCells, zh1 = row[23] * kBrickCells, zh2 = row[19] * kBrickCells,
                   zh3 = row[25] * kBrickCells;
    const uint32_t done = st->done;
    issue_point();
    if (done) return;
    if (!cb.live) return;
    const Col C = make_col(B, threadIdx.x);
    const uint32_t L = C.lane, f = (L >> 3) & 3u, j = L & 7u;
    const bool par = ((C.x + C.y) & 1u) != 0u, q = ((C.hx + C.hy) & 1u) != 0u;
    const uint32_t at = C.at, below = C.below, above = C.above, hat = C.h;
    // face f of lane L < 32: 0 -x, 1 +x (f & 1: the + side), 2 -y, 3 +y (f & 2: a y face); j runs along the face
    const bool fp = (f & 1u) != 0u, fy = (f & 2u) != 0u;
    const bool face = L < 32u;
    const uint32_t hl = C.hx + 8u * C.hy; // the face cell's index in its own brick
    // the cell beyond it, one further out from the brick
    const uint32_t deep = face ? (uint32_t)((int)hat + (fy ? (fp ? 8 : -8) : (fp ? 1 : -1))) : at;
    // the brick's own cell beside it
    const int own = face ? (int)(fy ? j + (fp ? 56u : 0u) : 8u * j + (fp ? 7u : 0u)) : (int)L;
    // the corner cell at the end of the face: x side -1 or 8, y side -1 or 8
    const bool cx8 = fy ? j == 7u : fp, cy8 = fy ? fp : j == 7u;
    const uint32_t corner = cy8 ? (cx8 ? e_pp : e_mp + 7u) : (cx8 ? e_pm + 56u : e_mm + 63u);
    const uint32_t end = face && (j == 0u || j == 7u) ? corner : at;
    // the face cell's column continues under plane 0 in the edge brick under its brick, over plane 7 over it
    const uint32_t zlb = fy ? (fp ? zl3 : zl2) : (fp ? zl1 : zl0), zhb = fy ? (fp ? zh3 : zh2) : (fp ? zh1 : zh0);
    const uint32_t zlo = face ? zlb + hl + 448u : at, zhi = face ? zhb + hl : at;
    double p[NP + 2], h[NP];
    float dv[NP], gv[NP];
    uint32_t mv[NP];
    p[0] = pld(pres, below);
    p[NP + 1] = pld(pres, above);
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const uint32_t i = at + 64u * k;
        mv[k] = pld(fmask, i);
        dv[k] = pld(div, i);
        gv[k] = pld(diag, i);
        p[k + 1] = pld(pres, i);
        h[k] = pld(pres, hat + 64u * k);
    }
    // the red cell under or over the brick: mask, divergence, diagonal, the cell beyond it; the face cell's end
    const uint32_t zc = par ? below : above;
    const uint32_t zm_ = pld(fmask, zc);
    const float zd = pld(div, zc), zg = pld(diag, zc);
    const double zbeyond = pld(pres, par ? below - 64u : above + 64u);
    const double zf = pld(pres, q ? zhi : zlo);
    // the face cell's red planes
    constexpr int NJ = NP / 2;
    uint32_t hm[NJ];
    float hd[NJ], hg[NJ];
    double hdeep[NJ], hend[NJ];
#pragma unroll
    for (int m = 0; m < NJ; ++m) {
        const uint32_t o = 64u * (2u * (uint32_t)m + q);
        hm[m] = pld(fmask, hat + o);
        hd[m] = pld(div, hat + o);
        hg[m] = pld(diag, hat + o);
        hdeep[m] = pld(pres, deep + o);
        hend[m] = pld(pres, end + o);
    }
    issue_point(); // every load is in flight before the first update waits for one
    // red: the brick's own cells, one per pair (plane 2j + par); the neighbours are the black cells of the pair
    double bv[NJ]; // the lane's black value of each pair (the own-lane neighbour of a red face cell)
#pragma unroll
    for (int jj = 0; jj < NJ; ++jj) {
        const int k0 = 2 * jj;
        bv[jj] = pick(par, p[k0 + 1], p[k0 + 2]);
        double nb[6];
        plane_nbrs(C, bv[jj], pick(q, h[k0], h[k0 + 1]), nb);
        nb[4] = pick(par, p[k0 + 1], p[k0]);
        nb[5] = pick(par, p[k0 + 3], p[k0 + 2]);
        const uint32_t m = pick(par, mv[k0 + 1], mv[k0]);
        const double old = pick(par, p[k0 + 2], p[k0 + 1]);
        const double v = sor_update(m, nb, pick(par, dv[k0 + 1], dv[k0]), pick(par, gv[k0 + 1], gv[k0]), old, w);
        const bool upd = (m & kSelfFluid) != 0u;
        p[k0 + 1] = upd && !par ? v : p[k0 + 1];
        p[k0 + 2] = upd && par ? v : p[k0 + 2];
    }
    // red: the face cells (lanes 0-31), plane k = 2m + q; the lanes beside one along its face offer their black
    // value of the pair, which is in plane k
#pragma unroll
    for (int m = 0; m < NJ; ++m) {
        const double hb = pick(q, h[2 * m], h[2 * m + 1]);
        const double a = __shfl_up(hb, 1u, 64), b = __shfl_down(hb, 1u, 64);
        const double ov = __shfl(bv[m], own, 64);
        const double lm = j == 0u ? hend[m] : a; // the face neighbour at j - 1
        const double lp = j == 7u ? hend[m] : b; // and at j + 1
        const double hc = pick(q, h[2 * m + 1], h[2 * m]);
        const double zm = pick(q, h[2 * m], m == 0 ? zf : h[2 * m - 1]);
        const double zp = pick(q, m == NJ - 1 ? zf : h[2 * m + 2], h[2 * m + 1]);
        double nb[6];
        if (f < 2u) {
            nb[0] = f == 0u ? hdeep[m] : ov;
            nb[1] = f == 0u ? ov : hdeep[m];
            nb[2] = lm;
            nb[3] = lp;
        } else {
            nb[0] = lm;
            nb[1] = lp;
            nb[2] = f == 2u ? hdeep[m] : ov;
            nb[3] = f == 2u ? ov : hdeep[m];
        }
        nb[4] = zm;
        nb[5] = zp;
        const bool upd = L < 32u && (hm[m] & kSelfFluid);
        const double v = sor_update(hm[m], nb, hd[m], hg[m], hc, w);
        h[2 * m] = upd && !q ? v : h[2 * m];
        h[2 * m + 1] = upd && q ? v : h[2 * m + 1];
    }
    // red: the cell under the brick (x + y odd) or over it (even); its in-plane neighbours are black, offered by the
    // lanes beside it (the plane under by even lanes, the plane over by odd ones) and by the face lanes (their `zf`)
    {
        double nb[6];
        plane_nbrs(C, pick(par, p[NP + 1], p[0]), zf, nb);
        nb[4] = pick(par, zbeyond, p[NP]);
        nb[5] = pick(par, p[1], zbeyond);
        const double old = pick(par, p[0], p[NP + 1]);
        const double v = sor_update(zm_, nb, zd, zg, old, w);
        const bool upd = (zm_ & kSelfFluid) != 0u;
        p[0] = upd && par ? v : p[0];
        p[NP + 1] = upd && !par ? v : p[NP + 1];
    }
    // black: the brick's own cells, one per pair (plane 2j + 1 - par); the neighbours are the red cells of the pair
#pragma unroll
    for (int jj = 0; jj < NJ; ++jj) {
        const int k0 = 2 * jj;
        double nb[6];
        plane_nbrs(C, pick(par, p[k0 + 2], p[k0 + 1]), pick(q, h[k0 + 1], h[k0]), nb);
        nb[4] = pick(par, p[k0], p[k0 + 1]);
        nb[5] = pick(par, p[k0 + 2], p[k0 + 3]);
        const uint32_t m = pick(par, mv[k0], mv[k0 + 1]);
        const double old = pick(par, p[k0 + 1], p[k0 + 2]);
        const double v = sor_update(m, nb, pick(par, dv[k0], dv[k0 + 1]), pick(par, gv[k0], gv[k0 + 1]), old, w);
        const bool upd = (m & kSelfFluid) != 0u;
        p[k0 + 1] = upd && par ? v : p[k0 + 1];
        p[k0 + 2] = upd && !par ? v : p[k0 + 2];
    }
#pragma unroll
    for (int k = 0; k < NP; ++k)
        if (mv[k] & kSelfFluid) pst(pout, at + 64u * k, p[k + 1]);
}

// Partials ||b||^2 and ||b - A p||^2: mode 0 the SOR convergence check (skipped once the solver is done), 1 always,
// 2 only once the solver is done (the true-residual check of a finished MGPCG solve).
__device__ __forceinline__ void sor_residual_body(const Blk &G, uint32_t n, const uint32_t *binfo, const uint8_t *fmask,
                                                              const float *div, const float *diag, const double *pres,
                                                              double *partials, const SolverState *st, int mode)
{
    const int32_t done = st->done;
    if ((mode == 0 && done) || (mode == 2 && !done)) return;
    const ColBrick cb = col_brick(G, n);
    if (!cb.live) return;
    const Col C = make_col(load_brick(binfo, cb.e), threadIdx.x);
    PresCol P;
    load_pres_col(C, fmask, div, diag, pres, P);
    double v[2] = {0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        double nb[6];
        col_nbrs64(C, P.p[k + 1], P.h[k], P.p[k], P.p[k + 2], nb);
        if (!(P.m[k] & kSelfFluid)) continue;
        double s = 0.0;
        for (int n = 0; n < 6; ++n)
            if (P.m[k] >> n & 1u) s += nb[n];
        const double dg = (double)(P.g[k] + kDiagEps);
        const double ri = -(double)P.div[k] - (dg * P.p[k + 1] - s);
        v[0] += (double)P.div[k] * (double)P.div[k];
        v[1] += ri * ri;
    }
    col_partial<2>(v, partials, cb.e, C.lane);
}

// ---- single-tank launches of the column passes ----
template <bool kFirst>
__global__ __launch_bounds__(kColMaxThreads) void k_update_black(uint32_t n, const uint32_t *binfo, const uint8_t *fmask,
                                                              const float *diag, double *pres, const float *r_in,
                                                              float *r_out, float *z, const float *d, const float *q,
                                                              const SolverState *st)
{
    update_black_body<kFirst>(this_blk(), n, binfo, fmask, diag, pres, r_in, r_out, z, d, q, st);
}

__global__ __launch_bounds__(kColMaxThreads) void k_restrict0(uint32_t n, const uint32_t *binfo, const uint8_t *fmask,
                                                           const float *diag, const float *xf, const float *bf,
                                                           PLevel L1, const SolverState *st)
{
    restrict0_body(this_blk(), n, binfo, fmask, diag, xf, bf, L1, st);
}

__global__ __launch_bounds__(kColMaxThreads) void k_prolong_black(uint32_t n, const uint32_t *binfo, const uint8_t *fmask,
                                                               const float *diag, const float *r, const float *z,
                                                               float *z1, PLevel L1, const SolverState *st)
{
    prolong_black_body(this_blk(), n, binfo, fmask, diag, r, z, z1, L1, st);
}

__global__ __launch_bounds__(kColMaxThreads) void k_red_dot(uint32_t n, const uint32_t *binfo, const uint8_t *fmask,
                                                         const float *diag, const float *r, float *z1,
                                                         double *partials, const SolverState *st)
{
    red_dot_body(this_blk(), n, binfo, fmask, diag, r, z1, partials, st);
}

template <bool kFirst>
__global__ __launch_bounds__(kColMaxThreads) void k_direction(uint32_t n, const uint32_t *binfo, const uint8_t *fmask,
                                                           const float *diag, const float *z1, const float *d_in,
                                                           float *d_out, float *q, const float *r, double *partials,
                                                           const SolverState *st)
{
    direction_body<kFirst>(this_blk(), n, binfo, fmask, diag, z1, d_in, d_out, q, r, partials, st);
}

__global__ __launch_bounds__(kColMaxThreads) void k_pcg_init(uint32_t n, const uint32_t *binfo, const uint8_t *fmask, const float *div,
                                                          const float *diag, const double *pres, float *r,
                                                          double *partials)
{
    pcg_init_body(this_blk(), n, binfo, fmask, div, diag, pres, r, partials);
}

__global__ __launch_bounds__(kColMaxThreads) void k_sor_pair(uint32_t n, const uint32_t *binfo, const uint8_t *fmask,
                                                          const float *div, const float *diag, const double *pres,
                                                          double *pout, double w, const SolverState *st)
{
    sor_pair_body(this_blk(), n, binfo, fmask, div, diag, pres, pout, w, st);
}

__global__ __launch_bounds__(kColMaxThreads) void k_sor_residual(uint32_t n, const uint32_t *binfo, const uint8_t *fmask,
                                                              const float *div, const float *diag, const double *pres,
                                                              double *partials, const SolverState *st, int mode)
{
    sor_residual_body(this_blk(), n, binfo, fmask, div, diag, pres, partials, st, mode);
}

// ---- the column passes over a batch of tanks (one launch for all of them) ----
template <bool kFirst>
__global__ __launch_bounds__(kColMaxThreads) void kb_update_black(const BatchTank *T, const uint32_t *range,
                                                                  uint32_t nt, int ri, int ro, int di)
{
    Blk G;
    const BatchTank &B = T[batch_tank(range, nt, blockIdx.x, G)];
    update_black_body<kFirst>(G, B.nsolve, B.binfo_f, B.fmask, B.diag, B.pres, B.r[ri], B.r[ro], B.z, B.d[di], B.q,
                              B.state);
}

__global__ __launch_bounds__(kColMaxThreads) void kb_restrict0(const BatchTank *T, const uint32_t *range,
                                                               uint32_t nt, int ri)
{
    Blk G;
    const BatchTank &B = T[batch_tank(range, nt, blockIdx.x, G)];
    restrict0_body(G, B.nsolve, B.binfo_f, B.fmask, B.diag, B.z, B.r[ri], B.L1, B.state);
}

__global__ __launch_bounds__(kColMaxThreads) void kb_prolong_black(const BatchTank *T, const uint32_t *range,
                                                                   uint32_t nt, int ri)
{
    Blk G;
    const BatchTank &B = T[batch_tank(range, nt, blockIdx.x, G)];
    prolong_black_body(G, B.nsolve, B.binfo_f, B.fmask, B.diag, B.r[ri], B.z, B.z1, B.L1, B.state);
}

__global__ __launch_bounds__(kColMaxThreads) void kb_red_dot(const BatchTank *T, const uint32_t *range, uint32_t nt,
                                                             int ri)
{
    Blk G;
    const BatchTank &B = T[batch_tank(range, nt, blockIdx.x, G)];
    red_dot_body(G, B.nsolve, B.binfo_f, B.fmask, B.diag, B.r[ri], B.z1, B.partials, B.state);
}

template <bool kFirst>
__global__ __launch_bounds__(kColMaxThreads) void kb_direction(const BatchTank *T, const uint32_t *range,
                                                               uint32_t nt, int din, int dout, int ri)
{
    Blk G;
    const BatchTank &B = T[batch_tank(range, nt, blockIdx.x, G)];
    direction_body<kFirst>(G, B.nsolve, B.binfo_f, B.fmask, B.diag, B.z1, B.d[din], B.d[dout], B.q, B.r[ri],
                           B.partials, B.state);
}

__global__ __launch_bounds__(kColMaxThreads) void kb_true_residual_if_done(const BatchTank *T, const uint32_t *range,
                                                                           uint32_t nt)
{
    Blk G;
    const BatchTank &B = T[batch_tank(range, nt, blockIdx.x, G)];
    sor_residual_body(G, B.nsolve, B.binfo_f, B.fmask, B.div, B.diag, B.pres, B.partials, B.state, 2);
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
// the column-form passes: one wavefront per brick holding fluid
inline uint32_t col_blocks(const slosh_ctx *c, uint32_t n) { return (n + c->col_bricks - 1u) / c->col_bricks; }
#define L0C_LAUNCH(kid, kernel, ...)                                                                             \
    launch_timed(c, slosh::kid, kernel, dim3(col_blocks(c, c->nsolve)), dim3(64u * c->col_bricks), 0u, c->nsolve,    \
                 c->binfo_f, c->fmask, __VA_ARGS__)

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
    c->nparts = c->nactive;
    launch_timed(c, K_PCG_INIT, k_pcg_init, dim3(col_blocks(c, c->nactive)), dim3(64u * c->col_bricks), 0, c->nactive,
                 c->binfo, c->fmask, c->div, c->diag, c->pres, r_out, c->partials);
}

void l0_update_black(slosh_ctx *c, const float *r_in, float *r_out, const float *d, int first)
{
    if (first) L0C_LAUNCH(K_UPDATE_BLACK, k_update_black<true>, c->diag, c->pres, r_in, r_out, c->z, d, c->q, c->state);
    else L0C_LAUNCH(K_UPDATE_BLACK, k_update_black<false>, c->diag, c->pres, r_in, r_out, c->z, d, c->q, c->state);
}

void l0_restrict(slosh_ctx *c, const float *r)
{
    L0C_LAUNCH(K_RESTRICT0, k_restrict0, c->diag, c->z, r, plevel_of(c, 1), c->state);
}

void l0_prolong_black(slosh_ctx *c, const float *r)
{
    L0C_LAUNCH(K_PROLONG_BLACK, k_prolong_black, c->diag, r, c->z, c->z1, plevel_of(c, 1), c->state);
}

void l0_red_dot(slosh_ctx *c, const float *r)
{
    c->nparts = c->nsolve;
    L0C_LAUNCH(K_RED_DOT, k_red_dot, c->diag, r, c->z1, c->partials, c->state);
}

void l0_direction(slosh_ctx *c, const float *d_in, float *d_out, const float *r, int first)
{
    c->nparts = c->nsolve;
    if (first) L0C_LAUNCH(K_DIRECTION, k_direction<true>, c->diag, c->z1, d_in, d_out, c->q, r, c->partials, c->state);
    else L0C_LAUNCH(K_DIRECTION, k_direction<false>, c->diag, c->z1, d_in, d_out, c->q, r, c->partials, c->state);
}

void l0_fcg_zq(slosh_ctx *c, const float *z)
{
    c->nparts = c->nsolve;
    L0_LAUNCH(K_FCG, k_fcg_zq, z, c->q, c->partials, c->state);
}

void l0_fcg_direction(slosh_ctx *c, const float *z, const float *d_in, float *d_out, const float *r, int first)
{
    c->nparts = c->nsolve;
    L0_LAUNCH(K_FCG, k_fcg_direction, c->diag, z, d_in, d_out, c->q, r, first, c->partials, c->state);
}

void l0_fcg_update(slosh_ctx *c, float *r, const float *d)
{
    c->nparts = c->nsolve;
    L0_LAUNCH(K_FCG, k_fcg_update, c->pres, r, d, c->q, c->partials, c->state);
}

void l0_sor_pair(slosh_ctx *c, const double *pin, double *pout, double w)
{
    L0C_LAUNCH(K_SOR_PAIR, k_sor_pair, c->div, c->diag, pin, pout, w, c->state);
}

void l0_sor_residual(slosh_ctx *c)
{
    c->nparts = c->nsolve;
    L0C_LAUNCH(K_SOR_RESIDUAL, k_sor_residual, c->div, c->diag, c->pres, c->partials, c->state, 0);
}

void l0_fcg_resid(slosh_ctx *c, const float *r, const float *z, float *out)
{
    L0_LAUNCH(K_FCG, k_fcg_resid, c->diag, r, z, out, c->state);
}

void l0_error_partials(slosh_ctx *c, const double *p_ref)
{
    c->nparts = c->nsolve;
    L0_LAUNCH(K_SOR_RESIDUAL, k_error_partials, c->pres, p_ref, c->partials);
}

void l0_true_residual(slosh_ctx *c)
{
    c->nparts = c->nsolve;
    L0C_LAUNCH(K_SOR_RESIDUAL, k_sor_residual, c->div, c->diag, c->pres, c->partials, c->state, 1);
}

void l0_true_residual_if_done(slosh_ctx *c)
{
    c->nparts = c->nsolve;
    L0C_LAUNCH(K_SOR_RESIDUAL, k_sor_residual, c->div, c->diag, c->pres, c->partials, c->state, 2);
}

// The column passes over a batch: the workgroup count of the column range, 64 threads per brick.
#define L0B_LAUNCH(kernel, ...)                                                                                  \
    hipLaunchKernelGGL(kernel, dim3(B.total[kRangeCol]), dim3(64u * B.col_bricks), 0u, B.stream, B.d_tanks,          \
                       B.range(kRangeCol), B.ntanks, __VA_ARGS__)

void b_update_black(const BatchRun &B, int ri, int ro, int di, bool first)
{
    if (first) L0B_LAUNCH(kb_update_black<true>, ri, ro, di);
    else L0B_LAUNCH(kb_update_black<false>, ri, ro, di);
}

void b_restrict(const BatchRun &B, int ri) { L0B_LAUNCH(kb_restrict0, ri); }

void b_prolong_black(const BatchRun &B, int ri) { L0B_LAUNCH(kb_prolong_black, ri); }

void b_red_dot(const BatchRun &B, int ri) { L0B_LAUNCH(kb_red_dot, ri); }

void b_direction(const BatchRun &B, int din, int dout, int ri, bool first)
{
    if (first) L0B_LAUNCH(kb_direction<true>, din, dout, ri);
    else L0B_LAUNCH(kb_direction<false>, din, dout, ri);
}

void b_true_residual_if_done(const BatchRun &B)
{
    hipLaunchKernelGGL(kb_true_residual_if_done, dim3(B.total[kRangeCol]), dim3(64u * B.col_bricks), 0u, B.stream,
                       B.d_tanks, B.range(kRangeCol), B.ntanks);
}

} // namespace slosh
