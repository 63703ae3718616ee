// unet.hip -- the learned preconditioner: a 3-level 3-D U-Net on the brick pool, BF16 on MFMA, FP32 accumulate.
//
// No reference counterpart (the reference preconditions CG with the multigrid V-cycle only, SloshPressureMG.usf);
// the network is this library's own component. Definition (also the layer table below):
//   input, level 0, 4 channels (padded to 16): r / s, fluid, air, solid       s = RMS of r over fluid cells
//   E0a conv3 4->16 ReLU, E0b 16->16 ReLU = s0                                level 0 (8^3 per brick)
//   mean-pool; E1a 16->32 ReLU, E1b 32->32 ReLU = s1                          level 1 (4^3 per brick)
//   mean-pool; B2a 32->64 ReLU, B2b 64->64 ReLU, D1p 64->32                   level 2 (2^3 per brick)
//   t1 = ReLU(up(D1p) + s1); D1b 32->32 ReLU; D0p 32->16                      level 1
//   t0 = ReLU(up(D0p) + s0); D0b 16->16 ReLU; head 1x1 16->1 on fluid cells   level 0   -> z = s * head
// Convolutions are 3x3x3 correlations (tap t = (dx+1) + 3(dy+1) + 9(dz+1)) with zero padding outside allocated
// bricks. Activations are stored channel-last in BF16 per level in the pool layout (slot * cells_per_brick + local,
// the null slot all zero), weights as BF16 [tap][cout][cin]; the FP32 master weights live in the flat parameter
// vector (per layer W[27][cout][cin] then bias[cout]).
//
// A convolution is an implicit GEMM per workgroup: M = the output cells of its bricks (level 0: one brick, 512;
// level 1: four, 256; level 2: eight, 64), N = output channels, K = 27 taps x input channels. The input block of the workgroup's
// bricks plus a one-cell halo is staged in LDS from the brick's 3x3x3 neighbourhood; A fragments (16 cells x 16
// channels) come from LDS, B fragments (16 input x 16 output channels of one tap) from the L2-resident weights,
// and v_mfma_f32_16x16x16_bf16 accumulates in FP32 (8^3 bricks: v_mfma_f32_16x16x32_bf16 over pairs of taps, B
// fragments held in registers, A fragments reused across the three z taps). The backward pass w.r.t. the input is the same kernel with the
// weights mirrored in the tap and transposed in the channels; weight gradients are FP32 reductions.
#include <math.h>

#include <new>

#include "slosh_ctx.h"

namespace slosh {

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

inline uint32_t blocks_for(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

constexpr uint32_t kSelfFluid = 1u << 6;

// Bricks per workgroup of a pass at a level with S cells per brick axis: 512, 256 and 128 output cells, so that a
// workgroup's weight fragments feed at least 2 M-tiles per wavefront at the coarse levels.
__host__ __device__ constexpr int group_bricks(int S) { return S == 2 ? 16 : (S == 4 ? 4 : 1); }
// The forward/input-gradient convolution groups 8 bricks at 2^3 (64 output cells): 64 channels then stage 71 KB of
// LDS, two workgroups per CU instead of one (16 bricks: 141 KB).
__host__ __device__ constexpr int conv_group_bricks(int S) { return S == 2 ? 8 : group_bricks(S); }

// LDS row stride (in BF16 elements) of one staged cell: the channels plus 4 of padding, so that the 16 rows of an A
// fragment fall into distinct LDS banks.
__host__ __device__ constexpr int lds_stride(int C) { return C + 4; }

// Group-relative output cell of row r (0..15) of M tile t. In 8^3 bricks a tile holds x = 0..7 of rows y and y + 4
// of one z plane: the halves lie 40 staged cells apart, so that for every tap the rows' 8-byte A reads fall into
// distinct LDS bank pairs (with rows y and y + 1, 10 cells apart, cells c and c + 16 of a fragment share a pair;
// measured 8% slower).
// In 2^3 bricks (8 per workgroup) tile t holds plane z = t & 1 of bricks 2 (t >> 1) + {0, 1, 4, 5}, 2x2 cells each:
// with the staged boxes 4 dwords apart modulo 32 banks, those four bricks' rows fall into distinct bank pairs, and
// the tiles of one brick set come in z pairs for the reuse of A reads across the z taps. In 4^3 bricks tile t is
// the x-z plane y = t & 3 of brick t >> 2 (rows x = r & 3, z = r >> 2): its rows lie 0..3 and 36..39, ... staged
// cells apart, distinct bank pairs at 16 and 32 channels (x-y planes would pair rows 2 and 12, 18 cells apart).
template <int S> __device__ __forceinline__ uint32_t tile_vox(uint32_t t, uint32_t r)
{
    if constexpr (S == 8) return (r & 7u) + 8u * ((t & 3u) + 4u * (r >> 3)) + 64u * (t >> 2);
    if constexpr (S == 2) return (2u * (t >> 1) + ((r >> 2) & 1u) + 4u * (r >> 3)) * 8u + (r & 3u) + 4u * (t & 1u);
    return (t >> 2) * 64u + (r & 3u) + 4u * (t & 3u) + 16u * (r >> 2);
}

// M tile of the i-th fragment of M group mw (of MTG tiles). In 8^3 bricks group w takes rows y = w and w + 4 of
// every z plane (tile 4 i + w holds plane i), so that one A fragment read at box plane b serves the three output
// planes b - 2, b - 1 and b of the taps dz = 2, 1, 0: 10 reads per (dx, dy) instead of 24.
template <int S, int MTG> __device__ __forceinline__ uint32_t wave_tile(uint32_t mw, int i)
{
    if constexpr (S == 8) return (uint32_t)i * 4u + mw;
    return mw * MTG + (uint32_t)i;
}

// Pool index (cell) of block-relative coordinates (u, v, w) in [-1, S + 1)^3 through the 27-neighbourhood slots.
template <int S> __device__ __forceinline__ uint32_t nb_cell(const uint32_t *nb27, int u, int v, int w)
{
    const int ox = u < 0 ? -1 : (u >= S ? 1 : 0), oy = v < 0 ? -1 : (v >= S ? 1 : 0), oz = w < 0 ? -1 : (w >= S ? 1 : 0);
    const uint32_t slot = nb27[(ox + 1) + 3 * (oy + 1) + 9 * (oz + 1)];
    return slot * (uint32_t)(S * S * S) + (uint32_t)((u - ox * S) + S * ((v - oy * S) + S * (w - oz * S)));
}

// ---- convolution (forward, and backward w.r.t. the input with mirrored/transposed weights) -----------------
// 8^3 bricks: at most 128 VGPRs, so that the 4 workgroups per CU that 40 KB of LDS allows are resident together;
// 2^3 bricks: at most 256, for the 2 that 71 KB allow (64 input channels then spill about 20 registers: measured
// 48.7 us per launch against 52.1 for one workgroup of 16 bricks per CU). Level 1 is left uncapped (16 -> 32
// channels would spill under a cap of 128).
template <int CIN, int COUT, int S, bool RELU, bool BIAS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(S == 8 ? 4 : (S == 2 ? 2 : 1)))) void k_conv(
    const uint32_t *binfo, uint32_t nbricks, const uint16_t *in, const uint16_t *w, const float *bias, uint16_t *out,
    const int32_t *skip)
{
    if (*skip) return;
    constexpr int BG = conv_group_bricks(S), SB = S + 2, BOX = SB * SB * SB, CPB = S * S * S, M = BG * CPB;
    constexpr int MT = M / 16, NT = COUT / 16, KB = CIN / 16, CH8 = CIN / 8, LS = lds_stride(CIN);
    // LDS elements per staged box: 2^3 bricks pad 8 (4 dwords), so that consecutive boxes start 4 banks apart
    constexpr int BSTR = BOX * LS + (S == 2 ? 8 : 0);
    static_assert(MT % 4 == 0 && COUT % 16 == 0 && CIN % 16 == 0, "conv tiling");
    static_assert(S != 2 || (BG == 8 && (BOX * LS / 2) % 32 == 0), "2^3 tiles: 8 bricks, boxes 4 banks apart");
    __shared__ __attribute__((aligned(16))) uint16_t s_in[BG * BSTR];
    __shared__ uint32_t s_nb[BG][28];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t grp = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t row = lane & 15u, quad = lane >> 4;
    // Wavefront w takes N tile w % NT and M group w / NT (the M tiles split into 4 / NT groups), so that its B
    // fragments of every tap stay in registers, loaded once before the box: their latency overlaps the staging instead of stalling each
    // tap. 8^3 bricks (CIN = COUT = 16): K = 32 blocks pair the taps (dx, dy) = 2p and 2p + 1 of one dz (quads 0-1
    // the first, 2-3 the second; the second of p = 4 is zero), 15 fragments in 60 VGPRs. Other levels: one K = 16
    // fragment per tap and 16 input channels (at most 108 in 216 VGPRs).
    constexpr bool kS8 = S == 8;
    constexpr int WM = 4 / NT, MTG = MT / WM;
    static_assert(NT == 1 || NT == 2 || NT == 4, "one N tile per wavefront");
    static_assert(!kS8 || (KB == 1 && NT == 1 && MTG == 8 && BG == 1), "8^3 bricks: 16 channels in and out");
    const uint32_t nw = wave % (uint32_t)NT, mw = wave / (uint32_t)NT;
    v8bf wpair[kS8 ? 15 : 1];
    v4s wtap[kS8 ? 1 : 27 * KB];
    if constexpr (kS8) {
        const uint32_t hi = quad >> 1, c0 = (quad & 1u) * 8u;
#pragma unroll
        for (int p = 0; p < 5; ++p)
#pragma unroll
            for (int dz = 0; dz < 3; ++dz) {
                const uint32_t dxy = 2u * (uint32_t)p + hi;
                uint4 v = make_uint4(0u, 0u, 0u, 0u);
                if (dxy < 9u) v = *(const uint4 *)(w + ((size_t)((dxy + 9u * (uint32_t)dz) * COUT + row) * CIN + c0));
                wpair[p * 3 + dz] = __builtin_bit_cast(v8bf, v);
            }
    } else {
#pragma unroll
        for (int tap = 0; tap < 27; ++tap)
#pragma unroll
            for (int kb = 0; kb < KB; ++kb)
                wtap[tap * KB + kb] =
                    *(const v4s *)(w + ((size_t)(tap * COUT + nw * 16u + row) * CIN + kb * 16 + quad * 4u));
    }
    for (uint32_t k = tid; k < (uint32_t)(BG * 28); k += 256u) {
        const uint32_t bg = k / 28u, e = grp * BG + bg;
        s_nb[bg][k % 28u] = e < nbricks ? binfo[(size_t)e * kBinfoWords + k % 28u] : 0u;
    }
    __syncthreads();
    // the box in rounds of up to 8 16-byte chunks per thread: every chunk of a round is loaded before any is stored
    constexpr uint32_t kChunks = (uint32_t)(BG * BOX * CH8), kRound = 8u * 256u;
    for (uint32_t k0 = 0; k0 < kChunks; k0 += kRound) {
        uint4 sv[8];
#pragma unroll
        for (uint32_t q = 0; q < 8u; ++q) {
            const uint32_t k = k0 + tid + 256u * q;
            sv[q] = make_uint4(0u, 0u, 0u, 0u);
            if (k >= kChunks) continue;
            const uint32_t c8 = k % CH8, qq = k / CH8, bg = qq / BOX, bi = qq % BOX;
            const int u = (int)(bi % SB) - 1, v = (int)((bi / SB) % SB) - 1, ww = (int)(bi / (SB * SB)) - 1;
            const uint32_t cell = nb_cell<S>(s_nb[bg], u, v, ww);
            sv[q] = *(const uint4 *)(in + (size_t)cell * CIN + c8 * 8);
        }
#pragma unroll
        for (uint32_t q = 0; q < 8u; ++q) {
            const uint32_t k = k0 + tid + 256u * q;
            if (k >= kChunks) continue;
            const uint32_t c8 = k % CH8, qq = k / CH8, bg = qq / BOX, bi = qq % BOX;
            uint2 *dst = (uint2 *)(s_in + ((size_t)bg * BSTR + bi * LS + c8 * 8));
            dst[0] = make_uint2(sv[q].x, sv[q].y);
            dst[1] = make_uint2(sv[q].z, sv[q].w);
        }
    }
    __syncthreads();
    v4f acc[MTG];
    for (int i = 0; i < MTG; ++i) acc[i] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    if constexpr (kS8) {
        // A rows: x = row & 7, y = mw + 4 (row >> 3); fragment i is output plane z = i, and the box plane b (cells
        // at z = b - 1) feeds plane b - dz through tap dz. Lanes of the zero tap read the last tap's cells.
        const uint32_t hi = quad >> 1;
        const uint32_t rowbase = ((row & 7u) + SB * (mw + 4u * (row >> 3))) * LS + (quad & 1u) * 8u;
#pragma unroll
        for (int p = 0; p < 5; ++p) {
            const uint32_t dxy = 2u * (uint32_t)p + hi < 8u ? 2u * (uint32_t)p + hi : 8u;
            const uint32_t poff = rowbase + ((dxy % 3u) + SB * (dxy / 3u)) * LS;
#pragma unroll
            for (int b = 0; b < SB; ++b) {
                const uint2 lo = *(const uint2 *)(s_in + poff + (uint32_t)(SB * SB * LS * b));
                const uint2 up = *(const uint2 *)(s_in + poff + (uint32_t)(SB * SB * LS * b) + 4u);
                const v8bf a = __builtin_bit_cast(v8bf, make_uint4(lo.x, lo.y, up.x, up.y));
#pragma unroll
                for (int dz = 0; dz < 3; ++dz)
                    if (b - dz >= 0 && b - dz < MTG)
                        acc[b - dz] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, wpair[p * 3 + dz], acc[b - dz], 0, 0, 0);
            }
        }
    } else {
        // 4^3 and 2^3 bricks: a wavefront's M tiles come in sets of S planes of the same bricks (tile S j + k: plane
        // k across axis Q, y at 4^3 and z at 2^3), so one A read at box plane b serves planes b, b - 1 and b - 2
        // through the taps 0, 1, 2 along Q (S + 2 reads per two other tap offsets and K block instead of 3 S)
        static_assert((S == 4 || S == 2) && MTG % S == 0, "whole plane sets per wavefront");
        constexpr int Q = S == 4 ? 1 : 2;                          // the reuse axis
        constexpr int QSTR = Q == 1 ? SB : SB * SB;                // its box stride
#pragma unroll
        for (int j = 0; j < MTG / S; ++j) {
            const uint32_t vox = tile_vox<S>(wave_tile<S, MTG>(mw, S * j), row), bg = vox / CPB, l = vox % CPB;
            // the lane's cell with the reuse coordinate at 0
            const uint32_t cx = l % S, cy = Q == 1 ? 0u : (l / S) % S, cz = Q == 1 ? l / (S * S) : 0u;
            const uint32_t rowbase = bg * BSTR + (cx + SB * (cy + SB * cz)) * LS + quad * 4u;
#pragma unroll
            for (int o = 0; o < 9; ++o) {
                // the two other tap offsets: (dx, dz) at 4^3, (dx, dy) at 2^3
                const int dx = o % 3, d2 = o / 3;
                const int ooff = Q == 1 ? dx + SB * SB * d2 : dx + SB * d2;
#pragma unroll
                for (int kb = 0; kb < KB; ++kb)
#pragma unroll
                    for (int b = 0; b < SB; ++b) {
                        const v4s a = *(const v4s *)(s_in + rowbase + (uint32_t)((ooff + QSTR * b) * LS + kb * 16));
#pragma unroll
                        for (int dq = 0; dq < 3; ++dq)
                            if (b - dq >= 0 && b - dq < S) {
                                const int tap = Q == 1 ? dx + 3 * dq + 9 * d2 : dx + 3 * d2 + 9 * dq;
                                acc[S * j + b - dq] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(
                                    a, wtap[tap * KB + kb], acc[S * j + b - dq], 0, 0, 0);
                            }
                    }
            }
        }
    }
    // the output tile goes through LDS (the staged input is no longer read), then to memory in 16-byte chunks
    static_assert(M * COUT <= BG * BSTR, "output tile fits the input's LDS");
    __syncthreads();
    {
        const uint32_t co = nw * 16u + row;
        const float bv = BIAS ? bias[co] : 0.0f;
        for (int i = 0; i < MTG; ++i)
            for (int j = 0; j < 4; ++j) {
                const uint32_t vox = tile_vox<S>(wave_tile<S, MTG>(mw, i), quad * 4u + (uint32_t)j);
                float val = acc[i][j] + bv;
                if (RELU) val = fmaxf(val, 0.0f);
                s_in[vox * COUT + co] = f2bf(val);
            }
    }
    __syncthreads();
    constexpr int CO8 = COUT / 8;
    for (uint32_t k = tid; k < (uint32_t)(M * CO8); k += 256u) {
        const uint32_t vox = k / CO8, c8 = k % CO8, bg = vox / CPB, l = vox % CPB;
        if (grp * BG + bg >= nbricks) continue;
        *(uint4 *)(out + ((size_t)s_nb[bg][13] * CPB + l) * COUT + c8 * 8) = *(const uint4 *)(s_in + vox * COUT + c8 * 8);
    }
}

// ---- elementwise passes ---------------------------------------------------------------------------------------
// Thread per (cell, 8-channel chunk) of the allocated cells of a level (cells_per_brick CPB, C channels).
template <int C> __device__ __forceinline__ bool cell_chunk(const uint32_t *binfo, uint32_t nbricks, uint32_t cpb,
                                                            uint32_t i, uint32_t &slot, uint32_t &l, uint32_t &c8)
{
    constexpr uint32_t CH8 = C / 8;
    c8 = i % CH8;
    const uint32_t q = i / CH8, e = q / cpb;
    l = q % cpb;
    if (e >= nbricks) return false;
    slot = binfo[(size_t)e * kBinfoWords + 13];
    return true;
}

union bf8 {
    uint4 v;
    uint16_t h[8];
};
__device__ __forceinline__ bf8 ld8(const uint16_t *p)
{
    bf8 r;
    r.v = *(const uint4 *)p;
    return r;
}
__device__ __forceinline__ void st8(uint16_t *p, const bf8 &v) { *(uint4 *)p = v.v; }

// Level l (S cells per axis) -> level l + 1: mean over the 2x2x2 children (same brick), child order x fastest.
template <int C, int S>
__global__ void k_pool(const uint32_t *binfo, uint32_t nbricks, const uint16_t *in, uint16_t *out, const int32_t *skip)
{
    if (*skip) return;
    constexpr int SC = S / 2, CPBC = SC * SC * SC, CPB = S * S * S;
    uint32_t slot, l, c8;
    if (!cell_chunk<C>(binfo, nbricks, CPBC, blockIdx.x * blockDim.x + threadIdx.x, slot, l, c8)) return;
    const uint32_t cx = l % SC, cy = (l / SC) % SC, cz = l / (SC * SC);
    float sum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t oi = 0; oi < 8u; ++oi) {
        const uint32_t f = (2 * cx + (oi & 1u)) + S * ((2 * cy + ((oi >> 1) & 1u)) + S * (2 * cz + (oi >> 2)));
        const bf8 v = ld8(in + ((size_t)slot * CPB + f) * C + c8 * 8);
        for (int k = 0; k < 8; ++k) sum[k] += bf2f(v.h[k]);
    }
    bf8 o;
    for (int k = 0; k < 8; ++k) o.h[k] = f2bf(sum[k] / 8.0f);
    st8(out + ((size_t)slot * CPBC + l) * C + c8 * 8, o);
}

__device__ __forceinline__ uint32_t parent_local(uint32_t l, int S)
{
    const int SC = S / 2;
    const uint32_t x = l % S, y = (l / S) % S, z = l / (S * S);
    return (x >> 1) + SC * ((y >> 1) + SC * (z >> 1));
}

// t = ReLU(up(c) + s) at level l (S per axis); c lives at level l + 1.
template <int C, int S>
__global__ void k_upadd(const uint32_t *binfo, uint32_t nbricks, const uint16_t *c, const uint16_t *s, uint16_t *t,
                        const int32_t *skip)
{
    if (*skip) return;
    constexpr int CPB = S * S * S, CPBC = CPB / 8;
    uint32_t slot, l, c8;
    if (!cell_chunk<C>(binfo, nbricks, CPB, blockIdx.x * blockDim.x + threadIdx.x, slot, l, c8)) return;
    const bf8 a = ld8(c + ((size_t)slot * CPBC + parent_local(l, S)) * C + c8 * 8);
    const bf8 b = ld8(s + ((size_t)slot * CPB + l) * C + c8 * 8);
    bf8 o;
    for (int k = 0; k < 8; ++k) o.h[k] = f2bf(fmaxf(bf2f(a.h[k]) + bf2f(b.h[k]), 0.0f));
    st8(t + ((size_t)slot * CPB + l) * C + c8 * 8, o);
}

// Scale of the input: per-brick partials of sum r^2 and fluid count over fluid cells.
__global__ __launch_bounds__(256) void k_unet_norm(const uint32_t *binfo, const uint8_t *fmask, const float *r,
                                                   double *partials, const int32_t *skip)
{
    if (*skip) return;
    const uint32_t base = binfo[(size_t)xcd_remap(blockIdx.x, gridDim.x) * kBinfoWords + 13] * kBrickCells;
    double v0 = 0.0, v1 = 0.0;
    for (uint32_t l = threadIdx.x; l < kBrickCells; l += 256u) {
        const uint32_t i = base + l;
        if (!(fmask[i] & kSelfFluid)) continue;
        v0 += (double)r[i] * (double)r[i];
        v1 += 1.0;
    }
    __shared__ double sh[2][256];
    sh[0][threadIdx.x] = v0;
    sh[1][threadIdx.x] = v1;
    __syncthreads();
    for (uint32_t s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = sh[0][0];
        partials[2 * blockIdx.x + 1] = sh[1][0];
    }
}

// Sums n partial pairs in a fixed order; mode 0: scale (s, 1/s) of the input; mode 1: the loss terms
// (sum e^2, sum rhat^2).
__global__ __launch_bounds__(256) void k_unet_finalize(const double *partials, uint32_t n, int mode, UNetScalars *sc,
                                                       const int32_t *skip)
{
    if (*skip) return;
    __shared__ double sh[2][256];
    double a = 0.0, b = 0.0;
    for (uint32_t k = threadIdx.x; k < n; k += 256u) {
        a += partials[2 * k];
        b += partials[2 * k + 1];
    }
    sh[0][threadIdx.x] = a;
    sh[1][threadIdx.x] = b;
    __syncthreads();
    for (uint32_t s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    if (mode == 0) {
        const double rr = sh[0][0], nf = sh[1][0] > 0.0 ? sh[1][0] : 1.0;
        sc->s = rr > 0.0 ? (float)sqrt(rr / nf) : 0.0f;
        sc->s_inv = rr > 0.0 ? (float)(1.0 / sqrt(rr / nf)) : 0.0f;
    } else {
        sc->err2 = sh[0][0];
        sc->rhat2 = sh[1][0];
        sc->loss_sum += sh[1][0] > 0.0 ? sh[0][0] / sh[1][0] : 0.0;
        sc->loss_count += 1.0;
    }
}

// Network input at level 0: channels (r / s on fluid cells, fluid, air, solid, 0 x 12).
__global__ __launch_bounds__(256) void k_unet_input(const uint32_t *binfo, const uint32_t *ctype, const float *r,
                                                    const UNetScalars *sc, uint16_t *x0, const int32_t *skip)
{
    if (*skip) return;
    const uint32_t base = binfo[(size_t)xcd_remap(blockIdx.x, gridDim.x) * kBinfoWords + 13] * kBrickCells;
    const float s_inv = sc->s_inv;
    for (uint32_t l = threadIdx.x; l < kBrickCells; l += 256u) {
        const uint32_t i = base + l, t = ctype[i] & kCellTypeMask;
        bf8 lo, hi;
        for (int k = 0; k < 8; ++k) {
            lo.h[k] = 0;
            hi.h[k] = 0;
        }
        const uint16_t one = 0x3F80u;
        if (t == kCellFluid) {
            lo.h[0] = f2bf(r[i] * s_inv);
            lo.h[1] = one;
        } else if (t == kCellAir) {
            lo.h[2] = one;
        } else {
            lo.h[3] = one;
        }
        st8(x0 + (size_t)i * 16, lo);
        st8(x0 + (size_t)i * 16 + 8, hi);
    }
}

// Head: zhat = sum_c W[c] d0[c] + b on fluid cells (0 elsewhere); z = s * zhat, plus zadd when given (the
// V-cycle's part of the hybrid preconditioner).
__global__ __launch_bounds__(256) void k_unet_head(const uint32_t *binfo, const uint8_t *fmask, const uint16_t *d0,
                                                   const uint16_t *wh, const float *bh, const UNetScalars *sc,
                                                   float *zhat, float *z, const float *zadd, const int32_t *skip)
{
    if (*skip) return;
    const uint32_t base = binfo[(size_t)xcd_remap(blockIdx.x, gridDim.x) * kBinfoWords + 13] * kBrickCells;
    float w[16];
    for (int k = 0; k < 16; ++k) w[k] = bf2f(wh[k]);
    const float b = bh[0], s = sc->s;
    for (uint32_t l = threadIdx.x; l < kBrickCells; l += 256u) {
        const uint32_t i = base + l;
        float zh = 0.0f;
        if (fmask[i] & kSelfFluid) {
            const bf8 a = ld8(d0 + (size_t)i * 16), c = ld8(d0 + (size_t)i * 16 + 8);
            float acc = 0.0f;
            for (int k = 0; k < 8; ++k) acc += w[k] * bf2f(a.h[k]);
            for (int k = 0; k < 8; ++k) acc += w[8 + k] * bf2f(c.h[k]);
            zh = acc + b;
        }
        if (zhat) zhat[i] = zh;
        if (z) z[i] = zadd ? zadd[i] + s * zh : s * zh;
    }
}

// ---- training ------------------------------------------------------------------------------------------------
// e = A zhat - rhat on fluid cells (rhat = the input's BF16 channel 0), partials sum e^2, sum rhat^2.
__global__ __launch_bounds__(256) void k_unet_loss_e(const uint32_t *binfo, const uint8_t *fmask, const float *diag,
                                                     const float *zhat, const uint16_t *x0, float *e, double *partials)
{
    const uint32_t *nb = binfo + (size_t)xcd_remap(blockIdx.x, gridDim.x) * kBinfoWords;
    const uint32_t base = nb[13] * kBrickCells;
    double v0 = 0.0, v1 = 0.0;
    for (uint32_t l = threadIdx.x; l < kBrickCells; l += 256u) {
        const uint32_t i = base + l, m = fmask[i];
        if (!(m & kSelfFluid)) {
            e[i] = 0.0f;
            continue;
        }
        const int x = (int)(l & 7u), y = (int)((l >> 3) & 7u), zz = (int)(l >> 6);
        float s = 0.0f;
        for (int n = 0; n < 6; ++n) {
            if (!(m >> n & 1u)) continue;
            int d[3];
            nb_offset(n, d);
            s += zhat[nb_cell<8>(nb, x + d[0], y + d[1], zz + d[2])];
        }
        const float rh = bf2f(x0[(size_t)i * 16]);
        const float ev = ((diag[i] + kDiagEps) * zhat[i] - s) - rh;
        e[i] = ev;
        v0 += (double)ev * (double)ev;
        v1 += (double)rh * (double)rh;
    }
    __shared__ double sh[2][256];
    sh[0][threadIdx.x] = v0;
    sh[1][threadIdx.x] = v1;
    __syncthreads();
    for (uint32_t s2 = 128; s2 > 0; s2 >>= 1) {
        if (threadIdx.x < s2) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + s2];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + s2];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = sh[0][0];
        partials[2 * blockIdx.x + 1] = sh[1][0];
    }
}

// Head backward: g = 2 A e / sum rhat^2 (dL/dzhat on fluid cells); g_d0 = W g masked by d0 > 0 (BF16); the head's
// weight and bias gradients accumulate into grad.
__global__ __launch_bounds__(256) void k_unet_head_bwd(const uint32_t *binfo, const uint8_t *fmask, const float *diag,
                                                       const float *e, const uint16_t *d0, const uint16_t *wh,
                                                       const UNetScalars *sc, uint16_t *g_d0, float *gw_head)
{
    const uint32_t *nb = binfo + (size_t)xcd_remap(blockIdx.x, gridDim.x) * kBinfoWords;
    const uint32_t base = nb[13] * kBrickCells;
    const float scale = sc->rhat2 > 0.0 ? (float)(2.0 / sc->rhat2) : 0.0f;
    float w[16], gw[17];
    for (int k = 0; k < 16; ++k) w[k] = bf2f(wh[k]);
    for (int k = 0; k < 17; ++k) gw[k] = 0.0f;
    for (uint32_t l = threadIdx.x; l < kBrickCells; l += 256u) {
        const uint32_t i = base + l, m = fmask[i];
        bf8 lo, hi;
        for (int k = 0; k < 8; ++k) {
            lo.h[k] = 0;
            hi.h[k] = 0;
        }
        if (m & kSelfFluid) {
            const int x = (int)(l & 7u), y = (int)((l >> 3) & 7u), zz = (int)(l >> 6);
            float s = 0.0f;
            for (int n = 0; n < 6; ++n) {
                if (!(m >> n & 1u)) continue;
                int d[3];
                nb_offset(n, d);
                s += e[nb_cell<8>(nb, x + d[0], y + d[1], zz + d[2])];
            }
            const float g = scale * ((diag[i] + kDiagEps) * e[i] - s);
            const bf8 a = ld8(d0 + (size_t)i * 16), c = ld8(d0 + (size_t)i * 16 + 8);
            for (int k = 0; k < 8; ++k) {
                const float va = bf2f(a.h[k]), vc = bf2f(c.h[k]);
                gw[k] += va * g;
                gw[8 + k] += vc * g;
                lo.h[k] = va > 0.0f ? f2bf(w[k] * g) : (uint16_t)0;
                hi.h[k] = vc > 0.0f ? f2bf(w[8 + k] * g) : (uint16_t)0;
            }
            gw[16] += g;
        }
        st8(g_d0 + (size_t)i * 16, lo);
        st8(g_d0 + (size_t)i * 16 + 8, hi);
    }
    for (int k = 0; k < 17; ++k) {
        float v = gw[k];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if ((threadIdx.x & 63u) == 0u) atomicAdd(&gw_head[k], v);
    }
}

// g *= (act > 0) at one level.
template <int C> __global__ void k_relu_mask(const uint32_t *binfo, uint32_t nbricks, uint32_t cpb, const uint16_t *act,
                                             uint16_t *g)
{
    uint32_t slot, l, c8;
    if (!cell_chunk<C>(binfo, nbricks, cpb, blockIdx.x * blockDim.x + threadIdx.x, slot, l, c8)) return;
    const size_t o = ((size_t)slot * cpb + l) * C + c8 * 8;
    const bf8 a = ld8(act + o);
    bf8 v = ld8(g + o);
    for (int k = 0; k < 8; ++k)
        if (!(bf2f(a.h[k]) > 0.0f)) v.h[k] = 0;
    st8(g + o, v);
}

// Backward of t = ReLU(up(c) + s) at level l: gu = g_t (t > 0) (in place), g_c = sum of gu over the 8 children.
template <int C, int S>
__global__ void k_upadd_bwd(const uint32_t *binfo, uint32_t nbricks, const uint16_t *t, uint16_t *g_t, uint16_t *g_c)
{
    constexpr int SC = S / 2, CPBC = SC * SC * SC, CPB = S * S * S;
    uint32_t slot, l, c8;
    if (!cell_chunk<C>(binfo, nbricks, CPBC, blockIdx.x * blockDim.x + threadIdx.x, slot, l, c8)) return;
    const uint32_t cx = l % SC, cy = (l / SC) % SC, cz = l / (SC * SC);
    float sum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t oi = 0; oi < 8u; ++oi) {
        const uint32_t f = (2 * cx + (oi & 1u)) + S * ((2 * cy + ((oi >> 1) & 1u)) + S * (2 * cz + (oi >> 2)));
        const size_t o = ((size_t)slot * CPB + f) * C + c8 * 8;
        const bf8 tv = ld8(t + o);
        bf8 gv = ld8(g_t + o);
        for (int k = 0; k < 8; ++k) {
            if (!(bf2f(tv.h[k]) > 0.0f)) gv.h[k] = 0;
            sum[k] += bf2f(gv.h[k]);
        }
        st8(g_t + o, gv);
    }
    bf8 o;
    for (int k = 0; k < 8; ++k) o.h[k] = f2bf(sum[k]);
    st8(g_c + ((size_t)slot * CPBC + l) * C + c8 * 8, o);
}

// Gradient at a skip tensor s (level l): g_s = (gu + g_p(parent) / 8) (s > 0), g_p the gradient of pool(s).
template <int C, int S>
__global__ void k_skip_grad(const uint32_t *binfo, uint32_t nbricks, const uint16_t *gu, const uint16_t *g_p,
                            const uint16_t *s, uint16_t *g_s)
{
    constexpr int CPB = S * S * S, CPBC = CPB / 8;
    uint32_t slot, l, c8;
    if (!cell_chunk<C>(binfo, nbricks, CPB, blockIdx.x * blockDim.x + threadIdx.x, slot, l, c8)) return;
    const size_t o = ((size_t)slot * CPB + l) * C + c8 * 8;
    const bf8 a = ld8(gu + o), p = ld8(g_p + ((size_t)slot * CPBC + parent_local(l, S)) * C + c8 * 8), sv = ld8(s + o);
    bf8 r;
    for (int k = 0; k < 8; ++k)
        r.h[k] = bf2f(sv.h[k]) > 0.0f ? f2bf(bf2f(a.h[k]) + bf2f(p.h[k]) / 8.0f) : (uint16_t)0;
    st8(g_s + o, r);
}

// Weight and bias gradients of a convolution: dW[t][co][ci] += sum_v in[v + t][ci] g[v][co], db[co] += sum_v g[v][co]
// (FP32; one workgroup per brick group, the input box and g staged in LDS, one atomic add per output). The input
// is stored with CIN channels, of which the weights use the first CINW.
template <int CIN, int COUT, int S, int CINW>
__global__ __launch_bounds__(256) void k_conv_dw(const uint32_t *binfo, uint32_t nbricks, const uint16_t *in,
                                                 const uint16_t *g, float *dw, float *db)
{
    constexpr int BG = group_bricks(S), SB = S + 2, BOX = SB * SB * SB, CPB = S * S * S, M = BG * CPB, CH8 = CIN / 8;
    __shared__ __attribute__((aligned(16))) uint16_t s_in[BG * BOX * CIN];
    __shared__ __attribute__((aligned(16))) uint16_t s_g[M * COUT];
    __shared__ uint32_t s_nb[BG][28];
    const uint32_t tid = threadIdx.x;
    const uint32_t grp = xcd_remap(blockIdx.x, gridDim.x);
    for (uint32_t k = tid; k < (uint32_t)(BG * 28); k += 256u) {
        const uint32_t bg = k / 28u, e = grp * BG + bg;
        s_nb[bg][k % 28u] = e < nbricks ? binfo[(size_t)e * kBinfoWords + k % 28u] : 0u;
    }
    __syncthreads();
    for (uint32_t k = tid; k < (uint32_t)(BG * BOX * CH8); k += 256u) {
        const uint32_t c8 = k % CH8, q = k / CH8, bg = q / BOX, bi = q % BOX;
        const int u = (int)(bi % SB) - 1, v = (int)((bi / SB) % SB) - 1, ww = (int)(bi / (SB * SB)) - 1;
        const uint32_t cell = nb_cell<S>(s_nb[bg], u, v, ww);
        *(uint4 *)(s_in + ((size_t)(bg * BOX + bi) * CIN + c8 * 8)) = *(const uint4 *)(in + (size_t)cell * CIN + c8 * 8);
    }
    for (uint32_t k = tid; k < (uint32_t)(M * COUT / 8); k += 256u) {
        const uint32_t c8 = k % (COUT / 8), vox = k / (COUT / 8), bg = vox / CPB, l = vox % CPB;
        *(uint4 *)(s_g + (size_t)vox * COUT + c8 * 8) =
            *(const uint4 *)(g + ((size_t)s_nb[bg][13] * CPB + l) * COUT + c8 * 8);
    }
    __syncthreads();
    const uint32_t nb_valid = nbricks - grp * BG < (uint32_t)BG ? nbricks - grp * BG : (uint32_t)BG;
    for (uint32_t o = tid; o < (uint32_t)(27 * COUT * CINW); o += 256u) {
        const uint32_t ci = o % CINW, co = (o / CINW) % COUT, tap = o / (CINW * COUT);
        const uint32_t toff = (tap % 3u) + SB * (((tap / 3u) % 3u) + SB * (tap / 9u));
        float acc = 0.0f;
        for (uint32_t bg = 0; bg < nb_valid; ++bg)
            for (uint32_t l = 0; l < (uint32_t)CPB; ++l) {
                const uint32_t lx = l % S, ly = (l / S) % S, lz = l / (S * S);
                const uint32_t bi = bg * BOX + lx + SB * (ly + SB * lz) + toff;
                acc += bf2f(s_in[(size_t)bi * CIN + ci]) * bf2f(s_g[(size_t)(bg * CPB + l) * COUT + co]);
            }
        atomicAdd(&dw[o], acc);
    }
    for (uint32_t co = tid; co < (uint32_t)COUT; co += 256u) {
        float acc = 0.0f;
        for (uint32_t v = 0; v < nb_valid * CPB; ++v) acc += bf2f(s_g[(size_t)v * COUT + co]);
        atomicAdd(&db[co], acc);
    }
}

// Adam on the FP32 master weights, then the BF16 copies (forward layout and, for the backward pass, mirrored taps
// with transposed channels). t = step number (from 1).
__global__ void k_adam(float *w, float *grad, float *m, float *v, uint32_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2, float scale)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = grad[i] * scale;
    const float mi = b1 * m[i] + (1.0f - b1) * g, vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    grad[i] = 0.0f;
}

// One layer's BF16 weights from the FP32 master: wf[t][co][cinp] (input channels padded to cinp with zeros) and
// wt[26 - t][ci][co] (the transposed, mirrored kernel of the input gradient; only for cin == cinp).
__global__ void k_weights_bf16(const float *w32, uint32_t taps, uint32_t cin, uint32_t cinp, uint32_t cout,
                               uint16_t *wf, uint16_t *wt)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= taps * cout * cinp) return;
    const uint32_t ci = i % cinp, co = (i / cinp) % cout, t = i / (cinp * cout);
    const float v = ci < cin ? w32[(t * cout + co) * cin + ci] : 0.0f;
    wf[i] = f2bf(v);
    if (wt && ci < cin) wt[((taps - 1u - t) * cin + ci) * cout + co] = f2bf(v);
}

} // namespace

// ---- host side -------------------------------------------------------------------------------------------------
// Layer table: (cin, cout, level, taps). Parameter offsets follow from it.
const UNetLayerDesc kUNetLayers[kUNetLayerCount] = {
    {4, 16, 0, 27},  {16, 16, 0, 27}, {16, 32, 1, 27}, {32, 32, 1, 27}, {32, 64, 2, 27}, {64, 64, 2, 27},
    {64, 32, 2, 27}, {32, 32, 1, 27}, {32, 16, 1, 27}, {16, 16, 0, 27}, {16, 1, 0, 1}};

size_t unet_param_count()
{
    size_t n = 0;
    for (const auto &L : kUNetLayers) n += (size_t)L.taps * L.cout * L.cin + L.cout;
    return n;
}

namespace {

template <typename T> void ualloc(T **p, size_t count)
{
    if (hipMalloc((void **)p, (count ? count : 1) * sizeof(T)) != hipSuccess) throw std::bad_alloc();
    SLOSH_HIP(hipMemset(*p, 0, (count ? count : 1) * sizeof(T)));
}

size_t level_cells(const slosh_ctx *c, int level) { return (size_t)c->geo.npool * (kBrickCells >> (3 * level)); }

template <int CIN, int COUT, int S, bool RELU, bool BIAS>
void conv(slosh_ctx *c, const uint16_t *in, const uint16_t *w, const float *bias, uint16_t *out)
{
    constexpr int BG = conv_group_bricks(S);
    const uint32_t groups = (c->nactive + BG - 1) / BG;
    KTimer timer(c, S == 8 ? K_UNET_CONV0 : (S == 4 ? K_UNET_CONV1 : K_UNET_CONV2));
    hipLaunchKernelGGL((k_conv<CIN, COUT, S, RELU, BIAS>), dim3(groups), dim3(256), 0, c->stream, c->binfo, c->nactive,
                       in, w, bias, out, c->unet.skip);
}

template <int CIN, int COUT, int S, int CINW = CIN>
void conv_dw(slosh_ctx *c, const uint16_t *in, const uint16_t *g, int layer)
{
    constexpr int BG = group_bricks(S);
    const uint32_t groups = (c->nactive + BG - 1) / BG;
    UNet &u = c->unet;
    launch_timed(c, K_UNET_TRAIN, (k_conv_dw<CIN, COUT, S, CINW>), dim3(groups), dim3(256), 0, c->binfo, c->nactive, in, g,
                 u.grad + u.w_off[layer], u.grad + u.b_off[layer]);
}

uint32_t chunks(const slosh_ctx *c, int level, int C) { return blocks_for((uint64_t)c->nactive * (kBrickCells >> (3 * level)) * (C / 8), 256); }

} // namespace

void unet_init(slosh_ctx *c)
{
    UNet &u = c->unet;
    if (u.w32) return;
    u.nparams = unet_param_count();
    size_t o = 0, ob = 0;
    for (int k = 0; k < kUNetLayerCount; ++k) {
        const UNetLayerDesc &L = kUNetLayers[k];
        u.w_off[k] = o;
        u.b_off[k] = o + (size_t)L.taps * L.cout * L.cin;
        o = u.b_off[k] + L.cout;
        const uint32_t cinp = L.cin < 16 && L.taps == 27 ? 16u : (uint32_t)L.cin;
        u.wf_off[k] = ob;
        ob += (size_t)L.taps * L.cout * cinp;
        u.wt_off[k] = ob;
        ob += (size_t)L.taps * L.cout * L.cin;
    }
    ualloc(&u.w32, u.nparams);
    ualloc(&u.wbf, ob);
    ualloc(&u.sc, 1);
    const size_t n0 = level_cells(c, 0), n1 = level_cells(c, 1), n2 = level_cells(c, 2);
    ualloc(&u.x0, n0 * 16);
    ualloc(&u.a0, n0 * 16);
    ualloc(&u.s0, n0 * 16);
    ualloc(&u.t0, n0 * 16);
    ualloc(&u.d0, n0 * 16);
    ualloc(&u.p0, n1 * 16);
    ualloc(&u.a1, n1 * 32);
    ualloc(&u.s1, n1 * 32);
    ualloc(&u.t1, n1 * 32);
    ualloc(&u.d1, n1 * 32);
    ualloc(&u.c1, n1 * 16);
    ualloc(&u.p1, n2 * 32);
    ualloc(&u.a2, n2 * 64);
    ualloc(&u.b2, n2 * 64);
    ualloc(&u.c2, n2 * 32);
    ualloc(&u.zhat, n0);
    ualloc(&u.partials, 2 * (size_t)c->geo.npool + 2);
    ualloc(&u.zero_flag, 1);
    u.skip = u.zero_flag;
}

void unet_free(slosh_ctx *c)
{
    UNet &u = c->unet;
    for (void *p : {(void *)u.w32, (void *)u.wbf, (void *)u.sc, (void *)u.x0, (void *)u.a0, (void *)u.s0, (void *)u.t0,
                    (void *)u.d0, (void *)u.p0, (void *)u.a1, (void *)u.s1, (void *)u.t1, (void *)u.d1, (void *)u.c1,
                    (void *)u.p1, (void *)u.a2, (void *)u.b2, (void *)u.c2, (void *)u.zhat, (void *)u.partials,
                    (void *)u.grad, (void *)u.m, (void *)u.v, (void *)u.e, (void *)u.g16a, (void *)u.g16b,
                    (void *)u.g16c, (void *)u.g16_1a, (void *)u.g16_1b, (void *)u.g32a, (void *)u.g32b, (void *)u.g32c,
                    (void *)u.g32_2a, (void *)u.g32_2b, (void *)u.g64a, (void *)u.g64b, (void *)u.zero_flag})
        if (p) (void)hipFree(p);
    u = UNet{};
}

// The BF16 copies of every layer from the FP32 master weights.
void unet_upload_bf16(slosh_ctx *c)
{
    UNet &u = c->unet;
    for (int k = 0; k < kUNetLayerCount; ++k) {
        const UNetLayerDesc &L = kUNetLayers[k];
        const uint32_t cinp = L.cin < 16 && L.taps == 27 ? 16u : (uint32_t)L.cin;
        const uint32_t n = (uint32_t)L.taps * L.cout * cinp;
        hipLaunchKernelGGL(k_weights_bf16, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, u.w32 + u.w_off[k],
                           (uint32_t)L.taps, (uint32_t)L.cin, cinp, (uint32_t)L.cout, u.wbf + u.wf_off[k],
                           (cinp == (uint32_t)L.cin && L.taps == 27) ? u.wbf + u.wt_off[k] : nullptr);
    }
    SLOSH_HIP(hipGetLastError());
}

// z = N(r) (+ zadd): the network applied to the level-0 vector r (pool layout); z on fluid cells (zadd, or 0,
// elsewhere). The normalised head output is kept in u.zhat (training). Every kernel returns at once while *u.skip
// is set (the flexible-CG solve's done flag, so chunks enqueued past convergence cost only their launches).
void unet_apply(slosh_ctx *c, const float *r, float *z, const float *zadd)
{
    UNet &u = c->unet;
    const uint32_t nb = c->nactive;
    const uint16_t *W = u.wbf;
    auto wf = [&](int k) { return W + u.wf_off[k]; };
    auto bias = [&](int k) { return u.w32 + u.b_off[k]; };
    {
        SLOSH_TIMED(c, K_UNET_IO);
        hipLaunchKernelGGL(k_unet_norm, dim3(nb), dim3(256), 0, c->stream, c->binfo, c->fmask, r, u.partials, u.skip);
        hipLaunchKernelGGL(k_unet_finalize, dim3(1), dim3(256), 0, c->stream, u.partials, nb, 0, u.sc, u.skip);
        hipLaunchKernelGGL(k_unet_input, dim3(nb), dim3(256), 0, c->stream, c->binfo, c->ctype, r, u.sc, u.x0, u.skip);
    }
    conv<16, 16, 8, true, true>(c, u.x0, wf(0), bias(0), u.a0);
    conv<16, 16, 8, true, true>(c, u.a0, wf(1), bias(1), u.s0);
    {
        launch_timed(c, K_UNET_IO, (k_pool<16, 8>), dim3(chunks(c, 1, 16)), dim3(256), 0, c->binfo, nb, u.s0, u.p0,
                     u.skip);
    }
    conv<16, 32, 4, true, true>(c, u.p0, wf(2), bias(2), u.a1);
    conv<32, 32, 4, true, true>(c, u.a1, wf(3), bias(3), u.s1);
    {
        launch_timed(c, K_UNET_IO, (k_pool<32, 4>), dim3(chunks(c, 2, 32)), dim3(256), 0, c->binfo, nb, u.s1, u.p1,
                     u.skip);
    }
    conv<32, 64, 2, true, true>(c, u.p1, wf(4), bias(4), u.a2);
    conv<64, 64, 2, true, true>(c, u.a2, wf(5), bias(5), u.b2);
    conv<64, 32, 2, false, true>(c, u.b2, wf(6), bias(6), u.c2);
    {
        launch_timed(c, K_UNET_IO, (k_upadd<32, 4>), dim3(chunks(c, 1, 32)), dim3(256), 0, c->binfo, nb, u.c2, u.s1,
                     u.t1, u.skip);
    }
    conv<32, 32, 4, true, true>(c, u.t1, wf(7), bias(7), u.d1);
    conv<32, 16, 4, false, true>(c, u.d1, wf(8), bias(8), u.c1);
    {
        launch_timed(c, K_UNET_IO, (k_upadd<16, 8>), dim3(chunks(c, 0, 16)), dim3(256), 0, c->binfo, nb, u.c1, u.s0,
                     u.t0, u.skip);
    }
    conv<16, 16, 8, true, true>(c, u.t0, wf(9), bias(9), u.d0);
    launch_timed(c, K_UNET_IO, k_unet_head, dim3(nb), dim3(256), 0, c->binfo, c->fmask, u.d0, wf(10), bias(10), u.sc,
                 u.zhat, z, zadd, u.skip);
}

// Accumulates dL/dparams for the system (fmask, diag) and residual r into u.grad; the loss terms go to u.sc.
void unet_backward(slosh_ctx *c, const float *r)
{
    UNet &u = c->unet;
    const uint32_t nb = c->nactive;
    const size_t n0 = level_cells(c, 0), n1 = level_cells(c, 1), n2 = level_cells(c, 2);
    if (!u.grad) {
        ualloc(&u.grad, u.nparams);
        ualloc(&u.m, u.nparams);
        ualloc(&u.v, u.nparams);
        ualloc(&u.e, n0);
        ualloc(&u.g16a, n0 * 16);
        ualloc(&u.g16b, n0 * 16);
        ualloc(&u.g16c, n0 * 16);
        ualloc(&u.g16_1a, n1 * 16);
        ualloc(&u.g16_1b, n1 * 16);
        ualloc(&u.g32a, n1 * 32);
        ualloc(&u.g32b, n1 * 32);
        ualloc(&u.g32c, n1 * 32);
        ualloc(&u.g32_2a, n2 * 32);
        ualloc(&u.g32_2b, n2 * 32);
        ualloc(&u.g64a, n2 * 64);
        ualloc(&u.g64b, n2 * 64);
    }
    unet_apply(c, r, nullptr);
    const uint16_t *W = u.wbf;
    auto wt = [&](int k) { return W + u.wt_off[k]; };
    SLOSH_TIMED(c, K_UNET_TRAIN);
    hipLaunchKernelGGL(k_unet_loss_e, dim3(nb), dim3(256), 0, c->stream, c->binfo, c->fmask, c->diag, u.zhat, u.x0, u.e,
                       u.partials);
    hipLaunchKernelGGL(k_unet_finalize, dim3(1), dim3(256), 0, c->stream, u.partials, nb, 1, u.sc, u.skip);
    hipLaunchKernelGGL(k_unet_head_bwd, dim3(nb), dim3(256), 0, c->stream, c->binfo, c->fmask, c->diag, u.e, u.d0,
                       W + u.wf_off[10], u.sc, u.g16a, u.grad + u.w_off[10]);
    // D0b (input t0): g16a = dL/d(pre-activation of d0)
    conv_dw<16, 16, 8>(c, u.t0, u.g16a, 9);
    conv<16, 16, 8, false, false>(c, u.g16a, wt(9), nullptr, u.g16b); // dL/dt0
    hipLaunchKernelGGL((k_upadd_bwd<16, 8>), dim3(chunks(c, 1, 16)), dim3(256), 0, c->stream, c->binfo, nb, u.t0,
                       u.g16b, u.g16_1a); // g16b = gu0, g16_1a = dL/dc1
    // D0p (input d1, no activation)
    conv_dw<32, 16, 4>(c, u.d1, u.g16_1a, 8);
    conv<16, 32, 4, false, false>(c, u.g16_1a, wt(8), nullptr, u.g32a); // dL/dd1
    hipLaunchKernelGGL((k_relu_mask<32>), dim3(chunks(c, 1, 32)), dim3(256), 0, c->stream, c->binfo, nb,
                       (uint32_t)(kBrickCells >> 3), u.d1, u.g32a);
    // D1b (input t1)
    conv_dw<32, 32, 4>(c, u.t1, u.g32a, 7);
    conv<32, 32, 4, false, false>(c, u.g32a, wt(7), nullptr, u.g32b); // dL/dt1
    hipLaunchKernelGGL((k_upadd_bwd<32, 4>), dim3(chunks(c, 2, 32)), dim3(256), 0, c->stream, c->binfo, nb, u.t1,
                       u.g32b, u.g32_2a); // g32b = gu1, g32_2a = dL/dc2
    // D1p (input b2)
    conv_dw<64, 32, 2>(c, u.b2, u.g32_2a, 6);
    conv<32, 64, 2, false, false>(c, u.g32_2a, wt(6), nullptr, u.g64a); // dL/db2
    hipLaunchKernelGGL((k_relu_mask<64>), dim3(chunks(c, 2, 64)), dim3(256), 0, c->stream, c->binfo, nb,
                       (uint32_t)(kBrickCells >> 6), u.b2, u.g64a);
    // B2b (input a2)
    conv_dw<64, 64, 2>(c, u.a2, u.g64a, 5);
    conv<64, 64, 2, false, false>(c, u.g64a, wt(5), nullptr, u.g64b); // dL/da2
    hipLaunchKernelGGL((k_relu_mask<64>), dim3(chunks(c, 2, 64)), dim3(256), 0, c->stream, c->binfo, nb,
                       (uint32_t)(kBrickCells >> 6), u.a2, u.g64b);
    // B2a (input p1)
    conv_dw<32, 64, 2>(c, u.p1, u.g64b, 4);
    conv<64, 32, 2, false, false>(c, u.g64b, wt(4), nullptr, u.g32_2b); // dL/dp1
    hipLaunchKernelGGL((k_skip_grad<32, 4>), dim3(chunks(c, 1, 32)), dim3(256), 0, c->stream, c->binfo, nb, u.g32b,
                       u.g32_2b, u.s1, u.g32c); // g32c = dL/d(pre-activation of s1)
    // E1b (input a1)
    conv_dw<32, 32, 4>(c, u.a1, u.g32c, 3);
    conv<32, 32, 4, false, false>(c, u.g32c, wt(3), nullptr, u.g32a); // dL/da1
    hipLaunchKernelGGL((k_relu_mask<32>), dim3(chunks(c, 1, 32)), dim3(256), 0, c->stream, c->binfo, nb,
                       (uint32_t)(kBrickCells >> 3), u.a1, u.g32a);
    // E1a (input p0)
    conv_dw<16, 32, 4>(c, u.p0, u.g32a, 2);
    conv<32, 16, 4, false, false>(c, u.g32a, wt(2), nullptr, u.g16_1b); // dL/dp0
    hipLaunchKernelGGL((k_skip_grad<16, 8>), dim3(chunks(c, 0, 16)), dim3(256), 0, c->stream, c->binfo, nb, u.g16b,
                       u.g16_1b, u.s0, u.g16c); // g16c = dL/d(pre-activation of s0)
    // E0b (input a0)
    conv_dw<16, 16, 8>(c, u.a0, u.g16c, 1);
    conv<16, 16, 8, false, false>(c, u.g16c, wt(1), nullptr, u.g16a); // dL/da0
    hipLaunchKernelGGL((k_relu_mask<16>), dim3(chunks(c, 0, 16)), dim3(256), 0, c->stream, c->binfo, nb,
                       (uint32_t)kBrickCells, u.a0, u.g16a);
    // E0a (input x0): weights only (the padded input channels have zero weights and no gradient use)
    conv_dw<16, 16, 8, 4>(c, u.x0, u.g16a, 0);
    SLOSH_HIP(hipGetLastError());
    (void)n1;
    (void)n2;
}

void unet_adam(slosh_ctx *c, float lr, float scale)
{
    UNet &u = c->unet;
    u.adam_t++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)u.adam_t), c2 = 1.0f - powf(b2, (float)u.adam_t);
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(u.nparams, 256)), dim3(256), 0, c->stream, u.w32, u.grad, u.m, u.v,
                       (uint32_t)u.nparams, lr, b1, b2, 1e-8f, c1, c2, scale);
    unet_upload_bf16(c);
}

} // namespace slosh
