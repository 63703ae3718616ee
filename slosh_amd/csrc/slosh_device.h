// slosh_device.h -- device-side definitions shared by the gfx950 kernels of the solver.
//
// Grid storage is the reference's sparse brick pool (SloshDemo/Plugins/Slosh/Shaders/Private/Slosh.usf:98-199):
// the grid is tiled into 8^3 bricks, a brick table maps a brick to a pool slot of 512 cells (local index
// x + 8y + 64z), slot 0 is the all-zero null brick that unallocated and out-of-grid cells resolve to, and every
// cell owns its three min faces, stored as three planes of pool_cells entries. All arithmetic that the oracle
// restates is written in the same evaluation order; the library is compiled with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/slosh.h"

namespace slosh {

constexpr uint32_t kBrickCells = 512u;
constexpr uint32_t kNoBrick = 0xFFFFFFFFu;
constexpr uint32_t kCellAir = 0u;
constexpr uint32_t kCellFluid = 1u;
constexpr uint32_t kCellSolid = 2u;

// Ordering point for the compiler: memory operations written before it stay before it. The solver passes read their
// scalars (the brick table row, the solver's done flag, alpha or beta) and test the flag after this point, so all of
// them arrive in one memory latency; without it the compiler would sink the row's loads past the test (they are
// used only on the continuing path) and the row would wait for the flag. The coarse tail issues every global read
// before its test.
__device__ __forceinline__ void issue_point() { __asm__ volatile("" ::: "memory"); }

// Type of a coarse multigrid cell from its children (SloshPressureMG.usf:256-278, MGBuildTypeCS): AIR if any child
// is air, else FLUID if any child is fluid, else SOLID. With fluid_first, FLUID takes precedence over AIR.
__host__ __device__ constexpr uint8_t coarse_type(bool air, bool fluid, int fluid_first)
{
    return (uint8_t)(fluid_first ? (fluid ? kCellFluid : (air ? kCellAir : kCellSolid))
                                 : (air ? kCellAir : (fluid ? kCellFluid : kCellSolid)));
}
constexpr uint32_t kCellTypeMask = 3u;
constexpr uint32_t kCellDense = 4u;
constexpr uint32_t kColliderShift = 8u;
constexpr uint32_t kFaceValid = 1u;
constexpr uint32_t kFaceSolid = 2u;
constexpr uint32_t kFaceHasWeight = 4u;
constexpr uint32_t kFaceExtrapShift = 4u;
constexpr float kFixedScale = 1048576.0f;
constexpr float kFixedInv = 1.0f / 1048576.0f;
constexpr float kTargetPpc = 8.0f;
constexpr float kDiagEps = 1e-4f;
constexpr int kMaxLevels = 8;
constexpr uint32_t kCoarsestCells = 64u;
// P2G tile of one brick: nodes [brick_min - 1, brick_min + 9) per axis, for each of the 3 components.
constexpr int kTileDim = 10;
constexpr int kTileNodes = kTileDim * kTileDim * kTileDim;
constexpr int kTileInts = 6 * kTileNodes; // mom U,V,W then weight U,V,W

// Grid geometry and the brick pool, passed by value to every grid kernel.
struct Geo {
    int n[3];          // cells per axis
    int bd[3];         // bricks per axis
    uint32_t nbricks;
    uint32_t npool;    // pool slots including the null slot 0
    uint32_t pool_cells;
    float h, inv_h;
    const uint32_t *table;  // brick -> slot (0 = none)
    const uint32_t *active; // active bricks, ascending
};

__device__ __forceinline__ bool in_grid(const Geo &g, int x, int y, int z)
{
    return x >= 0 && y >= 0 && z >= 0 && x < g.n[0] && y < g.n[1] && z < g.n[2];
}

__device__ __forceinline__ uint32_t brick_of_cell(const Geo &g, int x, int y, int z)
{
    return (uint32_t)((x >> 3) + g.bd[0] * ((y >> 3) + g.bd[1] * (z >> 3)));
}

__device__ __forceinline__ uint32_t local_of_cell(int x, int y, int z)
{
    return (uint32_t)((x & 7) + 8 * ((y & 7) + 8 * (z & 7)));
}

// Slosh.usf:136-144 (CellIndex)
__device__ __forceinline__ uint32_t cell_index(const Geo &g, int x, int y, int z)
{
    uint32_t slot = 0u;
    if (in_grid(g, x, y, z)) slot = g.table[brick_of_cell(g, x, y, z)];
    return slot * kBrickCells + local_of_cell(x, y, z);
}

__device__ __forceinline__ void brick_coord(const Geo &g, uint32_t b, int c[3])
{
    c[0] = (int)(b % (uint32_t)g.bd[0]);
    c[1] = (int)((b / (uint32_t)g.bd[0]) % (uint32_t)g.bd[1]);
    c[2] = (int)(b / (uint32_t)(g.bd[0] * g.bd[1]));
}

__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ __forceinline__ int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }
__device__ __forceinline__ float len3(float x, float y, float z) { return sqrtf(x * x + y * y + z * z); }
__device__ __forceinline__ void cross3(const float a[3], const float b[3], float o[3])
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// Slosh.usf:357 / :518: clamped cell of a position.
__device__ __forceinline__ void pos_cell(const Geo &g, float px, float py, float pz, int c[3])
{
    c[0] = clampi((int)floorf(px * g.inv_h), 0, g.n[0] - 1);
    c[1] = clampi((int)floorf(py * g.inv_h), 0, g.n[1] - 1);
    c[2] = clampi((int)floorf(pz * g.inv_h), 0, g.n[2] - 1);
}

// SloshCommon.ush:254-265
__device__ __forceinline__ uint32_t pcg_hash(uint32_t v)
{
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
__device__ __forceinline__ float hash_to_unit(uint32_t h) { return (float)(h & 0x00FFFFFFu) / 16777216.0f; }

// SloshCommon.ush:137-146
__device__ __forceinline__ void quat_rotate(const float q[4], const float v[3], float out[3])
{
    float t[3], c[3];
    t[0] = 2.0f * (q[1] * v[2] - q[2] * v[1]);
    t[1] = 2.0f * (q[2] * v[0] - q[0] * v[2]);
    t[2] = 2.0f * (q[0] * v[1] - q[1] * v[0]);
    c[0] = q[1] * t[2] - q[2] * t[1];
    c[1] = q[2] * t[0] - q[0] * t[2];
    c[2] = q[0] * t[1] - q[1] * t[0];
    for (int k = 0; k < 3; ++k) out[k] = v[k] + q[3] * t[k] + c[k];
}

// SloshCommon.ush:178-195
__device__ inline float collider_sdf(const slosh_collider &c, const float p[3])
{
    float rel[3] = {p[0] - c.center[0], p[1] - c.center[1], p[2] - c.center[2]};
    float qi[4] = {-c.rotation[0], -c.rotation[1], -c.rotation[2], c.rotation[3]};
    float l[3];
    quat_rotate(qi, rel, l);
    if (c.type == 0u) return len3(l[0], l[1], l[2]) - c.extent[0];
    if (c.type == 1u) {
        float d[3], m[3];
        for (int k = 0; k < 3; ++k) {
            d[k] = fabsf(l[k]) - c.extent[k];
            m[k] = fmaxf(d[k], 0.0f);
        }
        return len3(m[0], m[1], m[2]) + fminf(fmaxf(d[0], fmaxf(d[1], d[2])), 0.0f);
    }
    l[2] -= clampf(l[2], -c.extent[1], c.extent[1]);
    return len3(l[0], l[1], l[2]) - c.extent[0];
}

// SloshCommon.ush:197-205
__device__ inline void collider_normal(const slosh_collider &c, const float p[3], float eps, float n[3])
{
    for (int a = 0; a < 3; ++a) {
        float pp[3] = {p[0], p[1], p[2]}, pm[3] = {p[0], p[1], p[2]};
        pp[a] = p[a] + eps;
        pm[a] = p[a] - eps;
        n[a] = collider_sdf(c, pp) - collider_sdf(c, pm);
    }
    float l = len3(n[0], n[1], n[2]);
    if (l > 1e-8f) {
        for (int a = 0; a < 3; ++a) n[a] = n[a] / l;
    } else {
        n[0] = 0.0f;
        n[1] = 0.0f;
        n[2] = 1.0f;
    }
}

// SloshCommon.ush:232-235, one component.
__device__ __forceinline__ float collider_velocity_axis(const slosh_collider &c, const float p[3], int axis)
{
    float r[3] = {p[0] - c.center[0], p[1] - c.center[1], p[2] - c.center[2]};
    const float *w = c.angular_velocity;
    float cr0 = w[1] * r[2] - w[2] * r[1];
    float cr1 = w[2] * r[0] - w[0] * r[2];
    float cr2 = w[0] * r[1] - w[1] * r[0];
    float cr = axis == 0 ? cr0 : (axis == 1 ? cr1 : cr2);
    return c.linear_velocity[axis] + cr;
}

// SloshCommon.ush:207-230 (ManipulatorSDF, ManipulatorWeight): the shape's signed distance (shape 3 everywhere: -1),
// and the effect's weight, 1 inside the shape fading linearly to 0 over `falloff` beyond it.
__device__ inline float manipulator_sdf(const slosh_manipulator &m, const float p[3])
{
    if (m.shape == 3u) return -1.0f;
    slosh_collider c{};
    for (int a = 0; a < 3; ++a) {
        c.center[a] = m.center[a];
        c.extent[a] = m.extent[a];
    }
    for (int k = 0; k < 4; ++k) c.rotation[k] = m.rotation[k];
    c.type = m.shape;
    return collider_sdf(c, p);
}

__device__ inline float manipulator_weight(const slosh_manipulator &m, const float p[3])
{
    const float d = manipulator_sdf(m, p);
    if (d <= 0.0f) return 1.0f;
    return m.falloff > 0.0f ? clampf(1.0f - d / m.falloff, 0.0f, 1.0f) : 0.0f;
}

// Slosh.usf:1523-1559: the force fields (every kind but drains) acting on the particle's new velocity nv at x.
__device__ inline void apply_force_fields(const slosh_manipulator *man, uint32_t nman, const float x[3], float nv[3],
                                          float dt)
{
    for (uint32_t k = 0; k < nman; ++k) {
        const slosh_manipulator m = man[k];
        if (m.kind == 0u) continue;
        const float w = manipulator_weight(m, x);
        if (w <= 0.0f) continue;
        const float r[3] = {x[0] - m.center[0], x[1] - m.center[1], x[2] - m.center[2]};
        const float f = m.strength * w * dt;
        if (m.kind == 1u) {
            for (int a = 0; a < 3; ++a) nv[a] = nv[a] + m.axis[a] * f;
        } else if (m.kind == 2u) {
            const float rl = len3(r[0], r[1], r[2]);
            for (int a = 0; a < 3; ++a) nv[a] = nv[a] + (rl > 1e-5f ? r[a] / rl : 0.0f) * f;
        } else if (m.kind == 3u) {
            const float ra = r[0] * m.axis[0] + r[1] * m.axis[1] + r[2] * m.axis[2];
            float perp[3];
            for (int a = 0; a < 3; ++a) perp[a] = r[a] - ra * m.axis[a];
            const float pl = len3(perp[0], perp[1], perp[2]);
            if (pl > 1e-5f) {
                float pu[3], tg[3];
                for (int a = 0; a < 3; ++a) pu[a] = perp[a] / pl;
                cross3(m.axis, pu, tg);
                for (int a = 0; a < 3; ++a) nv[a] = nv[a] + (tg[a] - pu[a] * m.param0) * f;
            }
        } else if (m.kind == 4u) {
            const float e = expf(-m.strength * w * dt);
            for (int a = 0; a < 3; ++a) nv[a] = nv[a] * e;
        }
    }
}

// Slosh.usf:1639-1654: whether a drain removes the particle (index i) at x: at once (param0 = 0), else with
// probability 1 - exp(-param0 dt) from the PCG hash of the particle, the frame and the drain.
__device__ inline bool drained_at(const slosh_manipulator *man, uint32_t nman, const float x[3], float dt, uint32_t i,
                                  uint32_t frame_seed)
{
    for (uint32_t k = 0; k < nman; ++k) {
        const slosh_manipulator m = man[k];
        if (m.kind != 0u || manipulator_sdf(m, x) > 0.0f) continue;
        const float chance = m.param0 > 0.0f ? 1.0f - expf(-m.param0 * dt) : 1.0f;
        if (chance >= 1.0f || hash_to_unit(pcg_hash(i * 2654435761u + frame_seed * 97u + k)) < chance) return true;
    }
    return false;
}

// SloshCommon.ush:244-252
__device__ __forceinline__ float ghost_theta(float rf, float ra)
{
    float denom = rf - ra;
    if (denom <= 1e-4f) return 1.0f;
    return clampf((rf - 0.5f) / denom, 0.05f, 1.0f);
}

__device__ __forceinline__ bool is_red(int x, int y, int z) { return (((unsigned)(x + y + z)) & 1u) == 0u; }

// Neighbour order -x, +x, -y, +y, -z, +z (Slosh.usf:732-735).
__device__ __forceinline__ void nb_offset(int nb, int d[3])
{
    d[0] = 0;
    d[1] = 0;
    d[2] = 0;
    d[nb >> 1] = (nb & 1) ? 1 : -1;
}

} // namespace slosh
