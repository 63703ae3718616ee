/*
 * oracle.c -- CPU restatement of the reference's per-step liquid solver. TEST INFRASTRUCTURE ONLY: the checker
 * of the HIP solver, loaded by the tests, the smoke check and the CPU-baseline leg of the benchmark, never by the
 * product library.
 *
 * Reference files (paths under SloshDemo/Plugins/Slosh/):
 *   Shaders/Private/SloshCommon.ush    flags, collider SDF, ghost theta, PCG hash
 *   Shaders/Private/Slosh.usf          emission, brick pool, classification, P2G, normalise, divergence, RB-SOR,
 *                                      projection, extrapolation, G2P/advection
 *   Shaders/Private/SloshPressureMG.usf MGPCG kernels
 *   Shaders/Private/SloshSort.usf      counting sort (key and compaction semantics)
 *   Source/SloshCore/Private/SloshProxy.cpp  pass order, MG hierarchy, constants
 *
 * Semantics this build fixes where the reference is nondeterministic or silent (each restated below):
 *   - sort: stable by 8^3-brick key, dead particles last (reference: atomic rank over 4^3 keys, unstable);
 *   - brick pool: freed slots return to the pool, wanted bricks take the lowest free slots in ascending brick
 *     order; the active list is in ascending brick order (reference: atomic stack and append);
 *   - one pressure operator for both solvers: (CellDiag + 1e-4) p_c - sum_{fluid n} p_n = -Divergence;
 *   - MGPCG stops at ||r||_2 <= tol ||b||_2 with FP64 dots and alpha/beta; its coarsest level (<= 64 cells)
 *     is solved exactly in FP64 (Cholesky) instead of 16 red-black sweeps;
 *   - HLSL round() is round-half-to-even (rintf); lerp(1-f, f, o) with o in {0,1} is the exact select.
 * Parity of the FP32 bits against the reference's own GPU output is unpinned: no run of it exists.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define BRICK_CELLS 512u
#define NO_BRICK 0xFFFFFFFFu
#define CELL_AIR 0u
#define CELL_FLUID 1u
#define CELL_SOLID 2u
#define CELL_TYPE_MASK 3u
#define CELL_DENSE_FLAG 4u
#define CELL_COLLIDER_SHIFT 8u
#define FACE_VALID 1u
#define FACE_SOLID 2u
#define FACE_HAS_WEIGHT 4u
#define FACE_EXTRAP_SHIFT 4u
#define FIXED_SCALE 1048576.0f
#define FIXED_INV (1.0f / 1048576.0f)
#define TARGET_PPC 8.0f
#define DIAG_EPS 1e-4f
#define MG_MAX_LEVELS 8
#define MG_COARSEST_CELLS 64u

typedef struct mglevel {
    int dim[3];
    uint32_t cells;
    uint32_t *type;   /* levels >= 1 only */
    float *x, *b;     /* FP32 preconditioner (levels >= 1) */
    double *xd, *bd;  /* FP64 judge solve (levels >= 1) */
} mglevel;

struct osim {
    slosh_settings s;
    int n[3];
    float h, inv_h;
    int bd[3];
    uint32_t nbricks, pool_slots, npool;
    size_t pool_cells;
    /* particles, current order */
    uint32_t np, cap;
    slosh_particle *p, *tmp;
    /* brick pool */
    uint32_t *table, *slot_brick, *active, nactive;
    uint8_t *touch, *want;
    /* grid */
    int32_t *accum;
    float *vel[3];
    int cur;
    uint32_t *fflags;
    float *fweight;
    uint32_t *ctype;
    int32_t *ccount;
    double *pres; /* FP64: the pressure solution is the CG alpha-update accumulator */
    double *pres_prev; /* SLOSH_FLAG_WARM_START_EXTRAPOLATE: last solution per unit dt (NaN: none) */
    uint8_t *had;      /* scratch of stage_p2g: the cell was fluid in the last step */
    uint8_t *xmark;    /* scratch of stage_extrapolate: faces filled by the current layer */
    /* SLOSH_FLAG_DENSITY_PROJECTION: right-hand side, potential q and face displacements */
    float *drhs;
    double *dq;
    float *ddisp;
    int dens_applied;
    float *div, *diag;
    /* solver */
    int nlevels;
    mglevel lv[MG_MAX_LEVELS];
    float *r, *z, *d, *q;
    double *rd, *zd, *dd, *qd, *pd;
    /* coarsest-level direct solve */
    uint32_t cn;        /* fluid unknowns */
    uint32_t *cidx;     /* level index of unknown i */
    double *chol;       /* cn x cn lower factor */
    double *part;       /* per-block partial sums of the parallel reductions (fixed block order) */
    size_t part_cap;
    slosh_collider col[SLOSH_MAX_COLLIDERS];
    uint32_t ncol;
    slosh_manipulator man[SLOSH_MAX_MANIPULATORS];
    uint32_t nman;
    uint64_t drained_total;
    int32_t coupling[SLOSH_MAX_COLLIDERS * 16]; /* fixed-point collider coupling sums of the last step */
    uint32_t coupling_ncol;
    slosh_results res;
    uint64_t frame;
    float prev_dt;
};

/* ------------------------------------------------------------------------------------------------------------ */
/* SloshCommon.ush:254-265 -- PCG hash (Jarzynski & Olano 2020) and its unit float.                            */
uint32_t oracle_pcg_hash(uint32_t v)
{
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

float oracle_hash_to_unit(uint32_t h) { return (float)(h & 0x00FFFFFFu) / 16777216.0f; }

static float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
static int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }
static float len3(const float v[3]) { return sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

/* SloshCommon.ush:137-146 */
static void quat_rotate(const float q[4], const float v[3], float out[3])
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

/* SloshCommon.ush:178-195 */
float oracle_collider_sdf(const slosh_collider *c, const float p[3])
{
    float rel[3] = {p[0] - c->center[0], p[1] - c->center[1], p[2] - c->center[2]};
    float qi[4] = {-c->rotation[0], -c->rotation[1], -c->rotation[2], c->rotation[3]};
    float l[3];
    quat_rotate(qi, rel, l);
    if (c->type == 0u) return len3(l) - c->extent[0];
    if (c->type == 1u) {
        float d[3], m[3];
        for (int k = 0; k < 3; ++k) {
            d[k] = fabsf(l[k]) - c->extent[k];
            m[k] = fmaxf(d[k], 0.0f);
        }
        return len3(m) + fminf(fmaxf(d[0], fmaxf(d[1], d[2])), 0.0f);
    }
    l[2] -= clampf(l[2], -c->extent[1], c->extent[1]);
    return len3(l) - c->extent[0];
}

/* SloshCommon.ush:197-205 */
static void collider_normal(const slosh_collider *c, const float p[3], float eps, float n[3])
{
    for (int a = 0; a < 3; ++a) {
        float pp[3] = {p[0], p[1], p[2]}, pm[3] = {p[0], p[1], p[2]};
        pp[a] = p[a] + eps;
        pm[a] = p[a] - eps;
        n[a] = oracle_collider_sdf(c, pp) - oracle_collider_sdf(c, pm);
    }
    float l = len3(n);
    if (l > 1e-8f) {
        for (int a = 0; a < 3; ++a) n[a] = n[a] / l;
    } else {
        n[0] = 0.0f; n[1] = 0.0f; n[2] = 1.0f;
    }
}

/* SloshCommon.ush:232-235, the component along one axis. */
static float collider_velocity_axis(const slosh_collider *c, const float p[3], int axis)
{
    float r[3] = {p[0] - c->center[0], p[1] - c->center[1], p[2] - c->center[2]};
    const float *w = c->angular_velocity;
    float cr[3] = {w[1] * r[2] - w[2] * r[1], w[2] * r[0] - w[0] * r[2], w[0] * r[1] - w[1] * r[0]};
    return c->linear_velocity[axis] + cr[axis];
}

void oracle_collider_normal(const slosh_collider *c, const float p[3], float eps, float n[3])
{
    collider_normal(c, p, eps, n);
}

void oracle_collider_velocity(const slosh_collider *c, const float p[3], float v[3])
{
    for (int a = 0; a < 3; ++a) v[a] = collider_velocity_axis(c, p, a);
}

/* Slosh.usf:1578-1591 (AdvectParticle's collider response; adhesion off): a particle closer than `radius` to a
 * collider's surface is pushed out along the finite-difference normal (eps = cell_size / 4) to the radius, its
 * inward velocity relative to the moving surface is removed and the tangential part scaled by (1 - friction). */
void oracle_collider_response(const slosh_collider *cols, uint32_t ncol, float radius, float friction,
                              float cell_size, float x[3], float nv[3])
{
    for (uint32_t k = 0; k < ncol; ++k) {
        const slosh_collider *cc = &cols[k];
        float dd = oracle_collider_sdf(cc, x);
        if (dd < radius) {
            float nn[3];
            collider_normal(cc, x, 0.25f * cell_size, nn);
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
}

/* SloshCommon.ush:244-252 */
float oracle_ghost_theta(float rho_fluid, float rho_air)
{
    float denom = rho_fluid - rho_air;
    if (denom <= 1e-4f) return 1.0f;
    return clampf((rho_fluid - 0.5f) / denom, 0.05f, 1.0f);
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Addressing, Slosh.usf:106-199.                                                                                */
static int in_grid(const osim *o, int x, int y, int z)
{
    return x >= 0 && y >= 0 && z >= 0 && x < o->n[0] && y < o->n[1] && z < o->n[2];
}

static uint32_t brick_of_cell(const osim *o, int x, int y, int z)
{
    return (uint32_t)((x >> 3) + o->bd[0] * ((y >> 3) + o->bd[1] * (z >> 3)));
}

static uint32_t cell_index(const osim *o, int x, int y, int z)
{
    uint32_t slot = 0u;
    if (in_grid(o, x, y, z)) slot = o->table[brick_of_cell(o, x, y, z)];
    return slot * BRICK_CELLS + (uint32_t)((x & 7) + 8 * ((y & 7) + 8 * (z & 7)));
}

static uint32_t cell_type_at(const osim *o, int x, int y, int z)
{
    return in_grid(o, x, y, z) ? o->ctype[cell_index(o, x, y, z)] : CELL_SOLID;
}

static uint32_t face_index(const osim *o, int axis, int x, int y, int z)
{
    return (uint32_t)axis * (uint32_t)o->pool_cells + cell_index(o, x, y, z);
}

static void brick_coord(const osim *o, uint32_t b, int c[3])
{
    c[0] = (int)(b % (uint32_t)o->bd[0]);
    c[1] = (int)((b / (uint32_t)o->bd[0]) % (uint32_t)o->bd[1]);
    c[2] = (int)(b / (uint32_t)(o->bd[0] * o->bd[1]));
}

/* Clamped cell of a position (Slosh.usf:357, :518). */
static void pos_cell(const osim *o, const float p[3], int c[3])
{
    for (int a = 0; a < 3; ++a) c[a] = clampi((int)floorf(p[a] * o->inv_h), 0, o->n[a] - 1);
}

static const int NB[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};

/* ------------------------------------------------------------------------------------------------------------ */
static void *xcalloc(size_t n, size_t sz) { return n ? calloc(n, sz) : calloc(1, sz); }

/* Atomic integer add and byte store of the parallel particle loops (integer sums are order-free). */
static void add_i32(int32_t *p, int32_t v) { __atomic_fetch_add(p, v, __ATOMIC_RELAXED); }
static void set_u8(uint8_t *p) { __atomic_store_n(p, (uint8_t)1u, __ATOMIC_RELAXED); }

static void free_levels(osim *o)
{
    for (int l = 1; l < o->nlevels; ++l) {
        free(o->lv[l].type); free(o->lv[l].x); free(o->lv[l].b); free(o->lv[l].xd); free(o->lv[l].bd);
    }
    free(o->cidx); free(o->chol);
    o->cidx = NULL; o->chol = NULL;
}

void osim_destroy(osim *o)
{
    if (!o) return;
    free(o->p); free(o->tmp);
    free(o->table); free(o->slot_brick); free(o->active); free(o->touch); free(o->want);
    free(o->accum);
    for (int k = 0; k < 3; ++k) free(o->vel[k]);
    free(o->fflags); free(o->fweight); free(o->ctype); free(o->ccount);
    free(o->pres); free(o->pres_prev); free(o->had); free(o->xmark); free(o->div); free(o->diag);
    free(o->drhs); free(o->dq); free(o->ddisp);
    free(o->r); free(o->z); free(o->d); free(o->q);
    free(o->rd); free(o->zd); free(o->dd); free(o->qd); free(o->pd);
    free(o->part);
    free_levels(o);
    free(o);
}

/* SloshProxy.cpp:185-306 (Allocate): pool of ceil(budget * bricks) slots plus the null slot; MG hierarchy halved
 * (rounding up) until a level holds <= 64 cells or 8 levels exist (reference: <= 512 cells, SloshProxy.cpp:269-285). */
osim *osim_create(const slosh_settings *s)
{
    if (!s) return NULL;
    for (int a = 0; a < 3; ++a)
        if (s->grid_dim[a] < 4 || s->grid_dim[a] > 4096) return NULL;
    if (!(s->cell_size > 0.0f) || !(s->sparse_storage_budget > 0.0f) || s->sparse_storage_budget > 1.0f) return NULL;
    osim *o = (osim *)calloc(1, sizeof(osim));
    if (!o) return NULL;
    o->s = *s;
    for (int a = 0; a < 3; ++a) {
        o->n[a] = s->grid_dim[a];
        o->bd[a] = (s->grid_dim[a] + 7) / 8;
    }
    o->h = s->cell_size;
    o->inv_h = 1.0f / s->cell_size;
    o->nbricks = (uint32_t)o->bd[0] * (uint32_t)o->bd[1] * (uint32_t)o->bd[2];
    double want_slots = ceil((double)s->sparse_storage_budget * (double)o->nbricks);
    o->pool_slots = (uint32_t)want_slots;
    if (o->pool_slots < 1u) o->pool_slots = 1u;
    if (o->pool_slots > o->nbricks) o->pool_slots = o->nbricks;
    o->npool = o->pool_slots + 1u;
    o->pool_cells = (size_t)o->npool * BRICK_CELLS;
    o->cap = s->particle_capacity;
    size_t pc = o->pool_cells;
    o->p = (slosh_particle *)xcalloc(o->cap, sizeof(slosh_particle));
    o->tmp = (slosh_particle *)xcalloc(o->cap, sizeof(slosh_particle));
    o->table = (uint32_t *)xcalloc(o->nbricks, 4);
    o->slot_brick = (uint32_t *)xcalloc(o->npool, 4);
    o->active = (uint32_t *)xcalloc(o->nbricks, 4);
    o->touch = (uint8_t *)xcalloc(o->nbricks, 1);
    o->want = (uint8_t *)xcalloc(o->nbricks, 1);
    o->accum = (int32_t *)xcalloc(6 * pc, 4);
    for (int k = 0; k < 3; ++k) o->vel[k] = (float *)xcalloc(3 * pc, 4);
    o->fflags = (uint32_t *)xcalloc(3 * pc, 4);
    o->fweight = (float *)xcalloc(3 * pc, 4);
    o->ctype = (uint32_t *)xcalloc(pc, 4);
    o->ccount = (int32_t *)xcalloc(pc, 4);
    o->pres = (double *)xcalloc(pc, 8);
    if (o->s.flags & SLOSH_FLAG_WARM_START_EXTRAPOLATE) {
        o->pres_prev = (double *)xcalloc(pc, 8);
        o->had = (uint8_t *)xcalloc(pc, 1);
    }
    o->div = (float *)xcalloc(pc, 4);
    o->diag = (float *)xcalloc(pc, 4);
    o->xmark = (uint8_t *)xcalloc(3 * pc, 1);
    if (o->s.flags & SLOSH_FLAG_DENSITY_PROJECTION) {
        o->drhs = (float *)xcalloc(pc, 4);
        o->dq = (double *)xcalloc(pc, 8);
        o->ddisp = (float *)xcalloc(3 * pc, 4);
    }
    o->r = (float *)xcalloc(pc, 4);
    o->z = (float *)xcalloc(pc, 4);
    o->d = (float *)xcalloc(pc, 4);
    o->q = (float *)xcalloc(pc, 4);
    o->rd = (double *)xcalloc(pc, 8);
    o->zd = (double *)xcalloc(pc, 8);
    o->dd = (double *)xcalloc(pc, 8);
    o->qd = (double *)xcalloc(pc, 8);
    o->pd = (double *)xcalloc(pc, 8);
    if (!o->p || !o->tmp || !o->table || !o->slot_brick || !o->accum || !o->vel[2] || !o->pd || !o->xmark) {
        osim_destroy(o);
        return NULL;
    }
    for (uint32_t sl = 0; sl < o->npool; ++sl) o->slot_brick[sl] = NO_BRICK;
    /* MG hierarchy */
    o->nlevels = 1;
    o->lv[0].dim[0] = o->n[0]; o->lv[0].dim[1] = o->n[1]; o->lv[0].dim[2] = o->n[2];
    o->lv[0].cells = (uint32_t)o->n[0] * (uint32_t)o->n[1] * (uint32_t)o->n[2];
    while (o->lv[o->nlevels - 1].cells > MG_COARSEST_CELLS && o->nlevels < MG_MAX_LEVELS) {
        mglevel *pv = &o->lv[o->nlevels - 1];
        mglevel *lv = &o->lv[o->nlevels];
        for (int a = 0; a < 3; ++a) lv->dim[a] = (pv->dim[a] + 1) / 2;
        lv->cells = (uint32_t)lv->dim[0] * (uint32_t)lv->dim[1] * (uint32_t)lv->dim[2];
        lv->type = (uint32_t *)xcalloc(lv->cells, 4);
        lv->x = (float *)xcalloc(lv->cells, 4);
        lv->b = (float *)xcalloc(lv->cells, 4);
        lv->xd = (double *)xcalloc(lv->cells, 8);
        lv->bd = (double *)xcalloc(lv->cells, 8);
        o->nlevels++;
    }
    {
        size_t blocks = o->nbricks;
        if (o->nlevels > 1 && (size_t)(o->lv[1].cells + 511u) / 512u > blocks) blocks = (o->lv[1].cells + 511u) / 512u;
        o->part_cap = blocks + 1u;
        o->part = (double *)xcalloc(4u * o->part_cap, 8);
        if (!o->part) {
            osim_destroy(o);
            return NULL;
        }
    }
    o->res.solver_iterations = 0;
    return o;
}

/* Threads of the parallel loops (OpenMP); results do not depend on it: floating-point reductions sum fixed blocks
 * of cells in a fixed order, integer ones are order-free. */
void oracle_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int oracle_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int osim_set_colliders(osim *o, const slosh_collider *c, uint32_t n)
{
    if (n > SLOSH_MAX_COLLIDERS) return SLOSH_ECAPACITY;
    if (n && c) memcpy(o->col, c, n * sizeof(slosh_collider));
    o->ncol = n;
    return SLOSH_OK;
}

int osim_set_manipulators(osim *o, const slosh_manipulator *m, uint32_t n)
{
    if (n > SLOSH_MAX_MANIPULATORS) return SLOSH_ECAPACITY;
    if (n && m) memcpy(o->man, m, n * sizeof(slosh_manipulator));
    o->nman = n;
    return SLOSH_OK;
}

static void cross3(const float a[3], const float b[3], float out[3])
{
    out[0] = a[1] * b[2] - a[2] * b[1];
    out[1] = a[2] * b[0] - a[0] * b[2];
    out[2] = a[0] * b[1] - a[1] * b[0];
}

/* SloshCommon.ush:207-230 (ManipulatorSDF, ManipulatorWeight). */
static float manipulator_sdf(const slosh_manipulator *m, const float p[3])
{
    if (m->shape == 3u) return -1.0f;
    slosh_collider c;
    memset(&c, 0, sizeof(c));
    for (int a = 0; a < 3; ++a) {
        c.center[a] = m->center[a];
        c.extent[a] = m->extent[a];
    }
    for (int k = 0; k < 4; ++k) c.rotation[k] = m->rotation[k];
    c.type = m->shape;
    return oracle_collider_sdf(&c, p);
}

static float manipulator_weight(const slosh_manipulator *m, const float p[3])
{
    float d = manipulator_sdf(m, p);
    if (d <= 0.0f) return 1.0f;
    return m->falloff > 0.0f ? clampf(1.0f - d / m->falloff, 0.0f, 1.0f) : 0.0f;
}

/* Slosh.usf:1523-1559: force fields on the new velocity nv of a particle at x. */
static void force_fields(const osim *o, const float x[3], float nv[3], float dt)
{
    for (uint32_t k = 0; k < o->nman; ++k) {
        const slosh_manipulator *m = &o->man[k];
        if (m->kind == 0u) continue;
        float w = manipulator_weight(m, x);
        if (w <= 0.0f) continue;
        float r[3] = {x[0] - m->center[0], x[1] - m->center[1], x[2] - m->center[2]};
        float f = m->strength * w * dt;
        if (m->kind == 1u) {
            for (int a = 0; a < 3; ++a) nv[a] = nv[a] + m->axis[a] * f;
        } else if (m->kind == 2u) {
            float rl = len3(r);
            for (int a = 0; a < 3; ++a) nv[a] = nv[a] + (rl > 1e-5f ? r[a] / rl : 0.0f) * f;
        } else if (m->kind == 3u) {
            float ra = r[0] * m->axis[0] + r[1] * m->axis[1] + r[2] * m->axis[2];
            float perp[3];
            for (int a = 0; a < 3; ++a) perp[a] = r[a] - ra * m->axis[a];
            float pl = len3(perp);
            if (pl > 1e-5f) {
                float pu[3], tg[3];
                for (int a = 0; a < 3; ++a) pu[a] = perp[a] / pl;
                cross3(m->axis, pu, tg);
                for (int a = 0; a < 3; ++a) nv[a] = nv[a] + (tg[a] - pu[a] * m->param0) * f;
            }
        } else if (m->kind == 4u) {
            float e = expf(-m->strength * w * dt);
            for (int a = 0; a < 3; ++a) nv[a] = nv[a] * e;
        }
    }
}

/* Slosh.usf:1639-1654: whether a drain removes particle i at x (frame seed = frame * 16 + substep). */
static int drained_at(const osim *o, const float x[3], float dt, uint32_t i, uint32_t frame_seed)
{
    for (uint32_t k = 0; k < o->nman; ++k) {
        const slosh_manipulator *m = &o->man[k];
        if (m->kind != 0u || manipulator_sdf(m, x) > 0.0f) continue;
        float chance = m->param0 > 0.0f ? 1.0f - expf(-m->param0 * dt) : 1.0f;
        if (chance >= 1.0f || oracle_hash_to_unit(oracle_pcg_hash(i * 2654435761u + frame_seed * 97u + k)) < chance)
            return 1;
    }
    return 0;
}

static void touch_pos(osim *o, const float p[3])
{
    int c[3];
    pos_cell(o, p, c);
    set_u8(&o->touch[brick_of_cell(o, c[0], c[1], c[2])]);
}

/* Slosh.usf:280-358 (EmitParticlesCS). Mode 0 FillCells: particle I goes to cell I/8, sub-cell I%8 of a 2x2x2
 * stratification, jittered by 0.8 of the sub-cell; mode 1: a uniform point in the sphere of radius extent[0]
 * (cube-root radius, uniform direction); mode 2: a uniform point in the rotated box. Then the optional cone spread
 * and speed jitter of the velocity; clamped to [1.01h, (N-1.01)h]; C = 0; age 0. */
int osim_emit(osim *o, const slosh_emit_request *r)
{
    if (!r || r->shape > 2u) return SLOSH_EINVAL;
    uint64_t count = 0;
    if (r->shape == SLOSH_EMIT_FILL_CELLS) {
        for (int a = 0; a < 3; ++a)
            if (r->cell_dims[a] < 0) return SLOSH_EINVAL;
        count = (uint64_t)r->cell_dims[0] * (uint64_t)r->cell_dims[1] * (uint64_t)r->cell_dims[2] * 8u;
    } else {
        count = r->count;
    }
    if (count == 0) return SLOSH_OK;
    if ((uint64_t)o->np + count > (uint64_t)o->cap) return SLOSH_ECAPACITY;
    float lo = o->h * 1.01f;
    float hi[3];
    for (int a = 0; a < 3; ++a) hi[a] = ((float)o->n[a] - 1.01f) * o->h;
    uint32_t dx = (uint32_t)r->cell_dims[0], dy = (uint32_t)r->cell_dims[1];
#pragma omp parallel for schedule(static)
    for (uint32_t i = 0; i < (uint32_t)count; ++i) {
        uint32_t h0 = oracle_pcg_hash(i * 3u + r->seed * 7919u);
        uint32_t h1 = oracle_pcg_hash(h0);
        uint32_t h2 = oracle_pcg_hash(h1);
        float rnd[3] = {oracle_hash_to_unit(h0), oracle_hash_to_unit(h1), oracle_hash_to_unit(h2)};
        float pos[3];
        if (r->shape == SLOSH_EMIT_FILL_CELLS) {
            uint32_t cl = i / 8u, sub = i % 8u;
            int local[3] = {(int)(cl % dx), (int)((cl / dx) % dy), (int)(cl / (dx * dy))};
            float so[3] = {(float)(sub & 1u), (float)((sub >> 1u) & 1u), (float)((sub >> 2u) & 1u)};
            for (int a = 0; a < 3; ++a) {
                float in_cell = ((so[a] + 0.5f) + 0.8f * (rnd[a] - 0.5f)) * 0.5f;
                pos[a] = ((float)(r->cell_min[a] + local[a]) + in_cell) * o->h;
            }
        } else if (r->shape == SLOSH_EMIT_SPHERE) {
            uint32_t h3 = oracle_pcg_hash(h2);
            float cos_t = 2.0f * rnd[0] - 1.0f;
            float sin_t = sqrtf(clampf(1.0f - cos_t * cos_t, 0.0f, 1.0f));
            float phi = 6.28318530718f * rnd[1];
            float rad = r->extent[0] * powf(fmaxf(oracle_hash_to_unit(h3), 1e-6f), 1.0f / 3.0f);
            float dir[3] = {sin_t * cosf(phi), sin_t * sinf(phi), cos_t};
            for (int a = 0; a < 3; ++a) pos[a] = r->center[a] + rad * dir[a];
        } else {
            float l[3], w[3];
            for (int a = 0; a < 3; ++a) l[a] = (2.0f * rnd[a] - 1.0f) * r->extent[a];
            quat_rotate(r->rotation, l, w);
            for (int a = 0; a < 3; ++a) pos[a] = r->center[a] + w[a];
        }
        float vel[3] = {r->velocity[0], r->velocity[1], r->velocity[2]};
        float speed = len3(vel);
        if (speed > 1e-6f && (r->velocity_spread > 0.0f || r->speed_jitter > 0.0f)) {
            uint32_t h4 = oracle_pcg_hash(h2 ^ 0x9E3779B9u), h5 = oracle_pcg_hash(h4), h6 = oracle_pcg_hash(h5);
            float dir[3] = {vel[0] / speed, vel[1] / speed, vel[2] / speed};
            float up[3] = {fabsf(dir[2]) < 0.9f ? 0.0f : 1.0f, 0.0f, fabsf(dir[2]) < 0.9f ? 1.0f : 0.0f};
            float t0[3], t1[3];
            cross3(dir, up, t0);
            float tl = len3(t0);
            for (int a = 0; a < 3; ++a) t0[a] = t0[a] / tl;
            cross3(dir, t0, t1);
            float cos_a = 1.0f + oracle_hash_to_unit(h4) * (cosf(r->velocity_spread) - 1.0f);
            float sin_a = sqrtf(clampf(1.0f - cos_a * cos_a, 0.0f, 1.0f));
            float phi = 6.28318530718f * oracle_hash_to_unit(h5);
            float cp = cosf(phi), sp = sinf(phi);
            float scale = fmaxf(1.0f + r->speed_jitter * (2.0f * oracle_hash_to_unit(h6) - 1.0f), 0.0f);
            for (int a = 0; a < 3; ++a) vel[a] = (cos_a * dir[a] + sin_a * (cp * t0[a] + sp * t1[a])) * speed * scale;
        }
        slosh_particle *pp = &o->p[o->np + i];
        memset(pp, 0, sizeof(*pp));
        for (int a = 0; a < 3; ++a) {
            pp->pos[a] = clampf(pos[a], lo, hi[a]);
            pp->vel[a] = vel[a];
        }
        pp->age = 0.0f;
        touch_pos(o, pp->pos);
    }
    o->np += (uint32_t)count;
    return SLOSH_OK;
}

int osim_write_particles(osim *o, const slosh_particle *p, uint32_t n)
{
    if (n > o->cap) return SLOSH_ECAPACITY;
    if (n) memcpy(o->p, p, (size_t)n * sizeof(slosh_particle));
    o->np = n;
    memset(o->touch, 0, o->nbricks);
    for (uint32_t i = 0; i < n; ++i)
        if (o->p[i].age >= 0.0f) touch_pos(o, o->p[i].pos);
    return SLOSH_OK;
}

uint32_t osim_num_particles(const osim *o) { return o->np; }

int osim_read_particles(const osim *o, slosh_particle *dst, uint32_t cap)
{
    uint32_t n = o->np < cap ? o->np : cap;
    if (n) memcpy(dst, o->p, (size_t)n * sizeof(slosh_particle));
    return (int)n;
}

/* ------------------------------------------------------------------------------------------------------------ */
/* SloshSort.usf:37-47, 66-154: counting sort by key, restated as a stable sort; key = 8^3 brick of the clamped
 * cell (Slosh.usf:115-118), dead particles (!(age >= 0)) take the last key and are dropped from the live range. */
static void stage_sort(osim *o)
{
    uint32_t nk = o->nbricks + 1u;
    uint32_t *cnt = (uint32_t *)xcalloc(nk + 1u, 4);
    uint32_t *key = (uint32_t *)xcalloc(o->np, 4);
#pragma omp parallel for schedule(static)
    for (uint32_t i = 0; i < o->np; ++i) {
        uint32_t k = o->nbricks;
        if (o->p[i].age >= 0.0f) {
            int c[3];
            pos_cell(o, o->p[i].pos, c);
            k = brick_of_cell(o, c[0], c[1], c[2]);
        }
        key[i] = k;
        __atomic_fetch_add(&cnt[k + 1u], 1u, __ATOMIC_RELAXED);
    }
    for (uint32_t k = 0; k < nk; ++k) cnt[k + 1u] += cnt[k];
    uint32_t live = cnt[o->nbricks];
    for (uint32_t i = 0; i < o->np; ++i) o->tmp[cnt[key[i]]++] = o->p[i];
    slosh_particle *t = o->p;
    o->p = o->tmp;
    o->tmp = t;
    o->np = live;
    free(cnt);
    free(key);
}

/* Slosh.usf:1719-1831 (BrickFreeCS, BrickAllocCS, PoolArgsCS, ClearNewSlotsCS) with deterministic slots. */
static void stage_pool(osim *o)
{
#pragma omp parallel for schedule(static)
    for (uint32_t b = 0; b < o->nbricks; ++b) {
        int c[3];
        brick_coord(o, b, c);
        uint8_t w = 0;
        for (int dz = -1; dz <= 1 && !w; ++dz)
            for (int dy = -1; dy <= 1 && !w; ++dy)
                for (int dx = -1; dx <= 1 && !w; ++dx) {
                    int x = c[0] + dx, y = c[1] + dy, z = c[2] + dz;
                    if (x < 0 || y < 0 || z < 0 || x >= o->bd[0] || y >= o->bd[1] || z >= o->bd[2]) continue;
                    if (o->touch[(uint32_t)(x + o->bd[0] * (y + o->bd[1] * z))]) w = 1;
                }
        o->want[b] = w;
    }
    for (uint32_t b = 0; b < o->nbricks; ++b) {
        uint32_t sl = o->table[b];
        if (!o->want[b] && sl != 0u) {
            o->table[b] = 0u;
            o->slot_brick[sl] = NO_BRICK;
        }
    }
    uint32_t next_free = 1u;
    uint32_t overflow = 0u;
    for (uint32_t b = 0; b < o->nbricks; ++b) {
        if (!o->want[b] || o->table[b] != 0u) continue;
        while (next_free < o->npool && o->slot_brick[next_free] != NO_BRICK) next_free++;
        if (next_free >= o->npool) {
            overflow++;
            continue;
        }
        uint32_t sl = next_free++;
        o->table[b] = sl;
        o->slot_brick[sl] = b;
        size_t base = (size_t)sl * BRICK_CELLS;
        for (uint32_t l = 0; l < BRICK_CELLS; ++l) {
            o->pres[base + l] = 0.0;
            o->ccount[base + l] = 0;
            o->ctype[base + l] = CELL_AIR; /* a new slot has no fluid history (read by the warm start) */
        }
    }
    o->nactive = 0;
    for (uint32_t b = 0; b < o->nbricks; ++b)
        if (o->table[b] != 0u) o->active[o->nactive++] = b;
    memset(o->touch, 0, o->nbricks);
    o->res.active_bricks = o->nactive;
    o->res.pool_overflow_bricks = overflow;
}

/* Iterates the cells of active brick E (Slosh.usf:155-188 SparseCell); FOR_ACTIVE_CELLS walks every allocated cell
 * in active-list order. The parallel loops run FOR_BRICK_CELLS over the active bricks. */
#define FOR_BRICK_CELLS(o, E, IDX, X, Y, Z)                                                                     \
    for (uint32_t l_ = 0; l_ < BRICK_CELLS; ++l_)                                                                 \
        for (int bc_[3], once_ = (brick_coord((o), (o)->active[E], bc_), 1); once_; once_ = 0)                     \
            for (int X = bc_[0] * 8 + (int)(l_ & 7u), Y = bc_[1] * 8 + (int)((l_ >> 3) & 7u),                     \
                     Z = bc_[2] * 8 + (int)(l_ >> 6), ok_ = in_grid((o), X, Y, Z); ok_; ok_ = 0)                   \
                for (uint32_t IDX = (o)->table[(o)->active[E]] * BRICK_CELLS + l_, k_ = 1; k_; k_ = 0)
#define FOR_ACTIVE_CELLS(o, IDX, X, Y, Z)                                                                       \
    for (uint32_t e_ = 0; e_ < (o)->nactive; ++e_) FOR_BRICK_CELLS(o, e_, IDX, X, Y, Z)

/* Slosh.usf:448-491 (ScatterComponentTiled) for one particle and one axis: 8 staggered nodes, int32 fixed point. */
static void p2g_component(osim *o, int axis, const float xp[3], float up, const float cp[3])
{
    float off[3] = {0.5f, 0.5f, 0.5f};
    off[axis] = 0.0f;
    float g[3], fr[3];
    int base[3];
    for (int a = 0; a < 3; ++a) {
        g[a] = xp[a] * o->inv_h - off[a];
        base[a] = (int)floorf(g[a]);
        fr[a] = g[a] - (float)base[a];
    }
    for (uint32_t oi = 0; oi < 8u; ++oi) {
        int od[3] = {(int)(oi & 1u), (int)((oi >> 1) & 1u), (int)((oi >> 2) & 1u)};
        int node[3] = {base[0] + od[0], base[1] + od[1], base[2] + od[2]};
        if (!in_grid(o, node[0], node[1], node[2])) continue;
        float w3[3];
        for (int a = 0; a < 3; ++a) w3[a] = od[a] ? fr[a] : 1.0f - fr[a];
        float w = w3[0] * w3[1] * w3[2];
        if (w <= 0.0f) continue;
        float xn[3], dxn[3];
        for (int a = 0; a < 3; ++a) {
            xn[a] = ((float)node[a] + off[a]) * o->h;
            dxn[a] = xn[a] - xp[a];
        }
        float dot = cp[0] * dxn[0] + cp[1] * dxn[1] + cp[2] * dxn[2];
        int32_t mom = (int32_t)rintf(w * (up + dot) * FIXED_SCALE);
        int32_t wt = (int32_t)rintf(w * FIXED_SCALE);
        uint32_t cell = cell_index(o, node[0], node[1], node[2]);
        if (cell < BRICK_CELLS) continue; /* null brick: dropped */
        size_t f = (size_t)axis * o->pool_cells + cell;
        add_i32(&o->accum[f], mom);
        add_i32(&o->accum[3 * o->pool_cells + f], wt);
    }
}

/* Slosh.usf:365-395 (BuildSolidCS) + 493-563 (P2GCS) + 569-598 (MarkCellsCS). The warm start of fluid cells is
 * the last solution times warm_scale = dt/dt_prev (SloshProxy.cpp:575); with SLOSH_FLAG_WARM_START_EXTRAPOLATE (an
 * option of this build) a cell fluid in the last two solves starts from P = dt (q1 + (q1 - q2) dt/dt_prev), the
 * linear extrapolation of the pressure per unit dt (q1 = P_last/dt_prev, q2 the value before). */
static void stage_p2g(osim *o, float warm_scale, float dt, float dt_prev)
{
#pragma omp parallel for schedule(dynamic, 4)
    for (uint32_t e = 0; e < o->nactive; ++e) FOR_BRICK_CELLS(o, e, idx, x, y, z)
    {
        if (o->had) o->had[idx] = (o->ctype[idx] & CELL_TYPE_MASK) == CELL_FLUID;
        uint32_t packed = CELL_AIR;
        if (x == 0 || y == 0 || z == 0 || x == o->n[0] - 1 || y == o->n[1] - 1 || z == o->n[2] - 1) {
            packed = CELL_SOLID;
        } else {
            float pc[3] = {((float)x + 0.5f) * o->h, ((float)y + 0.5f) * o->h, ((float)z + 0.5f) * o->h};
            float best = 0.0f;
            for (uint32_t k = 0; k < o->ncol; ++k) {
                float dd = oracle_collider_sdf(&o->col[k], pc);
                if (dd < best) {
                    best = dd;
                    packed = CELL_SOLID | ((k + 1u) << CELL_COLLIDER_SHIFT);
                }
            }
        }
        o->ctype[idx] = packed;
        o->ccount[idx] = 0;
    }
    /* Build semantics: a particle deposits only when its own brick has storage (the per-brick gather runs over
     * allocated bricks). The reference also deposits the stencil of a particle in an unallocated brick onto
     * allocated neighbours; the two differ only after a pool overflow. */
#pragma omp parallel for schedule(static, 4096)
    for (uint32_t i = 0; i < o->np; ++i) {
        const slosh_particle *pp = &o->p[i];
        if (!(pp->age >= 0.0f)) continue;
        int c[3];
        pos_cell(o, pp->pos, c);
        uint32_t ci = cell_index(o, c[0], c[1], c[2]);
        if (ci < BRICK_CELLS) continue;
        p2g_component(o, 0, pp->pos, pp->vel[0], &pp->c[0]);
        p2g_component(o, 1, pp->pos, pp->vel[1], &pp->c[3]);
        p2g_component(o, 2, pp->pos, pp->vel[2], &pp->c[6]);
        add_i32(&o->ccount[ci], 1);
    }
    uint32_t fluid = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : fluid)
    for (uint32_t e = 0; e < o->nactive; ++e) FOR_BRICK_CELLS(o, e, idx, x, y, z)
    {
        (void)x; (void)y; (void)z;
        uint32_t packed = o->ctype[idx];
        if ((packed & CELL_TYPE_MASK) != CELL_SOLID) {
            int32_t count = o->ccount[idx];
            packed = count > 0 ? CELL_FLUID : CELL_AIR;
            if ((float)count >= 0.5f * TARGET_PPC) packed |= CELL_DENSE_FLAG;
            o->ctype[idx] = packed;
        }
        if ((packed & CELL_TYPE_MASK) == CELL_FLUID) {
            const double old = o->pres[idx];
            double p0 = old * (double)warm_scale;
            if (o->pres_prev) {
                double q1 = NAN;
                if (o->had[idx] && dt_prev > 0.0f) {
                    q1 = old / (double)dt_prev;
                    const double q2 = o->pres_prev[idx];
                    if (!isnan(q2)) p0 = (double)dt * (q1 + (q1 - q2) * ((double)dt / (double)dt_prev));
                }
                o->pres_prev[idx] = q1;
            }
            o->pres[idx] = p0;
            fluid++;
        } else {
            o->pres[idx] = 0.0;
            if (o->pres_prev) o->pres_prev[idx] = NAN;
        }
    }
    o->res.fluid_cells = fluid;
}

/* Slosh.usf:232-241 */
static float solid_face_velocity(const osim *o, uint32_t packed, int axis, const float fp[3])
{
    uint32_t id = (packed >> CELL_COLLIDER_SHIFT) & 0xFFu;
    if (id == 0u) return 0.0f;
    return collider_velocity_axis(&o->col[id - 1u], fp, axis);
}

/* Slosh.usf:604-651 (NormalizeFace, APIC mode): ratio, weight, gravity on fluid faces, solid velocities. */
static void stage_normalize(osim *o, float dt, const float g[3])
{
    float *vel = o->vel[0];
    o->cur = 0;
#pragma omp parallel for schedule(dynamic, 4)
    for (uint32_t e = 0; e < o->nactive; ++e) FOR_BRICK_CELLS(o, e, idx, x, y, z)
    {
        for (int a = 0; a < 3; ++a) {
            size_t f = (size_t)a * o->pool_cells + idx;
            int32_t mom = o->accum[f];
            int32_t wt = o->accum[3 * o->pool_cells + f];
            o->accum[f] = 0;
            o->accum[3 * o->pool_cells + f] = 0;
            float u = wt > 0 ? (float)mom / (float)wt : 0.0f;
            o->fweight[f] = (float)wt * FIXED_INV;
            uint32_t flags = wt > 0 ? FACE_HAS_WEIGHT : 0u;
            int c[3] = {x, y, z};
            int ca[3] = {x, y, z};
            ca[a] -= 1;
            uint32_t ta = cell_type_at(o, ca[0], ca[1], ca[2]);
            uint32_t tb = cell_type_at(o, c[0], c[1], c[2]);
            uint32_t sa = ta & CELL_TYPE_MASK, sb = tb & CELL_TYPE_MASK;
            float fp[3] = {(float)x + 0.5f, (float)y + 0.5f, (float)z + 0.5f};
            fp[a] = (float)c[a];
            for (int k = 0; k < 3; ++k) fp[k] = fp[k] * o->h;
            if (sa == CELL_SOLID && sb == CELL_SOLID) {
                u = solid_face_velocity(o, tb, a, fp);
            } else if (sa == CELL_SOLID || sb == CELL_SOLID) {
                u = solid_face_velocity(o, sb == CELL_SOLID ? tb : ta, a, fp);
                flags |= FACE_VALID | FACE_SOLID;
            } else if (sa == CELL_FLUID || sb == CELL_FLUID) {
                u += g[a] * dt;
                flags |= FACE_VALID;
            } else {
                u = 0.0f;
            }
            vel[f] = u;
            o->fflags[f] = flags;
        }
    }
}

/* Slosh.usf:657-686 (ViscosityJacobiFace) x viscosity_iterations, driver SloshProxy.cpp:603-617: Jacobi iterations of
 * (I - nu dt Laplacian) u = u* on valid, non-solid faces; u* in vel[0], iterates alternating in vel[1] and vel[2]. */
static void stage_viscosity(osim *o, float alpha, int iterations)
{
    memcpy(o->vel[1], o->vel[0], 3 * o->pool_cells * sizeof(float));
    int src = 1;
    for (int it = 0; it < iterations; ++it) {
        int dst = src == 1 ? 2 : 1;
        const float *us = o->vel[0], *vs = o->vel[src];
        float *vd = o->vel[dst];
#pragma omp parallel for schedule(dynamic, 4)
        for (uint32_t e = 0; e < o->nactive; ++e) FOR_BRICK_CELLS(o, e, idx, x, y, z)
        {
            for (int a = 0; a < 3; ++a) {
                size_t f = (size_t)a * o->pool_cells + idx;
                uint32_t flags = o->fflags[f];
                float center = vs[f];
                if ((flags & FACE_VALID) == 0u || (flags & FACE_SOLID) != 0u) {
                    vd[f] = center;
                    continue;
                }
                float sum = 0.0f;
                for (int k = 0; k < 6; ++k) {
                    int nx = x + NB[k][0], ny = y + NB[k][1], nz = z + NB[k][2];
                    float value = center;
                    if (in_grid(o, nx, ny, nz)) {
                        uint32_t ni = face_index(o, a, nx, ny, nz);
                        if ((o->fflags[ni] & FACE_VALID) != 0u) value = vs[ni];
                    }
                    sum += value;
                }
                vd[f] = (us[f] + alpha * sum) / (1.0f + 6.0f * alpha);
            }
        }
        src = dst;
    }
    o->cur = src;
}

/* Slosh.usf:211-224 */
static float cell_density(const osim *o, int x, int y, int z)
{
    if (!in_grid(o, x, y, z)) return 0.0f;
    float ws = 0.0f;
    for (int a = 0; a < 3; ++a) {
        int p[3] = {x, y, z};
        p[a] += 1;
        ws += o->fweight[face_index(o, a, x, y, z)] + o->fweight[face_index(o, a, p[0], p[1], p[2])];
    }
    return ws / (6.0f * TARGET_PPC);
}

/* Slosh.usf:227-230 */
static float air_face_weight(const osim *o, float rho_fluid, int x, int y, int z)
{
    return o->s.ghost_fluid ? 1.0f / oracle_ghost_theta(rho_fluid, cell_density(o, x, y, z)) : 1.0f;
}

/* Slosh.usf:692-741 (DivergenceCS). */
static void stage_divergence(osim *o, float dt)
{
    const float *vel = o->vel[o->cur];
#pragma omp parallel for schedule(dynamic, 4)
    for (uint32_t e = 0; e < o->nactive; ++e) FOR_BRICK_CELLS(o, e, idx, x, y, z)
    {
        o->ccount[idx] = 0;
        if ((o->ctype[idx] & CELL_TYPE_MASK) != CELL_FLUID) {
            o->div[idx] = 0.0f;
            o->diag[idx] = 0.0f;
            continue;
        }
        float dv = 0.0f;
        for (int a = 0; a < 3; ++a) {
            int p[3] = {x, y, z};
            p[a] += 1;
            dv += vel[face_index(o, a, p[0], p[1], p[2])] - vel[face_index(o, a, x, y, z)];
        }
        float ws = 0.0f;
        for (int a = 0; a < 3; ++a) {
            int p[3] = {x, y, z};
            p[a] += 1;
            ws += o->fweight[face_index(o, a, x, y, z)] + o->fweight[face_index(o, a, p[0], p[1], p[2])];
        }
        float rho = ws / (6.0f * TARGET_PPC);
        float excess = fmaxf(rho - (1.0f + o->s.density_deadband), 0.0f);
        /* the two-sided projection replaces the velocity correction (SloshProxy.cpp:421) */
        const float dcorr = (o->s.flags & SLOSH_FLAG_DENSITY_PROJECTION) ? 0.0f : o->s.density_correction;
        float target = fminf(dcorr * excess, 0.5f) * o->h / dt;
        o->div[idx] = dv - target;
        float dg = 0.0f;
        for (int nb = 0; nb < 6; ++nb) {
            int nx = x + NB[nb][0], ny = y + NB[nb][1], nz = z + NB[nb][2];
            uint32_t t = cell_type_at(o, nx, ny, nz) & CELL_TYPE_MASK;
            dg += t == CELL_FLUID ? 1.0f : (t == CELL_AIR ? air_face_weight(o, rho, nx, ny, nz) : 0.0f);
        }
        o->diag[idx] = dg;
    }
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Multigrid (SloshPressureMG.usf), level accessors. Level 0 is the brick pool, levels >= 1 are dense.           */
static uint32_t lv_index(const osim *o, int l, int x, int y, int z)
{
    if (l == 0) return cell_index(o, x, y, z);
    const mglevel *lv = &o->lv[l];
    return (uint32_t)(x + lv->dim[0] * (y + lv->dim[1] * z));
}

static int lv_inside(const osim *o, int l, int x, int y, int z)
{
    const int *dm = o->lv[l].dim;
    return x >= 0 && y >= 0 && z >= 0 && x < dm[0] && y < dm[1] && z < dm[2];
}

static uint32_t lv_type(const osim *o, int l, int x, int y, int z)
{
    if (!lv_inside(o, l, x, y, z)) return CELL_SOLID;
    if (l == 0) return o->ctype[cell_index(o, x, y, z)] & CELL_TYPE_MASK;
    return o->lv[l].type[lv_index(o, l, x, y, z)] & CELL_TYPE_MASK;
}

/* SloshPressureMG.usf:256-278 (MGBuildTypeCS): AIR if any child is air, else FLUID if any child is fluid, else
 * SOLID. With SLOSH_FLAG_MG_FLUID_COARSENING (a preconditioner option of this build) FLUID takes precedence. */
static void mg_build_types(osim *o)
{
    const int fluid_first = (o->s.flags & SLOSH_FLAG_MG_FLUID_COARSENING) != 0;
    for (int l = 1; l < o->nlevels; ++l) {
        mglevel *lv = &o->lv[l];
#pragma omp parallel for schedule(dynamic, 1)
        for (int z = 0; z < lv->dim[2]; ++z)
            for (int y = 0; y < lv->dim[1]; ++y)
                for (int x = 0; x < lv->dim[0]; ++x) {
                    int any_air = 0, any_fluid = 0;
                    for (uint32_t oi = 0; oi < 8u; ++oi) {
                        int fx = 2 * x + (int)(oi & 1u), fy = 2 * y + (int)((oi >> 1) & 1u), fz = 2 * z + (int)(oi >> 2);
                        if (!lv_inside(o, l - 1, fx, fy, fz)) continue;
                        uint32_t t = lv_type(o, l - 1, fx, fy, fz);
                        any_air |= t == CELL_AIR;
                        any_fluid |= t == CELL_FLUID;
                    }
                    lv->type[lv_index(o, l, x, y, z)] =
                        fluid_first ? (any_fluid ? CELL_FLUID : (any_air ? CELL_AIR : CELL_SOLID))
                                    : (any_air ? CELL_AIR : (any_fluid ? CELL_FLUID : CELL_SOLID));
                }
    }
}

/* Cells of a level, in a fixed order: level 0 through the active list, dense levels x fastest. `end` (when not 0)
 * bounds the walk to one block: an active brick of level 0, or 512 consecutive cells of a dense level. */
typedef struct lv_iter { int l, x, y, z; uint32_t idx; uint32_t e, li, end; } lv_iter;

static int lv_next(const osim *o, lv_iter *it)
{
    if (it->l == 0) {
        for (;;) {
            if (it->e >= o->nactive || (it->end && it->e >= it->end)) return 0;
            uint32_t l = it->li++;
            uint32_t e = it->e;
            if (it->li >= BRICK_CELLS) { it->li = 0; it->e++; }
            int bc[3];
            brick_coord(o, o->active[e], bc);
            it->x = bc[0] * 8 + (int)(l & 7u);
            it->y = bc[1] * 8 + (int)((l >> 3) & 7u);
            it->z = bc[2] * 8 + (int)(l >> 6);
            if (!in_grid(o, it->x, it->y, it->z)) continue;
            it->idx = o->table[o->active[e]] * BRICK_CELLS + l;
            return 1;
        }
    }
    const mglevel *lv = &o->lv[it->l];
    if (it->e >= lv->cells || (it->end && it->e >= it->end)) return 0;
    uint32_t i = it->e++;
    it->idx = i;
    it->x = (int)(i % (uint32_t)lv->dim[0]);
    it->y = (int)((i / (uint32_t)lv->dim[0]) % (uint32_t)lv->dim[1]);
    it->z = (int)(i / (uint32_t)(lv->dim[0] * lv->dim[1]));
    return 1;
}

/* Blocks of level l for the parallel loops, and the iterator of block b. */
static uint32_t lv_nblocks(const osim *o, int l) { return l == 0 ? o->nactive : (o->lv[l].cells + 511u) / 512u; }

static lv_iter lv_block(int l, uint32_t b)
{
    lv_iter it = {l, 0, 0, 0, 0, 0, 0, 0};
    it.e = l == 0 ? b : b * 512u;
    it.end = l == 0 ? b + 1u : b * 512u + 512u;
    return it;
}

/* Sum of the first n block partials of column k (stride 4), in block order. */
static double part_sum(const osim *o, uint32_t n, int k)
{
    double s = 0.0;
    for (uint32_t b = 0; b < n; ++b) s += o->part[4u * b + (uint32_t)k];
    return s;
}

static int is_red(int x, int y, int z) { return (((unsigned)(x + y + z)) & 1u) == 0u; }

/* Coarsest level: FP64 Cholesky of the fluid unknowns, factored once per solve. */
static void coarse_factor(osim *o)
{
    int l = o->nlevels - 1;
    free(o->cidx); free(o->chol);
    o->cidx = NULL; o->chol = NULL;
    uint32_t cells = o->lv[l].cells;
    uint32_t *cx = (uint32_t *)xcalloc(cells, 4);
    int *pos = (int *)xcalloc((size_t)cells * 3u, sizeof(int));
    uint32_t n = 0;
    lv_iter it = {l, 0, 0, 0, 0, 0, 0, 0};
    while (lv_next(o, &it)) {
        if (lv_type(o, l, it.x, it.y, it.z) != CELL_FLUID) continue;
        cx[n] = it.idx;
        pos[3 * n] = it.x; pos[3 * n + 1] = it.y; pos[3 * n + 2] = it.z;
        n++;
    }
    o->cn = n;
    o->cidx = cx;
    double *a = (double *)xcalloc((size_t)n * n, 8);
    for (uint32_t i = 0; i < n; ++i) {
        int x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
        float dg = DIAG_EPS;
        for (int nb = 0; nb < 6; ++nb) {
            int nx = x + NB[nb][0], ny = y + NB[nb][1], nz = z + NB[nb][2];
            if (!lv_inside(o, l, nx, ny, nz)) continue;
            uint32_t t = lv_type(o, l, nx, ny, nz);
            if (t != CELL_SOLID) dg += 1.0f;
            if (t == CELL_FLUID) {
                uint32_t ni = lv_index(o, l, nx, ny, nz);
                for (uint32_t j = 0; j < n; ++j)
                    if (cx[j] == ni) a[(size_t)i * n + j] = -1.0;
            }
        }
        if (l == 0) dg = o->diag[cx[i]] + DIAG_EPS;
        a[(size_t)i * n + i] = (double)dg;
    }
    for (uint32_t j = 0; j < n; ++j) {
        double s = a[(size_t)j * n + j];
        for (uint32_t k = 0; k < j; ++k) s -= a[(size_t)j * n + k] * a[(size_t)j * n + k];
        double djj = sqrt(s > 0.0 ? s : 1e-300);
        a[(size_t)j * n + j] = djj;
        for (uint32_t i = j + 1; i < n; ++i) {
            double t = a[(size_t)i * n + j];
            for (uint32_t k = 0; k < j; ++k) t -= a[(size_t)i * n + k] * a[(size_t)j * n + k];
            a[(size_t)i * n + j] = t / djj;
        }
    }
    o->chol = a;
    free(pos);
}

/* Solves L L^T x = b for the coarse unknowns (vectors indexed by unknown, not by cell). */
static void coarse_solve_d(const osim *o, const double *bvec, double *xvec)
{
    uint32_t n = o->cn;
    const double *a = o->chol;
    for (uint32_t i = 0; i < n; ++i) {
        double s = bvec[i];
        for (uint32_t k = 0; k < i; ++k) s -= a[(size_t)i * n + k] * xvec[k];
        xvec[i] = s / a[(size_t)i * n + i];
    }
    for (uint32_t ii = n; ii-- > 0;) {
        double s = xvec[ii];
        for (uint32_t k = ii + 1; k < n; ++k) s -= a[(size_t)k * n + ii] * xvec[k];
        xvec[ii] = s / a[(size_t)ii * n + ii];
    }
}

/* Instantiate the V-cycle in FP32 (the solver the GPU mirrors) and FP64 (the judge). */
#define REAL float
#define SFX(name) name##_f
#define LV_X(o, l) ((l) == 0 ? (o)->z : (o)->lv[l].x)
#define LV_B(o, l) ((l) == 0 ? (o)->r : (o)->lv[l].b)
#include "oracle_mg.inc"
#undef REAL
#undef SFX
#undef LV_X
#undef LV_B
#define REAL double
#define SFX(name) name##_d
#define LV_X(o, l) ((l) == 0 ? (o)->zd : (o)->lv[l].xd)
#define LV_B(o, l) ((l) == 0 ? (o)->rd : (o)->lv[l].bd)
#include "oracle_mg.inc"
#undef REAL
#undef SFX
#undef LV_X
#undef LV_B

/* Level-0 operator (SloshPressureMG.usf:142-157 with bGhostDiag): returns (diag + eps) and sum over fluid nbrs. */
static double apply_row_d(const osim *o, int x, int y, int z, uint32_t idx, const double *v)
{
    double s = 0.0;
    for (int nb = 0; nb < 6; ++nb) {
        int nx = x + NB[nb][0], ny = y + NB[nb][1], nz = z + NB[nb][2];
        if (lv_type(o, 0, nx, ny, nz) == CELL_FLUID) s += v[cell_index(o, nx, ny, nz)];
    }
    return (double)(o->diag[idx] + DIAG_EPS) * v[idx] - s;
}

/* r = b - A p of the FP64 pressure (FP32 r), the red initialisation z = r / diag of the first V-cycle sweep;
 * returns ||r||^2 and writes ||b||^2. */
static double mgpcg_residual(osim *o, double *bb_out)
{
    const uint32_t nb = lv_nblocks(o, 0);
#pragma omp parallel for schedule(dynamic, 16)
    for (uint32_t blk = 0; blk < nb; ++blk) {
        double bb = 0.0, rr = 0.0;
        lv_iter it = lv_block(0, blk);
        while (lv_next(o, &it)) {
            uint32_t i = it.idx;
            if ((o->ctype[i] & CELL_TYPE_MASK) != CELL_FLUID) {
                o->r[i] = 0.0f; o->z[i] = 0.0f; o->d[i] = 0.0f;
                continue;
            }
            double dgd, sumd;
            stencil_d(o, 0, it.x, it.y, it.z, i, o->pres, &dgd, &sumd);
            float ri = (float)(-(double)o->div[i] - (dgd * o->pres[i] - sumd));
            float dg = o->diag[i] + DIAG_EPS;
            o->r[i] = ri;
            o->z[i] = is_red(it.x, it.y, it.z) ? ri / dg : 0.0f;
            bb += (double)o->div[i] * (double)o->div[i];
            rr += (double)ri * (double)ri;
        }
        o->part[4u * blk] = bb;
        o->part[4u * blk + 1u] = rr;
    }
    *bb_out = part_sum(o, nb, 0);
    return part_sum(o, nb, 1);
}

/* ||b - A p||^2 of the FP64 pressure in FP64 (the true residual that decides convergence). */
static double mgpcg_true_rr(osim *o)
{
    const uint32_t nb = lv_nblocks(o, 0);
#pragma omp parallel for schedule(dynamic, 16)
    for (uint32_t blk = 0; blk < nb; ++blk) {
        double rr = 0.0;
        lv_iter it = lv_block(0, blk);
        while (lv_next(o, &it)) {
            uint32_t i = it.idx;
            if ((o->ctype[i] & CELL_TYPE_MASK) != CELL_FLUID) continue;
            double dgd, sumd;
            stencil_d(o, 0, it.x, it.y, it.z, i, o->pres, &dgd, &sumd);
            double ri = -(double)o->div[i] - (dgd * o->pres[i] - sumd);
            rr += ri * ri;
        }
        o->part[4u * blk] = rr;
    }
    return part_sum(o, nb, 0);
}

/* SloshProxy.cpp:807-1020 (AddPressureSolveMGPCG) with the kernels of SloshPressureMG.usf:284-543, restated with
 * FP64 dot products and alpha/beta updates and a relative-residual stop. The iteration stops when the residual it
 * carries meets the tolerance; the solve is accepted when the FP64 true residual ||b - A p|| / ||b|| of the
 * pressure does too, else it restarts from r = b - A p (at most three times, within max_iterations). */
static void solve_mgpcg_opts(osim *o, double tol, int maxit, int max_restarts)
{
    mg_build_types(o);
    coarse_factor(o);
    double bb = 0.0;
    double rr = mgpcg_residual(o, &bb);
    double bnorm = sqrt(bb);
    o->res.rhs_norm = bnorm;
    int iters = 0;
    double rel = bnorm > 0.0 ? sqrt(rr) / bnorm : 0.0;
    o->res.true_residual_rel = (float)rel;
    if (bnorm == 0.0 || (rel <= tol && sqrt(mgpcg_true_rr(o)) / bnorm <= tol)) {
        o->res.solver_iterations = 0;
        o->res.residual_rel = rel;
        o->res.converged = 1;
        return;
    }
    int conv = 0;
    const uint32_t nb = lv_nblocks(o, 0);
    for (int restart = 0;; ++restart) {
        double rho = vcycle_f(o);
#pragma omp parallel for schedule(dynamic, 16)
        for (uint32_t blk = 0; blk < nb; ++blk) {
            lv_iter it = lv_block(0, blk);
            while (lv_next(o, &it))
                if ((o->ctype[it.idx] & CELL_TYPE_MASK) == CELL_FLUID) o->d[it.idx] = o->z[it.idx];
        }
        int rec = 0;
        while (iters < maxit) {
#pragma omp parallel for schedule(dynamic, 16)
            for (uint32_t blk = 0; blk < nb; ++blk) {
                double dq = 0.0;
                lv_iter it = lv_block(0, blk);
                while (lv_next(o, &it)) {
                    uint32_t i = it.idx;
                    if ((o->ctype[i] & CELL_TYPE_MASK) != CELL_FLUID) continue;
                    /* q = A d with FP64 accumulation of the FP32 d (the terms cancel for smooth d) */
                    double sum = 0.0;
                    for (int k = 0; k < 6; ++k) {
                        int nx = it.x + NB[k][0], ny = it.y + NB[k][1], nz = it.z + NB[k][2];
                        if (lv_type(o, 0, nx, ny, nz) == CELL_FLUID) sum += (double)o->d[cell_index(o, nx, ny, nz)];
                    }
                    float dg = o->diag[i] + DIAG_EPS;
                    double qd = (double)dg * (double)o->d[i] - sum;
                    o->q[i] = (float)qd;
                    dq += (double)o->d[i] * qd;
                }
                o->part[4u * blk] = dq;
            }
            const double dq = part_sum(o, nb, 0);
            double alpha = fabs(dq) > 1e-30 ? rho / dq : 0.0;
#pragma omp parallel for schedule(dynamic, 16)
            for (uint32_t blk = 0; blk < nb; ++blk) {
                double rrb = 0.0;
                lv_iter it = lv_block(0, blk);
                while (lv_next(o, &it)) {
                    uint32_t i = it.idx;
                    if ((o->ctype[i] & CELL_TYPE_MASK) != CELL_FLUID) continue;
                    o->pres[i] = o->pres[i] + alpha * (double)o->d[i];
                    float ri = (float)((double)o->r[i] - alpha * (double)o->q[i]);
                    o->r[i] = ri;
                    rrb += (double)ri * (double)ri;
                    float dg = o->diag[i] + DIAG_EPS;
                    o->z[i] = is_red(it.x, it.y, it.z) ? ri / dg : 0.0f;
                }
                o->part[4u * blk] = rrb;
            }
            rr = part_sum(o, nb, 0);
            iters++;
            rel = sqrt(rr) / bnorm;
            if (rel <= tol) { rec = 1; break; }
            double rho_new = vcycle_f(o);
            double beta = fabs(rho) > 1e-30 ? rho_new / rho : 0.0;
            rho = rho_new;
#pragma omp parallel for schedule(dynamic, 16)
            for (uint32_t blk = 0; blk < nb; ++blk) {
                lv_iter it = lv_block(0, blk);
                while (lv_next(o, &it)) {
                    uint32_t i = it.idx;
                    if ((o->ctype[i] & CELL_TYPE_MASK) != CELL_FLUID) continue;
                    o->d[i] = (float)((double)o->z[i] + beta * (double)o->d[i]);
                }
            }
        }
        const double rel_true = sqrt(mgpcg_true_rr(o)) / bnorm;
        o->res.true_residual_rel = (float)rel_true;
        if (rec && rel_true <= tol) { conv = 1; break; }
        if (iters >= maxit || restart >= max_restarts) break;
        mgpcg_residual(o, &bb);
    }
    o->res.solver_iterations = iters;
    o->res.residual_rel = rel;
    o->res.converged = conv;
}

static void solve_mgpcg(osim *o)
{
    solve_mgpcg_opts(o, (double)o->s.tolerance, o->s.max_iterations > 0 ? o->s.max_iterations : 200, 3);
}

/* One colour sweep of Slosh.usf:856-892 (PressureRBGSCS) on the shared operator, relaxation w. */
static void rbsor_sweep(osim *o, unsigned colour, float w)
{
    const uint32_t nb = lv_nblocks(o, 0);
#pragma omp parallel for schedule(dynamic, 16)
    for (uint32_t blk = 0; blk < nb; ++blk) {
        lv_iter it = lv_block(0, blk);
        while (lv_next(o, &it)) {
            uint32_t i = it.idx;
            if ((o->ctype[i] & CELL_TYPE_MASK) != CELL_FLUID) continue;
            if ((((unsigned)(it.x + it.y + it.z)) & 1u) != colour) continue;
            double sum = 0.0;
            for (int q = 0; q < 6; ++q) {
                int nx = it.x + NB[q][0], ny = it.y + NB[q][1], nz = it.z + NB[q][2];
                if (lv_type(o, 0, nx, ny, nz) == CELL_FLUID) sum += o->pres[cell_index(o, nx, ny, nz)];
            }
            double dg = (double)(o->diag[i] + DIAG_EPS);
            double gs = (sum - (double)o->div[i]) / dg;
            double old = o->pres[i];
            o->pres[i] = old + (double)w * (gs - old);
        }
    }
}

/* Slosh.usf:856-892 (PressureRBGSCS) on the shared operator; residual checked every 8 iterations in FP64. */
static void solve_rbsor(osim *o)
{
    double tol = (double)o->s.tolerance;
    int maxit = o->s.max_iterations > 0 ? o->s.max_iterations : 1000;
    float w = o->s.sor_weight;
    const uint32_t nb = lv_nblocks(o, 0);
#pragma omp parallel for schedule(dynamic, 16)
    for (uint32_t blk = 0; blk < nb; ++blk) {
        double bbb = 0.0;
        lv_iter it = lv_block(0, blk);
        while (lv_next(o, &it))
            if ((o->ctype[it.idx] & CELL_TYPE_MASK) == CELL_FLUID) bbb += (double)o->div[it.idx] * (double)o->div[it.idx];
        o->part[4u * blk] = bbb;
    }
    double bnorm = sqrt(part_sum(o, nb, 0));
    o->res.rhs_norm = bnorm;
    int iters = 0, conv = 0;
    double rel = 0.0;
    for (;;) {
#pragma omp parallel for schedule(dynamic, 16)
        for (uint32_t blk = 0; blk < nb; ++blk) {
            double rrb = 0.0;
            lv_iter it = lv_block(0, blk);
            while (lv_next(o, &it)) {
                uint32_t i = it.idx;
                if ((o->ctype[i] & CELL_TYPE_MASK) != CELL_FLUID) continue;
                double dg, sum;
                stencil_d(o, 0, it.x, it.y, it.z, i, o->pres, &dg, &sum);
                double ri = -(double)o->div[i] - (dg * o->pres[i] - sum);
                rrb += ri * ri;
            }
            o->part[4u * blk] = rrb;
        }
        double rr = part_sum(o, nb, 0);
        rel = bnorm > 0.0 ? sqrt(rr) / bnorm : 0.0;
        if (rel <= tol) { conv = 1; break; }
        if (iters >= maxit) break;
        for (int k = 0; k < 8 && iters < maxit; ++k, ++iters)
            for (unsigned colour = 0; colour < 2u; ++colour) rbsor_sweep(o, colour, w);
    }
    o->res.solver_iterations = iters;
    o->res.residual_rel = rel;
    o->res.converged = conv;
    o->res.true_residual_rel = (float)rel;
}

/* Slosh.usf:995-1052 (ProjectFace) without the collider-coupling accumulations. */
static void stage_project(osim *o)
{
    float *vel = o->vel[o->cur];
#pragma omp parallel for schedule(dynamic, 4)
    for (uint32_t e = 0; e < o->nactive; ++e) FOR_BRICK_CELLS(o, e, idx, x, y, z)
    {
        for (int a = 0; a < 3; ++a) {
            size_t f = (size_t)a * o->pool_cells + idx;
            int ca[3] = {x, y, z};
            ca[a] -= 1;
            uint32_t ta = cell_type_at(o, ca[0], ca[1], ca[2]) & CELL_TYPE_MASK;
            uint32_t tb = cell_type_at(o, x, y, z) & CELL_TYPE_MASK;
            uint32_t flags = o->fflags[f] & FACE_HAS_WEIGHT;
            if (ta == CELL_SOLID && tb == CELL_SOLID) {
            } else if (ta == CELL_SOLID || tb == CELL_SOLID) {
                flags |= FACE_VALID | FACE_SOLID;
            } else if (ta == CELL_FLUID || tb == CELL_FLUID) {
                double pa, pb;
                if (ta == CELL_FLUID && tb == CELL_FLUID) {
                    pa = o->pres[cell_index(o, ca[0], ca[1], ca[2])];
                    pb = o->pres[idx];
                } else if (ta == CELL_FLUID) {
                    pa = o->pres[cell_index(o, ca[0], ca[1], ca[2])];
                    pb = pa - pa * (double)air_face_weight(o, cell_density(o, ca[0], ca[1], ca[2]), x, y, z);
                } else {
                    pb = o->pres[idx];
                    pa = pb - pb * (double)air_face_weight(o, cell_density(o, x, y, z), ca[0], ca[1], ca[2]);
                }
                vel[f] -= (float)(pb - pa);
                flags |= FACE_VALID;
            }
            o->fflags[f] = flags;
        }
    }
    /* Slosh.usf:1222-1240 (ResidualCS): post-projection |div|/h over fluid cells. */
    float mx = 0.0f;
    uint32_t cnt = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(max : mx) reduction(+ : cnt)
    for (uint32_t e = 0; e < o->nactive; ++e) {
        double bsum = 0.0;
        FOR_BRICK_CELLS(o, e, idx, x, y, z)
        {
            if ((o->ctype[idx] & CELL_TYPE_MASK) != CELL_FLUID) continue;
            float dv = 0.0f;
            for (int a = 0; a < 3; ++a) {
                int p[3] = {x, y, z};
                p[a] += 1;
                dv += vel[face_index(o, a, p[0], p[1], p[2])] - vel[face_index(o, a, x, y, z)];
            }
            float rate = fabsf(dv) * o->inv_h;
            if (rate > mx) mx = rate;
            bsum += (double)rate;
            cnt++;
        }
        o->part[4u * e] = bsum;
    }
    const double sum = part_sum(o, o->nactive, 0);
    o->res.max_divergence = mx;
    o->res.mean_divergence = cnt ? (float)(sum / (double)cnt) : 0.0f;
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Collider coupling (Slosh.usf:909-1216, 1316-1370; driver SloshProxy.cpp:397, 562, 645-649, 715-718). Sums per
 * collider in 1/1024 units, integer adds, so the parallel order does not matter. Slots: 0-2 force, 3-5 torque,
 * 6 wetted faces, 7 submerged cells, 8-10 fluid velocity, 11 solid cells, 12-14 submerged offset.
 * The wetted faces' fluid velocity is read after the whole projection (the reference reads it while other faces
 * are still being projected). */
#define COUPLING_SCALE 1024.0f
static uint32_t collider_id(uint32_t packed) { return (packed >> CELL_COLLIDER_SHIFT) & 0xFFu; }
static void add_fixed(int32_t *p, float v) { add_i32(p, (int32_t)rintf(v * COUPLING_SCALE)); }

/* Slosh.usf:913-926 */
static void pressure_force(osim *o, uint32_t k, const float fp[3], const float into[3], float p_hat, float p2pa)
{
    const float pa = p_hat * p2pa;
    float f[3], r[3], t[3];
    for (int a = 0; a < 3; ++a) {
        f[a] = pa * o->h * o->h * into[a];
        r[a] = fp[a] - o->col[k].center[a];
    }
    cross3(r, f, t);
    for (int a = 0; a < 3; ++a) add_fixed(&o->coupling[k * 16u + (uint32_t)a], f[a]);
    for (int a = 0; a < 3; ++a) add_fixed(&o->coupling[k * 16u + 3u + (uint32_t)a], t[a]);
}

/* Slosh.usf:951-983: walks of up to 16 cells along +x, -x, +y, -y out of collider id's solid region. */
static int seepage_pressure(const osim *o, int x, int y, int z, uint32_t id, float *p_out)
{
    static const int dirs[4][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}};
    float sum = 0.0f, cnt = 0.0f;
    for (int d = 0; d < 4; ++d) {
        int px = x, py = y;
        for (int s = 0; s < 16; ++s) {
            px += dirs[d][0];
            py += dirs[d][1];
            const uint32_t t = cell_type_at(o, px, py, z);
            if ((t & CELL_TYPE_MASK) != CELL_SOLID || collider_id(t) != id) {
                if ((t & CELL_TYPE_MASK) == CELL_FLUID) {
                    sum += (float)o->pres[cell_index(o, px, py, z)];
                    cnt += 1.0f;
                }
                break;
            }
        }
    }
    *p_out = cnt > 0.0f ? sum / cnt : 0.0f;
    return cnt > 0.0f;
}

/* Slosh.usf:985-998 */
static void contact(osim *o, uint32_t packed, int x, int y, int z, const float fp[3], const float into[3], float p2pa)
{
    const uint32_t id = collider_id(packed);
    float p;
    if (id == 0u || !(o->col[id - 1u].flags & SLOSH_COLLIDER_TWO_WAY)) return;
    if (seepage_pressure(o, x, y, z, id, &p)) pressure_force(o, id - 1u, fp, into, p, p2pa);
}

/* Slosh.usf:1003-1028 (the coupling terms of ProjectFace) */
static void stage_coupling_faces(osim *o, float p2pa, int binary)
{
    const float *vel = o->vel[o->cur];
#pragma omp parallel for schedule(dynamic, 4)
    for (uint32_t e = 0; e < o->nactive; ++e) FOR_BRICK_CELLS(o, e, idx, x, y, z)
    {
        const uint32_t B = o->ctype[idx], tb = B & CELL_TYPE_MASK;
        for (int a = 0; a < 3; ++a) {
            int ca[3] = {x, y, z};
            ca[a] -= 1;
            const uint32_t A = cell_type_at(o, ca[0], ca[1], ca[2]), ta = A & CELL_TYPE_MASK;
            /* FacePosition (Slosh.usf:204-208) */
            float fp[3] = {((float)x + 0.5f) * o->h, ((float)y + 0.5f) * o->h, ((float)z + 0.5f) * o->h};
            const int cc[3] = {x, y, z};
            fp[a] = (float)cc[a] * o->h;
            float axis[3] = {0.0f, 0.0f, 0.0f}, naxis[3] = {0.0f, 0.0f, 0.0f};
            axis[a] = 1.0f;
            naxis[a] = -1.0f;
            if (ta == CELL_SOLID && tb == CELL_SOLID) {
                if (collider_id(A) != collider_id(B)) {
                    contact(o, B, x, y, z, fp, axis, p2pa);
                    contact(o, A, ca[0], ca[1], ca[2], fp, naxis, p2pa);
                }
            } else if (ta == CELL_SOLID || tb == CELL_SOLID) {
                uint32_t id = 0u;
                int fc[3] = {x, y, z};
                const float *into = axis;
                if (ta == CELL_FLUID && collider_id(B) != 0u) {
                    id = collider_id(B);
                    fc[0] = ca[0];
                    fc[1] = ca[1];
                    fc[2] = ca[2];
                } else if (tb == CELL_FLUID && collider_id(A) != 0u) {
                    id = collider_id(A);
                    into = naxis;
                }
                if (id == 0u || !(o->col[id - 1u].flags & SLOSH_COLLIDER_TWO_WAY)) continue;
                const uint32_t k = id - 1u, fi = cell_index(o, fc[0], fc[1], fc[2]);
                if (binary) pressure_force(o, k, fp, into, (float)o->pres[fi], p2pa);
                add_i32(&o->coupling[k * 16u + 6u], 1);
                /* Slosh.usf:898-907 (CellCenterVelocity) */
                for (int b = 0; b < 3; ++b) {
                    int up[3] = {fc[0], fc[1], fc[2]};
                    up[b] += 1;
                    const float v = 0.5f * (vel[face_index(o, b, fc[0], fc[1], fc[2])] +
                                            vel[face_index(o, b, up[0], up[1], up[2])]);
                    add_fixed(&o->coupling[k * 16u + 8u + (uint32_t)b], v);
                }
            }
        }
    }
}

static uint32_t type_of(const osim *o, int x, int y, int z) { return cell_type_at(o, x, y, z) & CELL_TYPE_MASK; }

/* Slosh.usf:1138-1216 (CouplingForceCS with FluidPressureGradient 1094-1111 and BodySurfacePressure 1115-1136) */
static void stage_coupling_force(osim *o, float p2pa)
{
    const float h = o->h;
#pragma omp parallel for schedule(dynamic, 4)
    for (uint32_t e = 0; e < o->nactive; ++e) FOR_BRICK_CELLS(o, e, idx, x, y, z)
    {
        const uint32_t packed = o->ctype[idx], type = packed & CELL_TYPE_MASK;
        if (type == CELL_AIR) continue;
        const float X[3] = {((float)x + 0.5f) * h, ((float)y + 0.5f) * h, ((float)z + 0.5f) * h};
        const int C[3] = {x, y, z};
        float ph = 0.0f, gr[3] = {0.0f, 0.0f, 0.0f};
        int have = 0;
        if (type == CELL_FLUID) {
            ph = (float)o->pres[idx];
            for (int a = 0; a < 3; ++a) {
                int p[3] = {x, y, z}, m[3] = {x, y, z};
                p[a] += 1;
                m[a] -= 1;
                const int bp = type_of(o, p[0], p[1], p[2]) == CELL_FLUID, bm = type_of(o, m[0], m[1], m[2]) == CELL_FLUID;
                const float pp = bp ? (float)o->pres[cell_index(o, p[0], p[1], p[2])] : 0.0f;
                const float pm = bm ? (float)o->pres[cell_index(o, m[0], m[1], m[2])] : 0.0f;
                gr[a] = bp && bm ? (pp - pm) * 0.5f : (bp ? pp - ph : (bm ? ph - pm : 0.0f));
            }
            for (int a = 0; a < 3; ++a) gr[a] = gr[a] * o->inv_h;
            have = 1;
        } else if (collider_id(packed) != 0u) {
            float sum = 0.0f, cnt = 0.0f;
            for (int nb = 0; nb < 6; ++nb) {
                int d[3] = {0, 0, 0};
                d[nb >> 1] = (nb & 1) ? 1 : -1;
                const int n1[3] = {C[0] + d[0], C[1] + d[1], C[2] + d[2]};
                if (type_of(o, n1[0], n1[1], n1[2]) != CELL_FLUID) continue;
                const float pn = (float)o->pres[cell_index(o, n1[0], n1[1], n1[2])];
                const int n2[3] = {n1[0] + d[0], n1[1] + d[1], n1[2] + d[2]};
                const float slope =
                    type_of(o, n2[0], n2[1], n2[2]) == CELL_FLUID ? (float)o->pres[cell_index(o, n2[0], n2[1], n2[2])] - pn : 0.0f;
                sum += pn - 0.5f * slope;
                cnt += 1.0f;
            }
            ph = cnt > 0.0f ? sum / cnt : 0.0f;
            have = cnt > 0.0f;
        }
        if (!have) continue;
        for (uint32_t k = 0; k < o->ncol; ++k) {
            const slosh_collider *cc = &o->col[k];
            if (!(cc->flags & SLOSH_COLLIDER_TWO_WAY) || fabsf(oracle_collider_sdf(cc, X)) > h) continue;
            if (type == CELL_SOLID && collider_id(packed) != k + 1u) continue;
            float sp = ph;
            if (type == CELL_FLUID) {
                const float phi = oracle_collider_sdf(cc, X);
                float n[3];
                collider_normal(cc, X, 0.25f * h, n);
                sp += gr[0] * (-phi * n[0]) + gr[1] * (-phi * n[1]) + gr[2] * (-phi * n[2]);
            }
            const float scale = -sp * p2pa * h * h;
            float F[3] = {0.0f, 0.0f, 0.0f}, T[3] = {0.0f, 0.0f, 0.0f};
            for (int nb = 0; nb < 6; ++nb) {
                float nrm[3] = {0.0f, 0.0f, 0.0f};
                nrm[nb >> 1] = (nb & 1) ? 1.0f : -1.0f;
                if (type == CELL_SOLID) {
                    int d[3] = {0, 0, 0};
                    d[nb >> 1] = (nb & 1) ? 1 : -1;
                    const uint32_t nbp = cell_type_at(o, x + d[0], y + d[1], z + d[2]);
                    if ((nbp & CELL_TYPE_MASK) == CELL_SOLID && collider_id(nbp) != k + 1u) continue;
                }
                float fpos[3], df[3], r[3], t[3];
                for (int a = 0; a < 3; ++a) fpos[a] = X[a] + 0.5f * h * nrm[a];
                const float w = clampf(0.5f + oracle_collider_sdf(cc, fpos) * o->inv_h, 0.0f, 1.0f);
                for (int a = 0; a < 3; ++a) {
                    df[a] = scale * w * nrm[a];
                    r[a] = fpos[a] - cc->center[a];
                }
                cross3(r, df, t);
                for (int a = 0; a < 3; ++a) {
                    F[a] += df[a];
                    T[a] += t[a];
                }
            }
            for (int a = 0; a < 3; ++a) add_fixed(&o->coupling[k * 16u + (uint32_t)a], F[a]);
            for (int a = 0; a < 3; ++a) add_fixed(&o->coupling[k * 16u + 3u + (uint32_t)a], T[a]);
        }
    }
}

/* Slosh.usf:1316-1370 (SubmergedCS): walks of up to 12 cells along +x, -x, +y, -y; the exit cell must be dense fluid
 * with no air in the three cells below it along -z. */
static void stage_submerged(osim *o)
{
    static const int dirs[4][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}};
#pragma omp parallel for schedule(dynamic, 4)
    for (uint32_t e = 0; e < o->nactive; ++e) FOR_BRICK_CELLS(o, e, idx, x, y, z)
    {
        const uint32_t packed = o->ctype[idx], id = collider_id(packed);
        if ((packed & CELL_TYPE_MASK) != CELL_SOLID || id == 0u) continue;
        int32_t *acc = &o->coupling[(id - 1u) * 16u];
        add_i32(&acc[11], 1);
        int sub = 0;
        for (int d = 0; d < 4; ++d) {
            int px = x, py = y;
            for (int s = 0; s < 12; ++s) {
                px += dirs[d][0];
                py += dirs[d][1];
                const uint32_t t = cell_type_at(o, px, py, z);
                if ((t & CELL_TYPE_MASK) != CELL_SOLID || collider_id(t) != id) {
                    int liquid = (t & CELL_TYPE_MASK) == CELL_FLUID && (t & CELL_DENSE_FLAG) != 0u;
                    for (int below = 1; below <= 3; ++below)
                        liquid = liquid && type_of(o, px, py, z - below) != CELL_AIR;
                    sub = sub || liquid;
                    break;
                }
            }
        }
        if (!sub) continue;
        add_i32(&acc[7], 1);
        const slosh_collider *cc = &o->col[id - 1u];
        const float off[3] = {((float)x + 0.5f) * o->h - cc->center[0], ((float)y + 0.5f) * o->h - cc->center[1],
                              ((float)z + 0.5f) * o->h - cc->center[2]};
        for (int a = 0; a < 3; ++a) add_fixed(&acc[12 + a], off[a]);
    }
}

static void stage_coupling(osim *o, float dt)
{
    memset(o->coupling, 0, sizeof(o->coupling));
    o->coupling_ncol = 0;
    if (!o->ncol) return;
    /* SloshProxy.cpp:562: PressureToPascal = density h / dt */
    const float rho = o->s.fluid_density > 0.0f ? o->s.fluid_density : 1000.0f;
    const float p2pa = rho * o->h / dt;
    const int binary = (o->s.flags & SLOSH_FLAG_BINARY_COUPLING) != 0u;
    stage_coupling_faces(o, p2pa, binary);
    if (!binary) stage_coupling_force(o, p2pa);
    stage_submerged(o);
    o->coupling_ncol = o->ncol;
}

int osim_coupling_raw(const osim *o, int32_t *dst, uint32_t cap)
{
    const uint32_t m = o->coupling_ncol < cap ? o->coupling_ncol : cap;
    memcpy(dst, o->coupling, (size_t)m * 16u * sizeof(int32_t));
    return (int)o->coupling_ncol;
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Two-sided density projection (Slosh.usf:743-854; driver SloshProxy.cpp:656-702).                             */
/* Slosh.usf:753-773 (CellDensityInterior): P2G weight of the faces no solid shares, per rest count of those faces. */
static float cell_density_interior(const osim *o, int x, int y, int z)
{
    float ws = 0.0f, faces = 0.0f;
    for (int a = 0; a < 3; ++a) {
        int m[3] = {x, y, z}, p[3] = {x, y, z};
        m[a] -= 1;
        p[a] += 1;
        if ((cell_type_at(o, m[0], m[1], m[2]) & CELL_TYPE_MASK) != CELL_SOLID) {
            ws += o->fweight[face_index(o, a, x, y, z)];
            faces += 1.0f;
        }
        if ((cell_type_at(o, p[0], p[1], p[2]) & CELL_TYPE_MASK) != CELL_SOLID) {
            ws += o->fweight[face_index(o, a, p[0], p[1], p[2])];
            faces += 1.0f;
        }
    }
    return faces > 0.0f ? ws / (faces * TARGET_PPC) : 1.0f;
}

/* Slosh.usf:775-805 (DensityProjectionRhsCS): the solver's "divergence" -relax * clamp(rho - 1, -0.5, 0.5) of fluid
 * cells, a cell touching air only expanding; q from 0 (kept when the pressure solver is off: q as written). */
static void stage_density_rhs(osim *o, float relax, int keep_q)
{
#pragma omp parallel for schedule(dynamic, 4)
    for (uint32_t e = 0; e < o->nactive; ++e) FOR_BRICK_CELLS(o, e, idx, x, y, z)
    {
        if (!keep_q) o->dq[idx] = 0.0;
        if ((o->ctype[idx] & CELL_TYPE_MASK) != CELL_FLUID) {
            o->drhs[idx] = 0.0f;
            continue;
        }
        float dev = cell_density_interior(o, x, y, z) - 1.0f;
        int air = 0;
        for (int k = 0; k < 6; ++k)
            air = air || (cell_type_at(o, x + NB[k][0], y + NB[k][1], z + NB[k][2]) & CELL_TYPE_MASK) == CELL_AIR;
        if (air) dev = fmaxf(dev, 0.0f);
        o->drhs[idx] = -relax * clampf(dev, -0.5f, 0.5f);
    }
}

/* Slosh.usf:808-828 (DensityFaceDisplacement): -(q_b - q_a) per face (q of a non-fluid cell 0), 0 at solids. */
static void stage_density_disp(osim *o)
{
#pragma omp parallel for schedule(dynamic, 4)
    for (uint32_t e = 0; e < o->nactive; ++e) FOR_BRICK_CELLS(o, e, idx, x, y, z)
    {
        uint32_t tb = o->ctype[idx] & CELL_TYPE_MASK;
        for (int a = 0; a < 3; ++a) {
            size_t f = (size_t)a * o->pool_cells + idx;
            int ca[3] = {x, y, z};
            ca[a] -= 1;
            uint32_t ta = cell_type_at(o, ca[0], ca[1], ca[2]) & CELL_TYPE_MASK;
            if (ta == CELL_SOLID || tb == CELL_SOLID) {
                o->ddisp[f] = 0.0f;
                continue;
            }
            double qa = ta == CELL_FLUID ? o->dq[cell_index(o, ca[0], ca[1], ca[2])] : 0.0;
            double qb = tb == CELL_FLUID ? o->dq[idx] : 0.0;
            o->ddisp[f] = (float)(-(qb - qa));
        }
    }
}

/* SloshProxy.cpp:656-683: the projection's solve with the pressure solver on the same operator, Divergence := rhs,
 * Pressure := q: two MGPCG iterations (MGPCG and flexible CG), or density_projection_iterations RB-SOR sweeps at
 * omega 1.7; no stopping test. The step's pressure-solve results are kept. */
static void solve_density(osim *o)
{
    slosh_results keep = o->res;
    float *div = o->div;
    double *pres = o->pres;
    o->div = o->drhs;
    o->pres = o->dq;
    if (o->s.pressure_solver == SLOSH_SOLVER_RBSOR) {
        int iters = o->s.density_projection_iterations > 0 ? o->s.density_projection_iterations : 1;
        for (int it = 0; it < iters; ++it)
            for (unsigned colour = 0; colour < 2u; ++colour) rbsor_sweep(o, colour, 1.7f);
    } else {
        solve_mgpcg_opts(o, 0.0, 2, 0);
    }
    o->div = div;
    o->pres = pres;
    o->res = keep;
}

/* Slosh.usf:830-853 (SampleDensityDisplacement): face displacements interpolated per axis at xp (stencil nodes
 * clamped into the grid), shortened to half a cell, in metres. */
static void sample_displacement(const osim *o, const float xp[3], float out[3])
{
    float dv[3];
    for (int axis = 0; axis < 3; ++axis) {
        float off[3] = {0.5f, 0.5f, 0.5f};
        off[axis] = 0.0f;
        int base[3];
        float fr[3];
        for (int a = 0; a < 3; ++a) {
            float g = xp[a] * o->inv_h - off[a];
            base[a] = (int)floorf(g);
            fr[a] = g - (float)base[a];
        }
        float value = 0.0f;
        for (uint32_t oi = 0; oi < 8u; ++oi) {
            int od[3] = {(int)(oi & 1u), (int)((oi >> 1) & 1u), (int)((oi >> 2) & 1u)};
            int node[3];
            float w3[3];
            for (int a = 0; a < 3; ++a) {
                node[a] = clampi(base[a] + od[a], 0, o->n[a] - 1);
                w3[a] = od[a] ? fr[a] : 1.0f - fr[a];
            }
            value += w3[0] * w3[1] * w3[2] * o->ddisp[face_index(o, axis, node[0], node[1], node[2])];
        }
        dv[axis] = value;
    }
    float len = sqrtf(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
    float s = len > 0.5f ? 0.5f / len : 1.0f;
    for (int a = 0; a < 3; ++a) out[a] = (len > 0.5f ? dv[a] * s : dv[a]) * o->h;
}

/* Slosh.usf:1263-1310 (ExtrapolateFace), layer `layer`: src = vel[cur], dst = next buffer. */
static void stage_extrapolate(osim *o, uint32_t layer)
{
    const float *src = o->vel[o->cur];
    int dsti = (o->cur + 1) % 3;
    float *dst = o->vel[dsti];
    uint32_t known = FACE_VALID;
    for (uint32_t k = 0; k < layer; ++k) known |= 1u << (FACE_EXTRAP_SHIFT + k);
    const uint32_t bit = 1u << (FACE_EXTRAP_SHIFT + layer);
    /* faces filled by this layer get their bit after the pass: the pass reads only the bits of earlier layers */
#pragma omp parallel for schedule(dynamic, 4)
    for (uint32_t e = 0; e < o->nactive; ++e) FOR_BRICK_CELLS(o, e, idx, x, y, z)
    {
        for (int a = 0; a < 3; ++a) {
            size_t f = (size_t)a * o->pool_cells + idx;
            uint32_t flags = o->fflags[f];
            float center = src[f];
            if (flags & known) {
                dst[f] = center;
                continue;
            }
            float sum = 0.0f, cnt = 0.0f;
            for (int nb = 0; nb < 6; ++nb) {
                int nx = x + NB[nb][0], ny = y + NB[nb][1], nz = z + NB[nb][2];
                if (!in_grid(o, nx, ny, nz)) continue;
                uint32_t ni = face_index(o, a, nx, ny, nz);
                uint32_t nf = o->fflags[ni];
                if ((nf & known) != 0u && (nf & FACE_SOLID) == 0u) {
                    sum += src[ni];
                    cnt += 1.0f;
                }
            }
            if (cnt > 0.0f) {
                dst[f] = sum / cnt;
                o->xmark[f] = 1u;
            } else {
                dst[f] = 0.0f;
            }
        }
    }
#pragma omp parallel for schedule(dynamic, 4)
    for (uint32_t e = 0; e < o->nactive; ++e) FOR_BRICK_CELLS(o, e, idx, x, y, z)
    {
        (void)x; (void)y; (void)z;
        for (int a = 0; a < 3; ++a) {
            size_t f = (size_t)a * o->pool_cells + idx;
            if (o->xmark[f]) {
                o->fflags[f] |= bit;
                o->xmark[f] = 0u;
            }
        }
    }
    o->cur = dsti;
}

/* Slosh.usf:1376-1410 (GatherComponent, APIC branch). */
static void gather_component(const osim *o, const float *vel, int axis, const float xp[3], float *unew, float grad[3])
{
    float off[3] = {0.5f, 0.5f, 0.5f};
    off[axis] = 0.0f;
    float g[3], fr[3];
    int base[3];
    for (int a = 0; a < 3; ++a) {
        g[a] = xp[a] * o->inv_h - off[a];
        base[a] = (int)floorf(g[a]);
        fr[a] = g[a] - (float)base[a];
    }
    float u_acc = 0.0f, gr[3] = {0.0f, 0.0f, 0.0f};
    for (uint32_t oi = 0; oi < 8u; ++oi) {
        int od[3] = {(int)(oi & 1u), (int)((oi >> 1) & 1u), (int)((oi >> 2) & 1u)};
        int node[3];
        float w3[3], sg[3];
        for (int a = 0; a < 3; ++a) {
            node[a] = clampi(base[a] + od[a], 0, o->n[a] - 1);
            w3[a] = od[a] ? fr[a] : 1.0f - fr[a];
            sg[a] = od[a] ? 1.0f : -1.0f;
        }
        float w = w3[0] * w3[1] * w3[2];
        float gw[3] = {sg[0] * (w3[1] * w3[2]) * o->inv_h, sg[1] * (w3[0] * w3[2]) * o->inv_h,
                       sg[2] * (w3[0] * w3[1]) * o->inv_h};
        float u = vel[face_index(o, axis, node[0], node[1], node[2])];
        u_acc += w * u;
        for (int a = 0; a < 3; ++a) gr[a] += gw[a] * u;
    }
    *unew = u_acc;
    grad[0] = gr[0]; grad[1] = gr[1]; grad[2] = gr[2];
}

/* Slosh.usf:1417-1443 (SampleVelocityComponent). */
static float sample_component(const osim *o, const float *vel, int axis, const float xp[3], float fallback, int check)
{
    const uint32_t known = FACE_VALID | (0xFFFFFFFFu << FACE_EXTRAP_SHIFT);
    float off[3] = {0.5f, 0.5f, 0.5f};
    off[axis] = 0.0f;
    float g[3], fr[3];
    int base[3];
    for (int a = 0; a < 3; ++a) {
        g[a] = xp[a] * o->inv_h - off[a];
        base[a] = (int)floorf(g[a]);
        fr[a] = g[a] - (float)base[a];
    }
    float u = 0.0f;
    for (uint32_t oi = 0; oi < 8u; ++oi) {
        int od[3] = {(int)(oi & 1u), (int)((oi >> 1) & 1u), (int)((oi >> 2) & 1u)};
        int node[3];
        float w3[3];
        for (int a = 0; a < 3; ++a) {
            node[a] = clampi(base[a] + od[a], 0, o->n[a] - 1);
            w3[a] = od[a] ? fr[a] : 1.0f - fr[a];
        }
        uint32_t f = face_index(o, axis, node[0], node[1], node[2]);
        float uf = vel[f];
        if (check && (o->fflags[f] & known) == 0u) uf = fallback;
        u += w3[0] * w3[1] * w3[2] * uf;
    }
    return u;
}

/* Slosh.usf:1451-1662 (AdvectParticle, APIC; manipulators, adhesion and drains off) + 1664-1678 (touch). */
static void stage_g2p(osim *o, float dt, const float grav[3])
{
    const float *vel = o->vel[o->cur];
    float lo = o->h * 1.01f;
    float hi[3];
    for (int a = 0; a < 3; ++a) hi[a] = ((float)o->n[a] - 1.01f) * o->h;
    float friction = o->s.boundary_friction;
    float radius = o->s.particle_radius_cells * o->h;
    float maxspeed = 0.0f;
    uint32_t clamped = 0, drained = 0;
    const uint32_t frame_seed = (uint32_t)(o->frame * 16u);
#pragma omp parallel for schedule(static, 4096) reduction(max : maxspeed) reduction(+ : clamped, drained)
    for (uint32_t i = 0; i < o->np; ++i) {
        slosh_particle *pp = &o->p[i];
        if (!(pp->age >= 0.0f)) continue;
        float pic[3], gx[3], gy[3], gz[3];
        gather_component(o, vel, 0, pp->pos, &pic[0], gx);
        gather_component(o, vel, 1, pp->pos, &pic[1], gy);
        gather_component(o, vel, 2, pp->pos, &pic[2], gz);
        float nv[3] = {pic[0], pic[1], pic[2]};
        float cm[9] = {gx[0], gx[1], gx[2], gy[0], gy[1], gy[2], gz[0], gz[1], gz[2]};
        float v1[3] = {pic[0], pic[1], pic[2]};
        float hdt = 0.5f * dt;
        float mid[3];
        for (int a = 0; a < 3; ++a) mid[a] = clampf(pp->pos[a] + hdt * v1[a], lo, hi[a]);
        float midcells = hdt * len3(v1) * o->inv_h;
        float v2[3] = {v1[0], v1[1], v1[2]};
        if (midcells < 2.0f) {
            int check = midcells > 1.0f;
            float t = clampf(midcells - 1.0f, 0.0f, 1.0f);
            for (int a = 0; a < 3; ++a) {
                float sv = sample_component(o, vel, a, mid, v1[a], check);
                v2[a] = sv + t * (v1[a] - sv);
            }
        }
        float x[3];
        for (int a = 0; a < 3; ++a) x[a] = pp->pos[a] + dt * v2[a];
        if (o->dens_applied) {
            /* Slosh.usf:1506-1510: position-only shift, sampled where the density was measured */
            float pc[3], dx[3];
            for (int a = 0; a < 3; ++a) pc[a] = clampf(pp->pos[a], lo, hi[a]);
            sample_displacement(o, pc, dx);
            for (int a = 0; a < 3; ++a) x[a] = x[a] + dx[a];
        }
        int c[3];
        pos_cell(o, pp->pos, c);
        if (cell_index(o, c[0], c[1], c[2]) < BRICK_CELLS) {
            for (int a = 0; a < 3; ++a) {
                nv[a] = pp->vel[a] + grav[a] * dt;
                x[a] = pp->pos[a] + dt * nv[a];
            }
            memset(cm, 0, sizeof(cm));
        }
        if (o->nman) force_fields(o, x, nv, dt);
        oracle_collider_response(o->col, o->ncol, radius, friction, o->h, x, nv);
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
        float speed = len3(nv);
        if (speed > o->s.max_speed) clamped++;
        if (o->nman && drained_at(o, x, dt, i, frame_seed)) {
            /* Slosh.usf:1649-1652: the record becomes a dead slot (age -1), dropped by the next sort */
            memset(pp, 0, sizeof(*pp));
            pp->age = -1.0f;
            drained++;
            continue;
        }
        for (int a = 0; a < 3; ++a) {
            pp->pos[a] = x[a];
            pp->vel[a] = nv[a];
        }
        memcpy(pp->c, cm, sizeof(cm));
        pp->age += dt;
        if (speed > maxspeed) maxspeed = speed;
        touch_pos(o, pp->pos);
    }
    o->res.max_particle_speed = maxspeed;
    o->res.clamped_particles = clamped;
    o->drained_total += drained;
    o->res.drained_particles = o->drained_total;
}

/* ------------------------------------------------------------------------------------------------------------ */
/* One substep in the order of SloshProxy.cpp:544-713 (sort every step, one substep, fixed dt). */
int osim_step_until(osim *o, float dt, const float g[3], int stop_after)
{
    if (!(dt > 0.0f)) return SLOSH_EINVAL;
    o->frame++;
    o->res.sim_frame = o->frame;
    float warm = o->prev_dt > 0.0f ? dt / o->prev_dt : 1.0f;
    const float dt_prev = o->prev_dt;
    o->prev_dt = dt;
    stage_sort(o);
    o->res.live_particles = o->np;
    if (stop_after <= SLOSH_STAGE_SORT) return SLOSH_OK;
    stage_pool(o);
    if (stop_after <= SLOSH_STAGE_POOL) return SLOSH_OK;
    stage_p2g(o, warm, dt, dt_prev);
    if (stop_after <= SLOSH_STAGE_P2G) return SLOSH_OK;
    stage_normalize(o, dt, g);
    {
        const float alpha = o->s.kinematic_viscosity * dt / (o->h * o->h);
        if (alpha > 0.0f && o->s.viscosity_iterations > 0) stage_viscosity(o, alpha, o->s.viscosity_iterations);
    }
    if (stop_after <= SLOSH_STAGE_NORMALIZE) return SLOSH_OK;
    stage_divergence(o, dt);
    if (stop_after <= SLOSH_STAGE_DIVERGENCE) return SLOSH_OK;
    o->res.solver_iterations = 0;
    o->res.residual_rel = 0.0;
    o->res.converged = 1;
    if (o->s.pressure_solver == SLOSH_SOLVER_MGPCG) solve_mgpcg(o);
    else if (o->s.pressure_solver == SLOSH_SOLVER_RBSOR) solve_rbsor(o);
    if (stop_after <= SLOSH_STAGE_SOLVE) return SLOSH_OK;
    stage_project(o);
    stage_coupling(o, dt);
    o->dens_applied = 0;
    if ((o->s.flags & SLOSH_FLAG_DENSITY_PROJECTION) && o->nactive) {
        int interval = o->s.density_projection_interval > 0 ? o->s.density_projection_interval : 1;
        float relax = 1.0f - powf(1.0f - o->s.density_relaxation, (float)interval);
        if (relax > 0.0f && o->frame % (uint64_t)interval == 0u) {
            int keep = o->s.pressure_solver == SLOSH_SOLVER_NONE;
            stage_density_rhs(o, relax, keep);
            if (!keep) solve_density(o);
            stage_density_disp(o);
            o->dens_applied = 1;
        }
    }
    if (stop_after <= SLOSH_STAGE_PROJECT) return SLOSH_OK;
    for (int k = 0; k < o->s.extrapolation_layers; ++k) stage_extrapolate(o, (uint32_t)k);
    if (stop_after <= SLOSH_STAGE_EXTRAPOLATE) return SLOSH_OK;
    stage_g2p(o, dt, g);
    return SLOSH_OK;
}

int osim_get_results(const osim *o, slosh_results *out)
{
    *out = o->res;
    out->live_particles = o->np;
    return SLOSH_OK;
}

int osim_write_pressure(osim *o, const float *dense)
{
    for (int z = 0; z < o->n[2]; ++z)
        for (int y = 0; y < o->n[1]; ++y)
            for (int x = 0; x < o->n[0]; ++x) {
                uint32_t ci = cell_index(o, x, y, z);
                if (ci < BRICK_CELLS) continue;
                o->pres[ci] = (double)dense[(size_t)x + (size_t)o->n[0] * ((size_t)y + (size_t)o->n[1] * (size_t)z)];
            }
    return SLOSH_OK;
}

int osim_write_density_q(osim *o, const float *dense)
{
    if (!o->dq) return SLOSH_ESTATE;
    for (int z = 0; z < o->n[2]; ++z)
        for (int y = 0; y < o->n[1]; ++y)
            for (int x = 0; x < o->n[0]; ++x) {
                uint32_t ci = cell_index(o, x, y, z);
                if (ci < BRICK_CELLS) continue;
                o->dq[ci] = (double)dense[(size_t)x + (size_t)o->n[0] * ((size_t)y + (size_t)o->n[1] * (size_t)z)];
            }
    return SLOSH_OK;
}

int osim_read_grid(const osim *o, int field, void *dst, uint64_t bytes)
{
    size_t cells = (size_t)o->n[0] * o->n[1] * o->n[2];
    if (field == SLOSH_FIELD_BRICK_TABLE) {
        if (bytes < (uint64_t)o->nbricks * 4u) return SLOSH_ECAPACITY;
        memcpy(dst, o->table, (size_t)o->nbricks * 4u);
        return SLOSH_OK;
    }
    if (field == SLOSH_FIELD_PRESSURE_F64) {
        if (bytes < (uint64_t)cells * 8u) return SLOSH_ECAPACITY;
        double *outd = (double *)dst;
        for (int z = 0; z < o->n[2]; ++z)
            for (int y = 0; y < o->n[1]; ++y)
                for (int x = 0; x < o->n[0]; ++x) {
                    uint32_t ci = cell_index(o, x, y, z);
                    size_t di = (size_t)x + (size_t)o->n[0] * ((size_t)y + (size_t)o->n[1] * (size_t)z);
                    outd[di] = ci < BRICK_CELLS ? 0.0 : o->pres[ci];
                }
        return SLOSH_OK;
    }
    if (field == SLOSH_FIELD_DENSITY_Q) {
        if (!o->dq) return SLOSH_ESTATE;
        if (bytes < (uint64_t)cells * 4u) return SLOSH_ECAPACITY;
        float *outq = (float *)dst;
        for (int z = 0; z < o->n[2]; ++z)
            for (int y = 0; y < o->n[1]; ++y)
                for (int x = 0; x < o->n[0]; ++x) {
                    uint32_t ci = cell_index(o, x, y, z);
                    size_t di = (size_t)x + (size_t)o->n[0] * ((size_t)y + (size_t)o->n[1] * (size_t)z);
                    outq[di] = ci < BRICK_CELLS ? 0.0f : (float)o->dq[ci];
                }
        return SLOSH_OK;
    }
    int planes = 1;
    const uint32_t *src = NULL;
    switch (field) {
    case SLOSH_FIELD_DENSITY_RHS:
        if (!o->drhs) return SLOSH_ESTATE;
        src = (const uint32_t *)o->drhs;
        break;
    case SLOSH_FIELD_DENSITY_DISP:
        if (!o->ddisp) return SLOSH_ESTATE;
        src = (const uint32_t *)o->ddisp;
        planes = 3;
        break;
    case SLOSH_FIELD_CELL_FLAGS: src = o->ctype; break;
    case SLOSH_FIELD_PRESSURE: {
        if (bytes < (uint64_t)cells * 4u) return SLOSH_ECAPACITY;
        float *outp = (float *)dst;
        for (int z = 0; z < o->n[2]; ++z)
            for (int y = 0; y < o->n[1]; ++y)
                for (int x = 0; x < o->n[0]; ++x) {
                    uint32_t ci = cell_index(o, x, y, z);
                    size_t di = (size_t)x + (size_t)o->n[0] * ((size_t)y + (size_t)o->n[1] * (size_t)z);
                    outp[di] = ci < BRICK_CELLS ? 0.0f : (float)o->pres[ci];
                }
        return SLOSH_OK;
    }
    case SLOSH_FIELD_DIVERGENCE: src = (const uint32_t *)o->div; break;
    case SLOSH_FIELD_CELL_DIAG: src = (const uint32_t *)o->diag; break;
    case SLOSH_FIELD_VEL: src = (const uint32_t *)o->vel[o->cur]; planes = 3; break;
    case SLOSH_FIELD_FACE_FLAGS: src = o->fflags; planes = 3; break;
    case SLOSH_FIELD_FACE_WEIGHT: src = (const uint32_t *)o->fweight; planes = 3; break;
    case SLOSH_FIELD_ACCUM: src = (const uint32_t *)o->accum; planes = 6; break;
    case 8 /* cell counts (oracle only) */: src = (const uint32_t *)o->ccount; break;
    default: return SLOSH_EINVAL;
    }
    if (bytes < (uint64_t)cells * 4u * (uint64_t)planes) return SLOSH_ECAPACITY;
    uint32_t *out = (uint32_t *)dst;
    for (int pl = 0; pl < planes; ++pl)
        for (int z = 0; z < o->n[2]; ++z)
            for (int y = 0; y < o->n[1]; ++y)
                for (int x = 0; x < o->n[0]; ++x) {
                    uint32_t ci = cell_index(o, x, y, z);
                    size_t di = (size_t)pl * cells + (size_t)x + (size_t)o->n[0] * ((size_t)y + (size_t)o->n[1] * (size_t)z);
                    out[di] = ci < BRICK_CELLS ? 0u : src[(size_t)pl * o->pool_cells + ci];
                }
    return SLOSH_OK;
}

/* ------------------------------------------------------------------------------------------------------------ */
/* FP64 judge: MG-preconditioned CG from zero in FP64 on the level-0 operator of the current classification. */
int osim_solve_fp64(const osim *oc, double rel_tol, int max_iter, double *dense_out, double *rel_res_out)
{
    osim *o = (osim *)oc; /* scratch vectors only; no simulation state changes */
    mg_build_types(o);
    coarse_factor(o);
    const uint32_t nb = lv_nblocks(o, 0);
#pragma omp parallel for schedule(dynamic, 16)
    for (uint32_t blk = 0; blk < nb; ++blk) {
        double bbb = 0.0;
        lv_iter it = lv_block(0, blk);
        while (lv_next(o, &it)) {
            uint32_t i = it.idx;
            o->pd[i] = 0.0;
            o->dd[i] = 0.0;
            o->qd[i] = 0.0;
            if ((o->ctype[i] & CELL_TYPE_MASK) != CELL_FLUID) {
                o->rd[i] = 0.0; o->zd[i] = 0.0;
                continue;
            }
            double b = -(double)o->div[i];
            o->rd[i] = b;
            double dg = (double)(o->diag[i] + DIAG_EPS);
            o->zd[i] = is_red(it.x, it.y, it.z) ? b / dg : 0.0;
            bbb += b * b;
        }
        o->part[4u * blk] = bbb;
    }
    double bb = part_sum(o, nb, 0);
    double bnorm = sqrt(bb);
    int iters = 0;
    double rel = 0.0;
    if (bnorm > 0.0) {
        double rho = vcycle_d(o);
#pragma omp parallel for schedule(dynamic, 16)
        for (uint32_t blk = 0; blk < nb; ++blk) {
            lv_iter it = lv_block(0, blk);
            while (lv_next(o, &it)) o->dd[it.idx] = o->zd[it.idx];
        }
        for (iters = 1; iters <= max_iter; ++iters) {
#pragma omp parallel for schedule(dynamic, 16)
            for (uint32_t blk = 0; blk < nb; ++blk) {
                double dqb = 0.0;
                lv_iter it = lv_block(0, blk);
                while (lv_next(o, &it)) {
                    uint32_t i = it.idx;
                    if ((o->ctype[i] & CELL_TYPE_MASK) != CELL_FLUID) continue;
                    o->qd[i] = apply_row_d(o, it.x, it.y, it.z, i, o->dd);
                    dqb += o->dd[i] * o->qd[i];
                }
                o->part[4u * blk] = dqb;
            }
            double dq = part_sum(o, nb, 0);
            double alpha = rho / dq;
#pragma omp parallel for schedule(dynamic, 16)
            for (uint32_t blk = 0; blk < nb; ++blk) {
                double rrb = 0.0;
                lv_iter it = lv_block(0, blk);
                while (lv_next(o, &it)) {
                    uint32_t i = it.idx;
                    if ((o->ctype[i] & CELL_TYPE_MASK) != CELL_FLUID) continue;
                    o->pd[i] += alpha * o->dd[i];
                    o->rd[i] -= alpha * o->qd[i];
                    rrb += o->rd[i] * o->rd[i];
                    double dg = (double)(o->diag[i] + DIAG_EPS);
                    o->zd[i] = is_red(it.x, it.y, it.z) ? o->rd[i] / dg : 0.0;
                }
                o->part[4u * blk] = rrb;
            }
            double rr = part_sum(o, nb, 0);
            rel = sqrt(rr) / bnorm;
            if (rel <= rel_tol) break;
            double rho_new = vcycle_d(o);
            double beta = rho_new / rho;
            rho = rho_new;
#pragma omp parallel for schedule(dynamic, 16)
            for (uint32_t blk = 0; blk < nb; ++blk) {
                lv_iter it = lv_block(0, blk);
                while (lv_next(o, &it)) o->dd[it.idx] = o->zd[it.idx] + beta * o->dd[it.idx];
            }
        }
    }
    if (dense_out) {
        size_t cells = (size_t)o->n[0] * o->n[1] * o->n[2];
        memset(dense_out, 0, cells * sizeof(double));
        lv_iter it = {0, 0, 0, 0, 0, 0, 0, 0};
        while (lv_next(o, &it)) {
            if ((o->ctype[it.idx] & CELL_TYPE_MASK) != CELL_FLUID) continue;
            dense_out[(size_t)it.x + (size_t)o->n[0] * ((size_t)it.y + (size_t)o->n[1] * (size_t)it.z)] = o->pd[it.idx];
        }
    }
    if (rel_res_out) *rel_res_out = rel;
    return iters;
}

static double true_residual_impl(const osim *oc, const float *dense_p, const double *dense_pd);

double osim_true_residual(const osim *oc, const float *dense_p) { return true_residual_impl(oc, dense_p, NULL); }

double osim_true_residual_d(const osim *oc, const double *dense_p) { return true_residual_impl(oc, NULL, dense_p); }

static double true_residual_impl(const osim *oc, const float *dense_p, const double *dense_pd)
{
    osim *o = (osim *)oc;
    const uint32_t nb = lv_nblocks(o, 0);
#pragma omp parallel for schedule(dynamic, 16)
    for (uint32_t blk = 0; blk < nb; ++blk) {
        lv_iter it = lv_block(0, blk);
        while (lv_next(o, &it)) {
            size_t di = (size_t)it.x + (size_t)o->n[0] * ((size_t)it.y + (size_t)o->n[1] * (size_t)it.z);
            double pv = dense_p ? (double)dense_p[di] : (dense_pd ? dense_pd[di] : o->pres[it.idx]);
            o->pd[it.idx] = (o->ctype[it.idx] & CELL_TYPE_MASK) == CELL_FLUID ? pv : 0.0;
        }
    }
#pragma omp parallel for schedule(dynamic, 16)
    for (uint32_t blk = 0; blk < nb; ++blk) {
        double bbb = 0.0, rrb = 0.0;
        lv_iter it = lv_block(0, blk);
        while (lv_next(o, &it)) {
            uint32_t i = it.idx;
            if ((o->ctype[i] & CELL_TYPE_MASK) != CELL_FLUID) continue;
            double b = -(double)o->div[i];
            double r = b - apply_row_d(o, it.x, it.y, it.z, i, o->pd);
            bbb += b * b;
            rrb += r * r;
        }
        o->part[4u * blk] = bbb;
        o->part[4u * blk + 1u] = rrb;
    }
    const double bb = part_sum(o, nb, 0), rr = part_sum(o, nb, 1);
    return bb > 0.0 ? sqrt(rr / bb) : 0.0;
}
