/*
 * slosh.h -- C ABI of the MI355X per-step liquid solver (FLIP/APIC on a sparse 8^3-brick MAC grid).
 *
 * This header is the drop-in boundary that replaces the reference's render-thread proxy
 * (SloshDemo/Plugins/Slosh/Source/SloshCore/Public/SloshProxy.h:23-147) and its plain-data contract
 * (SloshDemo/Plugins/Slosh/Source/SloshCore/Public/SloshTypes.h:10-231). A UE-side shim re-implements
 * FSloshProxy over these calls; the solver kernels it drives are hand-written gfx950 HIP.
 *
 * ABI rules
 *   - Plain C: POD structs with fixed sizes (checked by static asserts in the library), plain pointers and
 *     byte counts, no exceptions across the boundary; every call returns a slosh_status.
 *   - One host thread per context. All device work of a context is ordered on one internal HIP stream;
 *     every call that returns host data is synchronous.
 *   - Caller-owned host buffers. The library owns all device memory of a context.
 *   - Settings outside their documented range are rejected with SLOSH_EINVAL (the reference clamped them
 *     silently, SloshProxy.cpp:344-347 and :411-427).
 *
 * Sim space: metres, origin at the domain's min corner. Cell (i,j,k) spans [i h,(i+1) h] x ... ; the outer cell
 * layer is solid wall. Any axis may be "up": gravity is a free vector passed to every step.
 */
#ifndef SLOSH_H_
#define SLOSH_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SLOSH_ABI_VERSION 2u

typedef struct slosh_ctx slosh_ctx; /* opaque; owns all device memory of one simulation domain */

typedef enum slosh_status {
    SLOSH_OK = 0,
    SLOSH_EINVAL = 1,    /* bad argument or setting out of range */
    SLOSH_ENOMEM = 2,    /* host or device allocation failed */
    SLOSH_EDEVICE = 3,   /* HIP runtime error, or no usable gfx950 device */
    SLOSH_ECAPACITY = 4, /* a request exceeds a fixed capacity (particle capacity, collider count, buffer size) */
    SLOSH_ESTATE = 5     /* call not valid in the current state */
} slosh_status;

/* Pressure solvers. Both solve the same operator (diag_c + 1e-4) p_c - sum_{fluid n} p_n = -div_c. */
enum {
    SLOSH_SOLVER_MGPCG = 0, /* multigrid-preconditioned CG, FP64 dots, to ||r||/||b|| <= tolerance */
    SLOSH_SOLVER_RBSOR = 1, /* red-black SOR with sor_weight, to tolerance or max_iterations */
    SLOSH_SOLVER_NONE = 2,  /* keep the (warm-started) pressure as is; used by parity tests of the other kernels */
    SLOSH_SOLVER_FCG_UNET = 3 /* flexible CG preconditioned by M(r) = V(r) + N(r - A V(r)): the MGPCG V-cycle and the
                                 learned U-Net's correction of the residual it leaves (slosh_unet_*); MGPCG is the
                                 fallback when it stalls and the judge of its result; grid dims multiples of 8 */
};

/* Solver settings: FSloshSolverSettings (SloshTypes.h:113-160), restricted to what the per-step path reads. */
typedef struct slosh_settings {
    int32_t grid_dim[3];          /* cells per axis, 4..4096, each a multiple of nothing in particular */
    float cell_size;              /* h in metres, > 0 */
    uint32_t particle_capacity;   /* maximum particle slots */
    float sparse_storage_budget;  /* brick pool capacity as a fraction of all 8^3 bricks, (0, 1] */
    int32_t pressure_solver;      /* SLOSH_SOLVER_* */
    float tolerance;              /* relative residual target ||r||_2 / ||b||_2 (e.g. 1e-6) */
    int32_t max_iterations;       /* MGPCG iteration cap / RB-SOR iteration cap */
    float sor_weight;             /* RB-SOR relaxation, [0.5, 1.95] */
    int32_t ghost_fluid;          /* 1: free-surface p = 0 at the 0.5 density iso level (theta weights) */
    int32_t extrapolation_layers; /* velocity extrapolation layers, 0..8 */
    float boundary_friction;      /* 0 free slip .. 1 no slip */
    float max_speed;              /* m/s; speeds above are counted (reference stats 9..13), not clamped */
    float particle_radius_cells;  /* collider contact radius in cells */
    float density_correction;     /* velocity-based density correction strength (0 in the parity profile) */
    float density_deadband;
    uint32_t flags;               /* SLOSH_FLAG_* */
    int32_t mg_coarse_sweeps;     /* MGPCG: red-black smoothing pairs per visit of the dense coarse levels (4 and
                                     coarser), 0..4; 0 and 1 give the reference's single pair (SloshProxy.cpp:904-959) */
    /* Two-sided density projection (SLOSH_FLAG_DENSITY_PROJECTION; SloshTypes.h:128-131): */
    float density_relaxation;             /* fraction of the density error removed per projection, [0, 1] */
    int32_t density_projection_iterations; /* RB-SOR sweeps of its solve (omega 1.7), 0..10000; MGPCG uses 2 iterations */
    int32_t density_projection_interval;  /* project every N steps (1..8; 0 means 1), relaxation compounded */
    float kinematic_viscosity;    /* m^2/s; > 0 runs viscosity_iterations Jacobi sweeps of (I - nu dt Laplacian) u = u*
                                     after the normalise pass (Slosh.usf:657-686, SloshProxy.cpp:603-617) */
    int32_t viscosity_iterations; /* 0..1000 */
    float fluid_density;          /* kg/m^3 (SloshTypes.h:121); converts the solver's pressure to pascals for the
                                     collider feedback, p_Pa = p * density * h / dt (SloshProxy.cpp:562); 0 means
                                     1000 */
    int32_t reserved[7];          /* must be zero */
} slosh_settings;

#define SLOSH_FLAG_NONE 0u
#define SLOSH_FLAG_CAPTURE_RESIDUALS 1u /* keep MGPCG residuals of each solve as training data for slosh_unet_train */
#define SLOSH_FLAG_MG_FP16 2u           /* MGPCG: multigrid levels >= 1 stored in FP16 (FP32 arithmetic, FP32 level-0
                                           residual, each V-cycle scaled to O(1) values); the result still meets the
                                           tolerance */
#define SLOSH_FLAG_MG_FLUID_COARSENING 4u /* MGPCG: a coarse cell with any fluid child is fluid (the reference's rule,
                                             SloshPressureMG.usf:256-278, makes it air when any child is air); the
                                             preconditioner keeps more of the free-surface band, the solution and
                                             its tolerance are unchanged */
#define SLOSH_FLAG_WARM_START_EXTRAPOLATE 8u /* pressure solvers: start from the linear extrapolation in time of the
                                                last two solutions (per unit dt) on cells that were fluid in both,
                                                instead of the last solution scaled by dt/dt_prev (the reference's
                                                warm start, SloshProxy.cpp:575 applied in Slosh.usf:569-598, kept
                                                on every other cell); the solve still stops at the tolerance */
#define SLOSH_FLAG_CAPTURE_VCYCLE_RESIDUALS 32u /* with SLOSH_FLAG_CAPTURE_RESIDUALS: capture r - A V(r), the
                                                   residual each V-cycle leaves, instead of r (the training data of
                                                   the hybrid preconditioner of SLOSH_SOLVER_FCG_UNET) */
#define SLOSH_FLAG_DENSITY_PROJECTION 64u /* the reference's two-sided density projection (Slosh.usf:743-854,
                                             SloshProxy.cpp:656-702): after the pressure projection, solve
                                             A q = relaxation * clamp(rho - 1) with the pressure operator and shift
                                             every particle by -grad q (at most half a cell); it replaces the
                                             velocity density correction (density_correction is then not applied) */
#define SLOSH_FLAG_BINARY_COUPLING 128u /* collider feedback: the pressure force from the binary fluid-solid cell faces
                                           (bSmoothCouplingForce = false, Slosh.usf:930-946) instead of the smoothed
                                           SDF-weighted estimator (Slosh.usf:1138-1216) */
#define SLOSH_FLAG_REPORT_FP64_ERROR 16u /* MGPCG: after each solve, solve the same system again from its solution to
                                            an FP64 true residual of 1e-12 (restarts form b - A p in FP64, so the FP32
                                            vectors and FP16 levels do not limit it) and report the first solution's
                                            relative L2 error against it (slosh_results.fp64_error_rel); the first
                                            solution is kept. A diagnostic: it costs a few solves per step */

/* Emission request: FSloshEmitRequest (SloshTypes.h:169-181). shape 0 = FillCells (8 jittered particles per cell),
 * 1 = `count` random points in the sphere of radius extent[0] about center, 2 = `count` random points in the box of
 * half extents `extent` rotated by `rotation` about center (Slosh.usf:280-358). A nonzero velocity_spread (cone half
 * angle, radians) or speed_jitter (+- fraction of the speed) varies each particle's velocity in every mode. */
enum { SLOSH_EMIT_FILL_CELLS = 0, SLOSH_EMIT_SPHERE = 1, SLOSH_EMIT_BOX = 2 };
typedef struct slosh_emit_request {
    uint32_t shape;
    uint32_t seed;         /* EmitSeed of the jitter hash; the reference increments its own from 1 (SloshProxy.cpp:529) */
    float center[3];
    float extent[3];
    float velocity[3];
    float rotation[4];     /* quaternion x,y,z,w */
    float velocity_spread; /* radians */
    float speed_jitter;
    int32_t cell_min[3];   /* FillCells: first cell; clamped by the caller as QueueFillLocalBox does */
    int32_t cell_dims[3];  /* FillCells: cells per axis */
    uint32_t count;        /* sphere/box: particles */
    uint32_t reserved;
} slosh_emit_request;

/* Analytic collider: FSloshColliderGPU (SloshTypes.h:32-44), 80 bytes. type 0 sphere, 1 box, 2 capsule. */
typedef struct slosh_collider {
    float center[3];
    uint32_t type;
    float rotation[4];
    float extent[3];
    uint32_t flags;
    float linear_velocity[3];
    float pad0;
    float angular_velocity[3];
    float pad1;
} slosh_collider;
#define SLOSH_MAX_COLLIDERS 64u
#define SLOSH_COLLIDER_TWO_WAY 1u /* collider flags: the body receives pressure force feedback (SloshCommon.ush:42) */

/* Per-collider feedback of the last step: FSloshColliderFeedback (SloshTypes.h:197-206), decoded as
 * PollReadbacks_RenderThread does for one substep (SloshProxy.cpp:755-768). Force, torque, wetted faces and fluid
 * velocity are zero for one-way colliders; the cell counts and the centroid are kept for every collider. */
typedef struct slosh_collider_feedback {
    float pressure_force[3];           /* N */
    float pressure_torque[3];          /* N m about the collider's centre */
    float mean_fluid_velocity[3];      /* m/s, fluid cell-centre velocity after the projection over wetted faces */
    float submerged_centroid_offset[3];/* m, mean offset of the submerged cells from the centre */
    int32_t wetted_faces;              /* fluid-solid faces of the body */
    int32_t submerged_cells;           /* body cells whose horizontal walk out of the body first meets dense,
                                          supported liquid */
    int32_t solid_cells;               /* body cells in allocated bricks */
    int32_t reserved;
} slosh_collider_feedback;

/* Particle manipulator: FSloshManipulatorGPU (SloshTypes.h:66-80), 80 bytes. kind 0 drain (removes the particles
 * inside: all at once when param0 = 0, else at rate param0 per second), 1 directional (acceleration strength along
 * axis), 2 radial (strength away from center), 3 swirl (strength around axis, param0 the inward pull ratio), 4 drag
 * (velocity times exp(-strength dt)). shape 0 sphere, 1 box, 2 capsule (as colliders), 3 everywhere. The effect is
 * full inside the shape and fades linearly to 0 over `falloff` metres beyond it (Slosh.usf:1523-1559, 1639-1654;
 * SloshCommon.ush:148-230). */
typedef struct slosh_manipulator {
    float center[3];
    uint32_t kind;
    float rotation[4];
    float extent[3];
    uint32_t shape;
    float axis[3];
    float strength;
    float falloff;
    float param0;
    float param1;
    uint32_t pad;
} slosh_manipulator;
#define SLOSH_MAX_MANIPULATORS 32u

/* Particle records returned by slosh_read_particles. */
enum { SLOSH_LAYOUT_F32_64 = 0, SLOSH_LAYOUT_REF_48 = 1 };
typedef struct slosh_particle { /* 64 B, FP32 APIC matrix */
    float pos[3];
    float age; /* < 0 marks a dead slot */
    float vel[3];
    float c[9]; /* rows Cx, Cy, Cz */
} slosh_particle;
typedef struct slosh_particle_ref48 { /* FSloshParticleGPU (SloshTypes.h:84-91): C as 9 halves in 5 uints */
    float pos[3];
    float age;
    float vel[3];
    uint32_t c[5];
} slosh_particle_ref48;

/* Grid fields for slosh_read_grid / slosh_write_grid. */
enum {
    SLOSH_FIELD_CELL_FLAGS = 0,  /* uint32: bits 0-1 type (0 air, 1 fluid, 2 solid), bit 2 dense, bits 8-15 collider+1 */
    SLOSH_FIELD_PRESSURE = 1,    /* float, scaled pressure p_hat = p dt / (rho h) */
    SLOSH_FIELD_DIVERGENCE = 2,  /* float, sum of face velocity differences (not divided by h) */
    SLOSH_FIELD_CELL_DIAG = 3,   /* float, ghost-fluid operator diagonal (without the 1e-4) */
    SLOSH_FIELD_VEL = 4,         /* float[3 planes]: face velocities U, V, W (each cell owns its min faces) */
    SLOSH_FIELD_FACE_FLAGS = 5,  /* uint32[3 planes] */
    SLOSH_FIELD_FACE_WEIGHT = 6, /* float[3 planes] */
    SLOSH_FIELD_ACCUM = 7,       /* int32[6 planes]: P2G momentum U,V,W then weight U,V,W (fixed point 2^20); defined
                                    between the P2G and normalise stages of a step */
    SLOSH_FIELD_BRICK_TABLE = 9, /* uint32 per 8^3 brick: pool slot, 0 = none (order argument ignored) */
    SLOSH_FIELD_PRESSURE_F64 = 10, /* double: the solver's FP64 pressure */
    SLOSH_FIELD_DENSITY_Q = 11,    /* float: density-projection potential q (cells), written as q by write_grid */
    SLOSH_FIELD_DENSITY_RHS = 12,  /* float: its right-hand side as the solver's "divergence" (-relaxation * error) */
    SLOSH_FIELD_DENSITY_DISP = 13  /* float[3 planes]: face displacements -(q_b - q_a) in cells */
};
enum { SLOSH_ORDER_DENSE = 0 /* x fastest over the whole grid, unallocated cells read 0 */ };

/* Diagnostics of the last step: FSloshResults (SloshTypes.h:208-231) plus solver convergence data. */
typedef struct slosh_results {
    uint64_t sim_frame;
    uint32_t live_particles;
    uint32_t fluid_cells;
    uint32_t active_bricks;        /* allocated bricks after this step's allocation */
    uint32_t pool_overflow_bricks; /* wanted bricks that found the pool full */
    float max_particle_speed;
    float max_divergence;          /* 1/s, post projection, over fluid cells */
    float mean_divergence;         /* 1/s */
    uint32_t clamped_particles;    /* particles above max_speed */
    int32_t solver_iterations;
    int32_t converged;             /* 1 when the FP64 true residual ||b - A p||/||b|| <= tolerance */
    double residual_rel;           /* final ||r||_2 / ||b||_2 of the solver's recurrence residual */
    double rhs_norm;               /* ||b||_2 */
    float step_ms;                 /* device time of the step (HIP events) */
    float solve_ms;                /* device time of the pressure solve */
    int32_t fallback_used;         /* flexible CG: 1 when MGPCG finished the solve, 2 when no network was set */
    float true_residual_rel;       /* ||b - A p||_2 / ||b||_2 of the FP64 pressure after the solve */
    double fp64_error_rel;         /* SLOSH_FLAG_REPORT_FP64_ERROR: ||p - p64|| / ||p64|| over fluid cells against the
                                      FP64-accurate solve of the same system; -1 when not computed */
    uint64_t drained_particles;    /* particles removed by drain manipulators since the last reset */
} slosh_results;

/* ---- lifecycle --------------------------------------------------------------------------------------------- */
/* Fills *s with the parity-profile defaults (APIC, ghost fluid on, density correction 0, MGPCG 1e-6). */
void slosh_default_settings(slosh_settings *s);
/* = FSloshProxy::Allocate (SloshProxy.cpp:185-306). Creates all device state on the current HIP device. */
slosh_status slosh_create(const slosh_settings *s, slosh_ctx **out);
void slosh_destroy(slosh_ctx *ctx);
/* = FSloshFrameInput::bReset: drops every particle and clears all grid state. */
slosh_status slosh_reset(slosh_ctx *ctx);

/* ---- per-frame inputs -------------------------------------------------------------------------------------- */
/* = the emit passes of Step_RenderThread (SloshProxy.cpp:502-540; Slosh.usf:272-358), run at once.
 * Returns SLOSH_ECAPACITY (emitting nothing) if the request does not fit particle_capacity. */
slosh_status slosh_emit(slosh_ctx *ctx, const slosh_emit_request *r);
/* = FSloshFrameInput::Colliders (SloshTypes.h:187); n <= SLOSH_MAX_COLLIDERS; kept until replaced. */
slosh_status slosh_set_colliders(slosh_ctx *ctx, const slosh_collider *c, uint32_t n);
/* = FSloshFrameInput::Manipulators (SloshTypes.h:189); n <= SLOSH_MAX_MANIPULATORS; kept until replaced. */
slosh_status slosh_set_manipulators(slosh_ctx *ctx, const slosh_manipulator *m, uint32_t n);

/* ---- the hot path -------------------------------------------------------------------------------------------- */
/* One substep of Step_RenderThread (SloshProxy.cpp:564-713) with a fixed dt and the given gravity (m/s^2):
 * sort -> brick pool -> classify -> P2G -> normalise -> divergence -> pressure -> project -> extrapolate -> G2P. */
slosh_status slosh_step(slosh_ctx *ctx, float dt, const float gravity[3]);

/* ---- pressure systems of their own -------------------------------------------------------------------------- */
/* No counterpart in the reference. For the training-set generator (64^3 windows of seeded scenes) and for callers
 * that solve pressure systems of their own with the library's solvers. Loads the system
 *     (diag_c + 1e-4) p_c - sum_{fluid neighbours n} p_n = -divergence_c   over the fluid cells c
 * given in dense order (x fastest) over the context's grid: types 0 air, 1 fluid, 2 solid; diag (>= 0: one per
 * fluid neighbour plus the ghost-fluid weights of air neighbours) and the finite divergence are read at fluid cells
 * only (SLOSH_EINVAL otherwise). Cells outside the grid count as solid, so a window cut from a larger
 * grid keeps the diagonal of its original neighbourhood: its matrix is a principal submatrix of the original one.
 * Drops the particles and the grid state (as slosh_reset), allocates the bricks holding fluid with their
 * neighbourhood (SLOSH_ECAPACITY, loading nothing, when the pool is too small) and sets the pressure to 0. */
slosh_status slosh_load_system(slosh_ctx *ctx, const uint8_t *types, const float *diag, const float *divergence);
/* No counterpart in the reference. Solves the current system (the last step's, or a loaded one) from the current
 * pressure with the configured solver, as the step's pressure stage does (capture flags included); the solver's
 * figures go to slosh_results, the pressure to SLOSH_FIELD_PRESSURE / SLOSH_FIELD_PRESSURE_F64. */
slosh_status slosh_solve(slosh_ctx *ctx);

/* ---- readback ------------------------------------------------------------------------------------------------ */
/* Copies the live particles (sorted order of the last step) into dst; *count receives the number written. */
slosh_status slosh_read_particles(slosh_ctx *ctx, void *dst, uint64_t dst_bytes, uint32_t *count, int32_t layout);
/* Copies a grid field in the given order. Per-face fields hold 3 planes (U, V, W) of grid_dim cells each. */
slosh_status slosh_read_grid(slosh_ctx *ctx, int32_t field, void *dst, uint64_t dst_bytes, int32_t order);
/* = FSloshProxy::GetResults (SloshProxy.cpp:177-183), synchronous. */
slosh_status slosh_get_results(slosh_ctx *ctx, slosh_results *out);
/* = FSloshResults::Colliders (SloshTypes.h:210): one entry per collider of the last step, in the order given to
 * slosh_set_colliders; *n receives the collider count and at most `cap` entries are written. Synchronous. */
slosh_status slosh_get_collider_feedback(slosh_ctx *ctx, slosh_collider_feedback *out, uint32_t cap, uint32_t *n);
/* No counterpart in the reference. For callers that cannot allocate device memory (tight render or serving loops,
 * harnesses that forbid hipMalloc): the same copy as slosh_read_particles, staged through one fixed library-wide
 * device buffer of SLOSH_NOALLOC_STAGING_BYTES per GPU (held until the process ends) instead of the context's
 * growable staging buffer, so the call allocates nothing. Returns SLOSH_ECAPACITY, copying nothing, when the live
 * particles times the record size exceed the buffer (1,048,576 records of 64 B, 1,398,101 of 48 B). Every context
 * on a GPU shares the buffer: calls that overlap on one GPU (other host threads) race. */
#define SLOSH_NOALLOC_STAGING_BYTES (64u << 20)
slosh_status slosh_read_particles_noalloc(slosh_ctx *ctx, void *dst, uint64_t dst_bytes, uint32_t *count,
                                          int32_t layout);
/* Zero-copy access for a renderer: packs the particles into a device buffer of the given layout and returns it
 * (valid until the next call that changes particles). = GetParticleBuffer_RenderThread (SloshProxy.h:41). */
slosh_status slosh_device_particles(slosh_ctx *ctx, int32_t layout, void **dev_ptr, uint32_t *count);

/* ---- state injection (checkpoint / resume, tests) ------------------------------------------------------------ */
/* Replaces all particles with n records of layout SLOSH_LAYOUT_F32_64 (no reference counterpart: the reference
 * only emits; this is the resume half of read_particles). The brick touch set is rebuilt from the positions. */
slosh_status slosh_write_particles(slosh_ctx *ctx, const void *src, uint32_t n);
/* Writes a per-cell float field (SLOSH_FIELD_PRESSURE or SLOSH_FIELD_DENSITY_Q) in dense order; cells of
 * unallocated bricks are ignored. Used to resume a warm start; with SLOSH_SOLVER_NONE the written values are kept by
 * the next step (parity tests of the passes that follow a solve). */
slosh_status slosh_write_grid(slosh_ctx *ctx, int32_t field, const void *src, uint64_t src_bytes, int32_t order);

/* ---- test hook ------------------------------------------------------------------------------------------------ */
/* Runs slosh_step but stops after the given stage so intermediate fields can be read back. */
enum {
    SLOSH_STAGE_SORT = 1,
    SLOSH_STAGE_POOL = 2,
    SLOSH_STAGE_P2G = 3,        /* classification + P2G + cell marking */
    SLOSH_STAGE_NORMALIZE = 4,
    SLOSH_STAGE_DIVERGENCE = 5,
    SLOSH_STAGE_SOLVE = 6,
    SLOSH_STAGE_PROJECT = 7,
    SLOSH_STAGE_EXTRAPOLATE = 8,
    SLOSH_STAGE_ALL = 9
};
slosh_status slosh_step_until(slosh_ctx *ctx, float dt, const float gravity[3], int32_t stop_after);

/* ---- profiling (no reference counterpart; the reference's equivalent is its RDG event scopes and GPU stat,
 *      SloshProxy.cpp:12, :342, :566) ---------------------------------------------------------------------------- */
typedef struct slosh_kernel_time {
    char name[32];      /* kernel name as a kernel trace shows it (k_g2p, k_apply_a, ...) */
    uint32_t launches;  /* timed launches since the last reset */
    uint32_t pad;
    double total_ms;    /* summed HIP-event time of those launches */
} slosh_kernel_time;
/* Selects the kernels to time with HIP events on the context's stream: bit i = kernel i of slosh_kernel_times'
 * order, ~0 = all, 0 = none (the default). A kernel launched alone carries the event pair on its dispatch (its own
 * begin and end); a multi-launch set-up stage is bracketed by events recorded around it. Timing turns the replay of
 * solver chunks as HIP graphs off. */
slosh_status slosh_set_kernel_timing(slosh_ctx *ctx, uint64_t mask);
/* Writes up to cap entries (one per kernel, fixed order) and sets *n to the kernel count; out == NULL resets
 * the totals. */
slosh_status slosh_kernel_times(slosh_ctx *ctx, slosh_kernel_time *out, uint32_t cap, uint32_t *n);

/* ---- learned preconditioner (no reference counterpart: the reference preconditions CG with its multigrid
 *      V-cycle only; this is the library's own component) --------------------------------------------------------
 * A 3-level 3-D U-Net (16/32/64 channels, 3x3x3 convolutions, BF16 weights on MFMA, FP32 accumulate) maps the
 * residual and the cell types to the preconditioned residual. Parameters: FP32, per layer W[27][cout][cin] then
 * bias[cout], layers E0a 4->16, E0b 16->16, E1a 16->32, E1b 32->32, B2a 32->64, B2b 64->64, D1p 64->32,
 * D1b 32->32, D0p 32->16, D0b 16->16, head 1x1 16->1. */
uint64_t slosh_unet_param_count(void);
slosh_status slosh_unet_set_weights(slosh_ctx *ctx, const float *w, uint64_t n);
slosh_status slosh_unet_get_weights(slosh_ctx *ctx, float *w, uint64_t n);
/* Applies the network to a residual given in dense order (x fastest) on the system of the last step; writes the
 * preconditioned residual (0 off fluid cells). Diagnostic and test entry point. */
slosh_status slosh_unet_apply(slosh_ctx *ctx, const float *r_dense, float *z_dense);
/* Training loss sum_fluid (A zhat - r/s)^2 / sum_fluid (r/s)^2 for the residual r (dense order) on the system of
 * the last step, and its gradient w.r.t. the parameters (n = slosh_unet_param_count floats). */
slosh_status slosh_unet_loss_grad(slosh_ctx *ctx, const float *r_dense, double *loss, float *grad);
typedef struct slosh_unet_train_config {
    uint32_t iterations;  /* Adam steps */
    uint32_t batch;       /* captured residuals per step */
    float learning_rate;
    uint32_t offset;      /* first captured residual used */
} slosh_unet_train_config;
typedef struct slosh_unet_train_report {
    double first_loss;    /* mean loss of the first step's batch */
    double last_loss;     /* mean loss of the last step's batch */
    uint32_t systems;     /* captured residuals available */
    uint32_t adam_steps;  /* total Adam steps since the weights were set */
} slosh_unet_train_report;
/* Copies captured residual k (0 = the solve's initial residual b - A p0) of the last solve in dense order; *n
 * receives the number of captured residuals. The training-data export of the learned preconditioner. */
slosh_status slosh_unet_read_capture(slosh_ctx *ctx, uint32_t k, float *dst, uint64_t dst_bytes, uint32_t *n);
/* Trains on the residuals captured by the last solve (settings flag SLOSH_FLAG_CAPTURE_RESIDUALS) of the current
 * system, with the library's forward/backward HIP kernels and Adam. */
slosh_status slosh_unet_train(slosh_ctx *ctx, const slosh_unet_train_config *cfg, slosh_unet_train_report *out);

/* ---- batches of independent tanks (no reference counterpart: the reference runs one volume per proxy) ----------
 * A batch steps existing contexts together. Mode 0: every tank's whole step runs on its own stream, tanks on several
 * host threads at once, longest first. Mode 1 (lockstep): the passes before and after the pressure solve run so,
 * while the MGPCG iterations of all tanks run in lockstep on one stream, each pass of an iteration one launch over
 * every tank's bricks (tanks whose solve cannot join, e.g. another solver, solve on their own stream). Solver
 * iteration chunks replay as HIP graphs. The configuration changes timing only: every tank's results equal its
 * single-tank run bit for bit. */
typedef struct slosh_batch slosh_batch;
typedef struct slosh_batch_config {
    uint32_t workers;      /* tanks stepped at once on host threads (1..64) */
    uint32_t solver_chunk; /* MGPCG iterations per read of the solver state (even, 2..32) */
    uint32_t graphs;       /* 1: replay iteration chunks as captured HIP graphs */
    uint32_t mode;         /* 0: each tank on its own stream; 1: lockstep MGPCG iterations over all tanks */
    uint32_t col_bricks;   /* bricks (wavefronts) per workgroup of the level-0 solver passes: 1, 2 or 4 */
} slosh_batch_config;
slosh_status slosh_batch_create(slosh_ctx *const *tanks, uint32_t n, slosh_batch **out);
void slosh_batch_destroy(slosh_batch *batch);
/* One substep of every tank; gravity holds 3 floats per tank. On an error status some tanks may have stopped
 * part-way through the step: reset them (slosh_reset) before stepping again. */
slosh_status slosh_batch_step(slosh_batch *batch, float dt, const float *gravity);
slosh_status slosh_batch_set_config(slosh_batch *batch, const slosh_batch_config *cfg);
slosh_status slosh_batch_get_config(const slosh_batch *batch, slosh_batch_config *cfg);
/* Times candidate configurations on steps_per_candidate real batch steps each (the tanks advance) and keeps the
 * fastest. */
slosh_status slosh_batch_autotune(slosh_batch *batch, float dt, const float *gravity, uint32_t steps_per_candidate,
                                  slosh_batch_config *chosen);

/* Library identification: ABI version and the offload target the kernels were built for. */
uint32_t slosh_abi_version(void);
const char *slosh_build_info(void);
const char *slosh_status_string(slosh_status s);

#ifdef __cplusplus
}
#endif
#endif /* SLOSH_H_ */
