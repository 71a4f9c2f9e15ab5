/*
 * libmcb — MI355X-native D3Q19 phase-field bubble lattice-Boltzmann step.
 *
 * C ABI (plain pointers, sizes and status codes; no HIP or torch types).
 * It replaces the reference's time step and driver surface
 * (kreeyto/MULTIC-BUBBLE-CUDA):
 *   - global device buffers + initializeVars()      src/var.cu:13-18, 57-114
 *   - init kernels launched by main()              src/main.cu:56-68, src/init.cu:3-80
 *   - the time loop body (10 launches, 3 syncs)    src/main.cu:74-182
 *   - copyAndSaveToBinary / generateSimulationInfoFile
 *                                                  src/auxFunctions.cu:12-67
 *   - freeMemory                                   src/auxFunctions.cu:4-10
 * The per-kernel operator face (src/kernels.cuh) is in mcb_ops.h.
 *
 * Ownership: the library owns device memory of a context; the caller owns
 * every host buffer. Errors: every call returns mcb_status and never aborts;
 * mcb_last_error() holds a message for the last failing call on a context.
 * Threading: a context is used by one host thread at a time; all its work is
 * ordered on the context's stream; read calls synchronise that stream.
 */
#ifndef MCB_H
#define MCB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCB_ABI_VERSION 1

typedef enum mcb_status {
    MCB_OK = 0,
    MCB_ERR_INVALID_ARG = 1,   /* bad parameter, size or pointer */
    MCB_ERR_NO_DEVICE = 2,     /* no usable gfx950 device */
    MCB_ERR_OUT_OF_MEMORY = 3, /* device allocation failed */
    MCB_ERR_HIP = 4,           /* HIP runtime or launch error */
    MCB_ERR_UNSUPPORTED = 5,   /* combination not implemented */
    MCB_ERR_IO = 6,            /* file could not be written */
    MCB_ERR_NONFINITE = 7,     /* NaN/Inf detected by a diagnostic */
    MCB_ERR_CAPACITY = 8       /* admission would exceed a memory budget */
} mcb_status;

/* Boundary handling. REFERENCE reproduces src/lbm.cu (interior-only
 * updates, ghost-layer fgBoundary + zero-gradient phi); PERIODIC updates
 * every cell with periodic wrap (the benchmark cases). */
typedef enum mcb_bc_mode { MCB_BC_PERIODIC = 0, MCB_BC_REFERENCE = 1 } mcb_bc_mode;

/* Lattice weight table. LITERAL = the reference's float literals
 * (src/var.cu:40-45, sum 1+2^-26); UNIT_SUM = nearest floats with exact unit
 * sum (removes a +1.49e-8/step mass gain). */
typedef enum mcb_weight_set { MCB_WEIGHTS_LITERAL = 0, MCB_WEIGHTS_UNIT_SUM = 1 } mcb_weight_set;

/* Population storage precision; arithmetic is FP32 in every form. */
typedef enum mcb_storage { MCB_STORE_FP32 = 0, MCB_STORE_FP16 = 1, MCB_STORE_BF16 = 2 } mcb_storage;

/* Kernel path. FUSED = 2 passes per step (phase recovery, then
 * normals/curvature/moments/collisions fused with LDS-staged stencils);
 * STAGED = the reference's per-kernel decomposition (mcb_ops.h kernels).
 * AUTO picks FUSED for PERIODIC cases, and for REFERENCE cases in FP32 at
 * tau_g = 1 without moving_wall (the reference preset); STAGED otherwise. */
typedef enum mcb_path { MCB_PATH_AUTO = 0, MCB_PATH_FUSED = 1, MCB_PATH_STAGED = 2 } mcb_path;

/* Interface normal n between steps on the FUSED FP32 path at tau_g = 1 (g in
 * moment form: the phase pass rebuilds g from phi, u and n of each source
 * cell). STORED: the collide pass writes n beside u. RECOMPUTED: the phase
 * pass recomputes n from phi with gradCalc's operations (src/lbm.cu:50-71),
 * the same bits, 24 fewer bytes per cell and step and more arithmetic. AUTO:
 * RECOMPUTED where the phase pass runs z-chunks of 32 planes or more (large
 * lattices), STORED otherwise. Results are identical in every mode; other
 * paths and forms ignore it. */
typedef enum mcb_normals { MCB_NORMALS_AUTO = 0, MCB_NORMALS_STORED = 1, MCB_NORMALS_RECOMPUTED = 2 } mcb_normals;

/* How the FUSED FP32 path keeps f between steps. POPULATIONS: the 19
 * post-collision populations (pull form). MOMENTS: the 13 values
 * collisionFluid reads (src/lbm.cu:258-287: rho, u, Pi^neq, F of each cell's
 * last collision); the collide pass rebuilds every f*_l from them with
 * collisionFluid's own operations, so results are bit-identical, with 52
 * bytes read and 52 written per cell instead of 76 + 76 (+ 12 for u). No
 * counterpart in the reference. MOMENTS needs FUSED, FP32, tau_g = 1 and
 * PERIODIC boundaries (MCB_ERR_UNSUPPORTED otherwise); f can be read, not
 * written, in this form. */
typedef enum mcb_f_form { MCB_F_POPULATIONS = 0, MCB_F_MOMENTS = 1 } mcb_f_form;

typedef struct mcb_params {
    int nx, ny, nz;     /* lattice sides (>= 3; reference 64 x 64 x 256, src/var.cu:3-4) */
    int bc_mode;        /* mcb_bc_mode */
    int weight_set;     /* mcb_weight_set */
    int storage;        /* mcb_storage */
    int path;           /* mcb_path */
    int device;         /* HIP device ordinal */
    float tau_f;        /* fluid relaxation time H_TAU (src/var.cu:21), > 0.5 */
    float tau_g;        /* phase relaxation time; 1 = the reference collisionPhase */
    float sigma;        /* surface tension H_SIGMA (src/var.cu:25) */
    float sharp_c;      /* interface sharpening H_SHARP_C; < 0 selects 0.15f*3.0f (src/var.cu:24) */
    int moving_wall;    /* 1: also run the reference's movingWallBC (dead code, src/lbm.cu:430-461) */
    int normals;        /* mcb_normals */
    int f_form;         /* mcb_f_form */
    int reserved[5];
} mcb_params;

typedef struct mcb_bubble {
    float cx, cy, cz;   /* centre in cell-index coordinates (reference: N*0.5f, src/init.cu:41-43) */
    float r;            /* radius (reference 20.0f, src/init.cu:39) */
} mcb_bubble;

/* Readable/writable fields. PHI is the phase buffer as the reference keeps
 * d_phi: after mcb_step it holds the phi computed at the start of the last
 * iteration (REFERENCE mode: with boundaryConditions applied). RHO, UX, UY,
 * UZ, FX, FY, FZ, NORMX, NORMY, NORMZ hold the values of the most recent
 * mcb_compute_macros call. F and G (19 planes, population-major, x fastest)
 * are the pre-collision populations in the reference layout (src/header.cuh:21).
 * On the FUSED path with REFERENCE boundaries, G is neither read nor written
 * and PHI is not written (MCB_ERR_UNSUPPORTED; the STAGED path keeps every
 * reference buffer); F is read and written exactly. */
typedef enum mcb_field {
    MCB_FIELD_PHI = 0, MCB_FIELD_RHO = 1, MCB_FIELD_UX = 2, MCB_FIELD_UY = 3, MCB_FIELD_UZ = 4,
    MCB_FIELD_FX = 5, MCB_FIELD_FY = 6, MCB_FIELD_FZ = 7,
    MCB_FIELD_NORMX = 8, MCB_FIELD_NORMY = 9, MCB_FIELD_NORMZ = 10,
    MCB_FIELD_F = 16, MCB_FIELD_G = 17
} mcb_field;

typedef struct mcb_ctx mcb_ctx;

/* Reference case constants (src/var.cu:3-4, 21-25; src/init.cu:39). */
mcb_status mcb_params_reference(mcb_params* p);
/* Check a parameter set without touching a device. */
mcb_status mcb_validate_params(const mcb_params* p);
/* Device bytes a context with these parameters allocates. */
size_t mcb_footprint_bytes(const mcb_params* p);

/* Allocate a context on p->device. stream: a hipStream_t or NULL (the
 * library then creates its own). Replaces initializeVars (src/var.cu:57-114). */
mcb_status mcb_create(const mcb_params* p, void* stream, mcb_ctx** out);
mcb_status mcb_destroy(mcb_ctx* ctx);

/* Initial state of src/main.cu:56-68: rho = 1, u = 0, f and g at
 * equilibrium, phi = max over bubbles of 0.5+0.5*tanh(2(R-r)/width)
 * (src/init.cu:39-46; width 3 in the reference). PERIODIC mode measures
 * minimum-image distances; REFERENCE mode leaves the faces at phi = 0. */
mcb_status mcb_init_bubbles(mcb_ctx* ctx, int n, const mcb_bubble* bubbles, float width);
/* Same, from an arbitrary host phi field of nx*ny*nz floats. */
mcb_status mcb_init_phi(mcb_ctx* ctx, const float* phi_host, size_t n);

/* Advance nsteps full iterations of the reference loop body. */
mcb_status mcb_step(mcb_ctx* ctx, int nsteps);
/* Advance nsteps and report the device time of those steps (hipEvents). */
mcb_status mcb_step_timed(mcb_ctx* ctx, int nsteps, double* ms);
/* Observables of the current state (phases phiCalc..momentiCalc of the
 * next iteration, without advancing): fills PHI, RHO, U*, F*, NORM*. */
mcb_status mcb_compute_macros(mcb_ctx* ctx);

mcb_status mcb_read_field(mcb_ctx* ctx, int field, float* host, size_t n);
mcb_status mcb_write_field(mcb_ctx* ctx, int field, const float* host, size_t n);

/* FP64 diagnostics over the fields of the last mcb_compute_macros:
 * sum of phi and kinetic energy sum(0.5*rho*|u|^2), reduced on the device
 * with a fixed-order tree. Returns MCB_ERR_NONFINITE if either is not finite. */
mcb_status mcb_reduce(mcb_ctx* ctx, double* sum_phi, double* kinetic);

mcb_status mcb_synchronize(mcb_ctx* ctx);

/* The FP64 diagnostics of mcb_reduce (sum of phi, kinetic energy
 * sum(0.5*rho*|u|^2), same kernels, same fixed-order tree, same bits) over
 * caller-owned device fields of NX*NY*NZ floats, written as two doubles to
 * the device buffer out; asynchronous on stream (a hipStream_t or NULL), on
 * the current device. It allocates nothing: the per-block partial sums sit
 * in a 16 KiB array that the library holds on every device it runs on until
 * the process ends. No counterpart in the reference (it computes no
 * diagnostics); it is for callers that cannot allocate device memory, such
 * as code under hipGraph capture or serving loops that forbid allocation.
 * Calls that overlap on one device share that array and race: order them.
 * Lattices of 2^30 cells or more are rejected (MCB_ERR_INVALID_ARG), as for
 * every context. */
mcb_status mcb_reduce_fields_noalloc(const float* phi, const float* rho, const float* ux, const float* uy,
                                     const float* uz, int NX, int NY, int NZ, double* out, void* stream);

/* Device time of each pass of the fused step, summed over nsteps steps
 * (hipEvents on the context stream around each launch):
 * out_ms[0] = phase recovery (k_phase), out_ms[1] = collision (k_collide).
 * The steps advance the state. No counterpart in the reference (its only
 * timer is a whole-program wall clock, src/main.cu:203-211). */
mcb_status mcb_step_kernel_times(mcb_ctx* ctx, int nsteps, double* out_ms, int n_out);

/* Tile of the fused collision pass: tx x ty cells per workgroup marching
 * zc planes; (tx, ty) must be one of mcb_tiling_at(0..mcb_tiling_count()-1). */
mcb_status mcb_set_tiling(mcb_ctx* ctx, int tx, int ty, int zc);
mcb_status mcb_get_tiling(const mcb_ctx* ctx, int* tx, int* ty, int* zc);
int mcb_tiling_count(void);
mcb_status mcb_tiling_at(int i, int* tx, int* ty);

/* Time every supported tile shape and z-chunk on a scratch lattice with these
 * parameters (on p->device) and remember the fastest; later contexts and
 * batch cases of the same sides and storage use it. No reference counterpart
 * (the reference launches fixed 8x8x8 blocks, src/main.cu:43-46). */
mcb_status mcb_autotune(const mcb_params* p, int* tx, int* ty, int* zc, double* ms_per_step);

/* ---- Batches: many independent periodic cases resident on one device ----
 * A batch step advances every case by one step with one launch per pass for
 * all cases (grouped by tile shape); two steps at a time are replayed from a
 * captured hipGraph. All cases of a batch share one storage form. Cases never
 * exchange data: a case's results do not depend on the rest of the batch.
 * No reference counterpart (the reference runs one lattice per process). */
typedef struct mcb_batch mcb_batch;
/* Device bytes one case occupies in a batch (populations and anchors). */
size_t mcb_case_bytes(const mcb_params* p);
mcb_status mcb_batch_create(int device, int storage, void* stream, mcb_batch** out);
mcb_status mcb_batch_destroy(mcb_batch* b);
/* Add a case with the initial state of mcb_init_bubbles / mcb_init_phi. */
mcb_status mcb_batch_add(mcb_batch* b, const mcb_params* p, int n, const mcb_bubble* bubbles, float width,
                         int* case_id);
mcb_status mcb_batch_add_phi(mcb_batch* b, const mcb_params* p, const float* phi, size_t n, int* case_id);
mcb_status mcb_batch_remove(mcb_batch* b, int case_id);
mcb_status mcb_batch_step(mcb_batch* b, int nsteps);
mcb_status mcb_batch_step_timed(mcb_batch* b, int nsteps, double* ms);
mcb_status mcb_batch_compute_macros(mcb_batch* b, int case_id);
mcb_status mcb_batch_read_field(mcb_batch* b, int case_id, int field, float* host, size_t n);
mcb_status mcb_batch_write_field(mcb_batch* b, int case_id, int field, const float* host, size_t n);
mcb_status mcb_batch_reduce(mcb_batch* b, int case_id, double* sum_phi, double* kinetic);
mcb_status mcb_batch_case_steps(mcb_batch* b, int case_id, long long* steps);
mcb_status mcb_batch_set_tiling(mcb_batch* b, int case_id, int tx, int ty, int zc);
mcb_status mcb_batch_synchronize(mcb_batch* b);
int mcb_batch_count(const mcb_batch* b);
size_t mcb_batch_bytes(const mcb_batch* b);
const char* mcb_batch_last_error(const mcb_batch* b);

/* ---- State images: checkpoint and restart ----
 * No counterpart in the reference: it writes phi snapshots only
 * (src/main.cu:184-189), which cannot restart a run (f and g are needed).
 * An image holds a context's (or batch case's) complete lattice state at a
 * step boundary in the form the next step reads (populations as stored, g as
 * populations or in moment form, the state's phi, the 16-bit mass-drift
 * state) plus its parameters and step count. Loading it into a context or
 * case with the same parameters (device and requested path aside) continues
 * the run bit for bit. The layout is this library's (version-checked), not
 * the reference's; F and G in the reference layout stay available through
 * mcb_read_field. */
/* Bytes of the context's state image. */
mcb_status mcb_state_size(mcb_ctx* ctx, size_t* bytes);
/* Write the image into buf (bytes >= mcb_state_size). */
mcb_status mcb_save_state(mcb_ctx* ctx, void* buf, size_t bytes);
/* Restore an image into a context created with the same parameters; a
 * context without an initial state is initialised by this call. */
mcb_status mcb_load_state(mcb_ctx* ctx, const void* buf, size_t bytes);
/* Parameters and step count recorded in an image (no device needed). */
mcb_status mcb_state_info(const void* buf, size_t bytes, mcb_params* params, long long* steps);
mcb_status mcb_batch_state_size(mcb_batch* b, int case_id, size_t* bytes);
mcb_status mcb_batch_save_state(mcb_batch* b, int case_id, void* buf, size_t bytes);
mcb_status mcb_batch_load_state(mcb_batch* b, int case_id, const void* buf, size_t bytes);
/* A new case with the image's parameters and state (a preempted case
 * resuming in another batch). */
mcb_status mcb_batch_add_state(mcb_batch* b, const void* buf, size_t bytes, int* case_id);

int mcb_abi_version(void);
const char* mcb_last_error(const mcb_ctx* ctx);
const char* mcb_status_string(mcb_status s);
/* Parameters the context was created with. */
mcb_status mcb_get_params(const mcb_ctx* ctx, mcb_params* out);

/* ---- Host-side helpers (no device needed) ---- */

/* A complete case: parameters, initial bubbles, step count and save cadence. */
#define MCB_MAX_BUBBLES 8
typedef struct mcb_case {
    mcb_params params;
    int n_bubbles;
    mcb_bubble bubbles[MCB_MAX_BUBBLES];
    float width;        /* tanh profile width (3 in every preset) */
    int nsteps;         /* NSTEPS (src/main.cu:35); the loop runs NSTEPS+1 iterations */
    int save_every;     /* MACRO_SAVE (src/main.cu:35) */
} mcb_case;

/* Named presets. "reference": src/var.cu + src/init.cu + src/main.cu
 * (64x64x256, ghost-layer boundaries, tau 0.505, one R=20 bubble at the
 * centre, 25000 steps, phi saved every 100). "cfg0" .. "cfg2": the periodic
 * benchmark cases (64^3 R=10 centred; 128^3 seeded R in [12,24] off-centre;
 * 256^3 two seeded non-overlapping bubbles R in [16,32]). "batch": one case
 * of the uneven batch (cube side in {32,48,64,96,128}, R in [0.15,0.3] side,
 * sigma in [0.01,0.1], tau_f in [0.52,0.8], tau_g 1, FP32, 2000 steps).
 * "serve": one serving request (side a multiple of 16 in [48,192], R in
 * [0.1,0.3] side, sigma in [0.01,0.1], tau_f in [0.52,0.8], tau_g 1, steps
 * log-uniform in [500,5000], FP16 storage). Bubbles of "batch" and "serve"
 * sit at a seeded position. Random draws use splitmix64(seed) mapped to
 * [0,1) by the top 53 bits, so every consumer of a (name, seed) pair gets
 * the same case; "batch" and "serve" are meant to be called with seeds
 * base + i for case i. */
mcb_status mcb_case_preset(const char* name, uint64_t seed, mcb_case* out);

/* ---- Serving: bubble cases arriving over time, continuously batched ----
 * No reference counterpart (the reference runs one case per process). */
typedef struct mcb_request {
    mcb_case c;             /* parameters (periodic), bubbles, width; c.nsteps steps are run */
    double arrival_s;       /* arrival time, seconds after the start of the run */
    float* phi_out;         /* optional host buffer of nx*ny*nz floats for the final phi */
} mcb_request;

typedef struct mcb_request_result {
    int status;             /* mcb_status of this request */
    int reserved;
    double arrival_s, admit_s, finish_s;   /* seconds after the start of the run */
    double sum_phi, kinetic;               /* FP64 diagnostics of the final state */
    long long lattice_updates;             /* cells * steps */
} mcb_request_result;

typedef struct mcb_serve_config {
    int device;
    int storage;            /* population storage of every request (MCB_STORE_*) */
    size_t hbm_budget;      /* device bytes the resident cases may hold; 0 = 90% of free memory */
    int max_cases;          /* resident cases at most; 0 = no limit */
    int tick_steps;         /* steps between scheduling decisions at most; 0 = 100 */
    int sequential;         /* 1: no batching, one request at a time (the baseline) */
    int reserved[7];
} mcb_serve_config;

typedef struct mcb_serve_stats {
    double makespan_s;      /* first arrival to last result */
    double lattice_updates; /* sum over completed requests */
    double mlups;           /* lattice_updates / makespan */
    double p50_s, p99_s, mean_s, max_s;    /* time to result (finish - arrival) */
    int completed, rejected;
    int peak_cases, rounds; /* most resident cases; scheduling rounds */
    double peak_bytes;      /* most device bytes held by resident cases */
    /* Host time of the loop by activity, seconds: admitting cases (allocation,
     * initial state), enqueueing steps (including batch rebuilds), waiting
     * for the device, reading out and releasing finished cases, idle (no
     * case resident, waiting for arrivals). */
    double admit_s, enqueue_s, wait_s, retire_s, idle_s;
} mcb_serve_stats;

/* Replay a request trace in real time on one device. A request is admitted
 * once it has arrived, in arrival order, when its footprint (mcb_case_bytes)
 * fits the budget next to the resident cases; resident cases advance together
 * as one batch, and new cases join between scheduling rounds while others
 * run. A round advances min(tick_steps, fewest steps left) steps, so a case
 * leaves at exactly its step count. A request larger than the whole budget is
 * rejected (MCB_ERR_CAPACITY in its result). results[i] belongs to reqs[i];
 * requests must be sorted by arrival time. */
mcb_status mcb_serve_run(const mcb_serve_config* cfg, const mcb_request* reqs, int n,
                         mcb_request_result* results, mcb_serve_stats* stats);
/* A trace of n "serve" presets (seeds seed*1000003 + i) with Poisson arrivals
 * at rate_per_s (exponential gaps from a second splitmix64 stream) and step
 * counts multiplied by step_scale (at least 1 step). */
mcb_status mcb_serve_trace(uint64_t seed, int n, double rate_per_s, double step_scale, mcb_request* out);

/* The initial phi of mcb_init_bubbles, on the host (nx*ny*nz floats). */
mcb_status mcb_host_init_phi(const mcb_params* p, int n, const mcb_bubble* bubbles, float width,
                             float* phi_out, size_t len);

/* <dir>/<id>_info.txt in the reference's format (src/auxFunctions.cu:24-40).
 * mlups <= 0 writes the reference's literal 1.187970e+01. */
mcb_status mcb_write_info_file(const char* path, const char* sim_id, const char* velocity_set,
                               int nx, int ny, int nz, int save_steps, int nsteps, float tau, double mlups);
/* <dir><id>_<var><step:06d>.bin, raw float32 x-fastest (src/auxFunctions.cu:49-67). */
mcb_status mcb_write_field_file(const char* dir, const char* sim_id, const char* var_name, int step,
                                const float* data, size_t n);

/* Device helpers for callers that drive mcb_ops.h on their own buffers. */
mcb_status mcb_device_count(int* count);
mcb_status mcb_device_alloc(int device, size_t bytes, void** dptr);
mcb_status mcb_device_free(void* dptr);
mcb_status mcb_memcpy_htod(void* dst, const void* src, size_t bytes);
mcb_status mcb_memcpy_dtoh(void* dst, const void* src, size_t bytes);
mcb_status mcb_device_synchronize(void);
/* Time of one plain float4 device-to-device copy of `bytes` (mean of iters),
 * the achievable-bandwidth yardstick for the roofline. */
mcb_status mcb_bench_copy(int device, size_t bytes, int iters, double* ms_per_copy);

#ifdef __cplusplus
}
#endif
#endif /* MCB_H */
