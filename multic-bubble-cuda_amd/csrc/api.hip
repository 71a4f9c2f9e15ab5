// Context API of libmcb (include/mcb.h): device memory, initial state, the
// time step on one HIP stream, readout and diagnostics.
//
// Two step paths share the same per-cell arithmetic (lattice.hpp):
//   FUSED  (PERIODIC, FP32): k_phase + k_collide per step, pull storage
//          with A/B population buffers (fused.hip).
//   STAGED (both modes): the reference's kernel sequence of src/main.cu:79-178
//          on the reference's buffers (ops.hip), one stream, no host syncs.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "common.hpp"
#include "fused.hpp"

using namespace mcb;

// Host-side parameter helpers (host.cpp).
namespace mcb {
mcb_status host_validate(const mcb_params* p, std::string* why);
Phys host_phys(const mcb_params& p);
}

struct mcb_ctx {
    mcb_params p{};
    Phys phys{};
    mcb_op_consts oc{};
    Geo geo{};
    int path = MCB_PATH_FUSED;
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::vector<hipEvent_t> tev;   // per-kernel timing events
    Tiling tiling{64, 4, 16};
    std::string err;
    std::vector<void*> allocs;

    // FUSED state
    float* f[2] = {nullptr, nullptr};
    float* g[2] = {nullptr, nullptr};
    int cur = 0;
    // shared scalar fields (FUSED: macro outputs; STAGED: the reference buffers)
    float* phi = nullptr;
    float *rho = nullptr, *ux = nullptr, *uy = nullptr, *uz = nullptr;
    float *ffx = nullptr, *ffy = nullptr, *ffz = nullptr;
    float *normx = nullptr, *normy = nullptr, *normz = nullptr;
    // STAGED-only buffers (src/var.cu:13-18)
    float *sf = nullptr, *sg = nullptr, *sg_out = nullptr;
    float *indicator = nullptr, *curvature = nullptr;
    float *pxx = nullptr, *pyy = nullptr, *pzz = nullptr, *pxy = nullptr, *pxz = nullptr, *pyz = nullptr;
    double* partial = nullptr;
    double* dsum = nullptr;
    bool initialised = false;
};

namespace {

mcb_status fail(mcb_ctx* c, mcb_status s, const std::string& msg) {
    if (c) c->err = msg;
    return s;
}

mcb_status hip_fail(mcb_ctx* c, hipError_t e, const char* where) {
    if (e == hipSuccess) return MCB_OK;
    return fail(c, hip_status(e), std::string(where) + ": " + hipGetErrorString(e));
}

#define MCB_HIP(c, call)                                          \
    do {                                                          \
        hipError_t e_ = (call);                                   \
        if (e_ != hipSuccess) return hip_fail((c), e_, #call);    \
    } while (0)

// Plane stride: cells rounded up to 256 elements, so every population
// plane starts 1 KiB-aligned.
long long plane_stride(long long n) { return (n + 255) / 256 * 256; }

// The fused kernels address all 19 planes of a buffer through one 32-bit
// buffer resource (lattices up to ~56.5M cells).
bool fused_fits(const mcb_params& p) {
    return plane_stride((long long)p.nx * p.ny * p.nz) <= kMaxFusedPlane;
}

int resolve_path(const mcb_params& p) {
    if (p.path == MCB_PATH_STAGED) return MCB_PATH_STAGED;
    if (p.path == MCB_PATH_FUSED) return MCB_PATH_FUSED;
    return p.bc_mode == MCB_BC_PERIODIC && fused_fits(p) ? MCB_PATH_FUSED : MCB_PATH_STAGED;
}

size_t footprint(const mcb_params& p, int path) {
    const long long n = (long long)p.nx * p.ny * p.nz;
    const long long P = plane_stride(n);
    size_t floats = 0;
    if (path == MCB_PATH_FUSED) floats = 4 * 19 * (size_t)P + 11 * (size_t)P;
    else floats = 3 * 19 * (size_t)P + 19 * (size_t)P;
    return floats * sizeof(float) + (2 * kReduceBlocks + 2) * sizeof(double);
}

mcb_status alloc(mcb_ctx* c, float** out, size_t floats) {
    void* ptr = nullptr;
    hipError_t e = hipMalloc(&ptr, floats * sizeof(float));
    if (e != hipSuccess) return hip_fail(c, e, "hipMalloc");
    c->allocs.push_back(ptr);
    *out = static_cast<float*>(ptr);
    return MCB_OK;
}

CollideArgs collide_args(mcb_ctx* c) {
    CollideArgs a{};
    a.geo = c->geo;
    a.phys = c->phys;
    a.phi = c->phi;
    a.fsrc = c->f[c->cur];
    a.gsrc = c->g[c->cur];
    a.fdst = c->f[c->cur ^ 1];
    a.gdst = c->g[c->cur ^ 1];
    a.rho = c->rho; a.ux = c->ux; a.uy = c->uy; a.uz = c->uz;
    a.fx = c->ffx; a.fy = c->ffy; a.fz = c->ffz;
    a.nx = c->normx; a.ny = c->normy; a.nz = c->normz;
    a.zc = c->tiling.zc;
    return a;
}

int step_mode(const mcb_ctx* c) { return c->p.tau_g == 1.0f ? COLLIDE_STEP : COLLIDE_STEP_BGK; }

// One iteration of the fused path, optionally bracketing each pass with
// events (ev[0..3]) for per-kernel timing.
mcb_status fused_step(mcb_ctx* c, bool timed) {
    if (timed) MCB_HIP(c, hipEventRecord(c->ev[0], c->stream));
    MCB_HIP(c, launch_phase(c->g[c->cur], c->phi, c->geo, c->stream));
    if (timed) MCB_HIP(c, hipEventRecord(c->ev[1], c->stream));
    MCB_HIP(c, launch_collide(collide_args(c), step_mode(c), c->tiling, c->stream));
    if (timed) MCB_HIP(c, hipEventRecord(c->ev[2], c->stream));
    c->cur ^= 1;
    return MCB_OK;
}

#define MCB_OP(c, call)                                                   \
    do {                                                                  \
        mcb_status s_ = (call);                                           \
        if (s_ != MCB_OK) return fail((c), s_, std::string(#call) + " failed"); \
    } while (0)

mcb_status staged_macros(mcb_ctx* c) {
    const int NX = c->p.nx, NY = c->p.ny, NZ = c->p.nz;
    const mcb_op_consts* k = &c->oc;
    void* s = c->stream;
    MCB_OP(c, mcb_phiCalc(c->phi, c->sg, NX, NY, NZ, k, s));
    MCB_OP(c, mcb_gradCalc(c->phi, c->normx, c->normy, c->normz, c->indicator, NX, NY, NZ, k, s));
    MCB_OP(c, mcb_curvatureCalc(c->curvature, c->indicator, c->normx, c->normy, c->normz, c->ffx, c->ffy, c->ffz,
                                NX, NY, NZ, k, s));
    MCB_OP(c, mcb_momentiCalc(c->ux, c->uy, c->uz, c->rho, c->ffx, c->ffy, c->ffz, c->sf, c->pxx, c->pyy, c->pzz,
                              c->pxy, c->pxz, c->pyz, NX, NY, NZ, k, s));
    return MCB_OK;
}

// src/main.cu:79-178 in program order on one stream.
mcb_status staged_step(mcb_ctx* c) {
    const int NX = c->p.nx, NY = c->p.ny, NZ = c->p.nz;
    const mcb_op_consts* k = &c->oc;
    void* s = c->stream;
    mcb_status st = staged_macros(c);
    if (st != MCB_OK) return st;
    MCB_OP(c, mcb_collisionFluid(c->sf, c->ux, c->uy, c->uz, c->ffx, c->ffy, c->ffz, c->rho, c->pxx, c->pyy, c->pzz,
                                 c->pxy, c->pxz, c->pyz, NX, NY, NZ, k, s));
    MCB_OP(c, mcb_collisionPhase(c->sg, c->ux, c->uy, c->uz, c->phi, c->normx, c->normy, c->normz, NX, NY, NZ, k, s));
    if (c->p.bc_mode == MCB_BC_REFERENCE && c->p.moving_wall) MCB_OP(c, mcb_movingWallBC(c->sf, c->rho, NX, NY, NZ, k, s));
    MCB_OP(c, mcb_streamingCalc(c->sg, c->sg_out, NX, NY, NZ, k, s));
    std::swap(c->sg, c->sg_out);
    if (c->p.bc_mode == MCB_BC_REFERENCE) {
        MCB_OP(c, mcb_fgBoundary(c->sf, c->rho, c->sg, c->phi, NX, NY, NZ, k, s));
        MCB_OP(c, mcb_boundaryConditions(c->phi, NX, NY, NZ, k, s));
    }
    return MCB_OK;
}

mcb_status check_ready(mcb_ctx* c) {
    if (!c) return MCB_ERR_INVALID_ARG;
    if (!c->initialised) return fail(c, MCB_ERR_INVALID_ARG, "context has no initial state (call mcb_init_*)");
    MCB_HIP(c, hipSetDevice(c->device));
    return MCB_OK;
}

float* scalar_field(mcb_ctx* c, int field) {
    switch (field) {
        case MCB_FIELD_PHI: return c->phi;
        case MCB_FIELD_RHO: return c->rho;
        case MCB_FIELD_UX: return c->ux;
        case MCB_FIELD_UY: return c->uy;
        case MCB_FIELD_UZ: return c->uz;
        case MCB_FIELD_FX: return c->ffx;
        case MCB_FIELD_FY: return c->ffy;
        case MCB_FIELD_FZ: return c->ffz;
        case MCB_FIELD_NORMX: return c->normx;
        case MCB_FIELD_NORMY: return c->normy;
        case MCB_FIELD_NORMZ: return c->normz;
        default: return nullptr;
    }
}

}  // namespace

extern "C" {

const char* mcb_status_string(mcb_status s) {
    switch (s) {
        case MCB_OK: return "ok";
        case MCB_ERR_INVALID_ARG: return "invalid argument";
        case MCB_ERR_NO_DEVICE: return "no usable device";
        case MCB_ERR_OUT_OF_MEMORY: return "out of device memory";
        case MCB_ERR_HIP: return "HIP runtime error";
        case MCB_ERR_UNSUPPORTED: return "unsupported combination";
        case MCB_ERR_IO: return "I/O error";
        case MCB_ERR_NONFINITE: return "non-finite value detected";
        case MCB_ERR_CAPACITY: return "memory budget exceeded";
    }
    return "unknown status";
}

const char* mcb_last_error(const mcb_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

size_t mcb_footprint_bytes(const mcb_params* p) {
    if (!p || host_validate(p, nullptr) != MCB_OK) return 0;
    return footprint(*p, resolve_path(*p));
}

mcb_status mcb_get_params(const mcb_ctx* ctx, mcb_params* out) {
    if (!ctx || !out) return MCB_ERR_INVALID_ARG;
    *out = ctx->p;
    return MCB_OK;
}

mcb_status mcb_create(const mcb_params* p, void* stream, mcb_ctx** out) {
    if (!out) return MCB_ERR_INVALID_ARG;
    *out = nullptr;
    std::string why;
    mcb_status st = host_validate(p, &why);
    if (st != MCB_OK) return st;
    const int path = resolve_path(*p);
    if (path == MCB_PATH_FUSED && (p->bc_mode != MCB_BC_PERIODIC || !fused_fits(*p))) return MCB_ERR_UNSUPPORTED;
    if (p->storage != MCB_STORE_FP32) return MCB_ERR_UNSUPPORTED;

    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return MCB_ERR_NO_DEVICE;
    if (p->device < 0 || p->device >= count) return MCB_ERR_NO_DEVICE;

    mcb_ctx* c = new (std::nothrow) mcb_ctx();
    if (!c) return MCB_ERR_OUT_OF_MEMORY;
    c->p = *p;
    c->path = path;
    c->device = p->device;
    c->phys = host_phys(*p);
    mcb_op_consts_from_params(p, &c->oc);
    c->geo.nx = p->nx; c->geo.ny = p->ny; c->geo.nz = p->nz;
    c->geo.nxy = p->nx * p->ny;
    c->geo.n = (long long)p->nx * p->ny * p->nz;

    auto bail = [&](mcb_status s) { mcb_destroy(c); return s; };
    if (hipSetDevice(c->device) != hipSuccess) return bail(MCB_ERR_NO_DEVICE);
    if (stream) {
        c->stream = static_cast<hipStream_t>(stream);
    } else {
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return bail(MCB_ERR_HIP);
        c->own_stream = true;
    }
    for (auto& e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) return bail(MCB_ERR_HIP);

    const long long n = c->geo.n;
    const long long P = plane_stride(n);
    if (path == MCB_PATH_FUSED) {
        c->geo.plane = P;
        for (int b = 0; b < 2; ++b) {
            if ((st = alloc(c, &c->f[b], 19 * (size_t)P)) != MCB_OK) return bail(st);
            if ((st = alloc(c, &c->g[b], 19 * (size_t)P)) != MCB_OK) return bail(st);
        }
        float** scal[] = {&c->phi, &c->rho, &c->ux, &c->uy, &c->uz, &c->ffx, &c->ffy, &c->ffz,
                          &c->normx, &c->normy, &c->normz};
        for (auto* q : scal)
            if ((st = alloc(c, q, (size_t)P)) != MCB_OK) return bail(st);
        // Default tiling: widest x tile the lattice fills.
        c->tiling = p->nx >= 64 ? Tiling{64, 8, 16} : (p->nx >= 32 ? Tiling{32, 8, 16} : Tiling{16, 16, 16});
    } else {
        // The reference layout: planes of exactly n cells (src/header.cuh:21).
        c->geo.plane = n;
        if ((st = alloc(c, &c->sf, 19 * (size_t)n)) != MCB_OK) return bail(st);
        if ((st = alloc(c, &c->sg, 19 * (size_t)n)) != MCB_OK) return bail(st);
        if ((st = alloc(c, &c->sg_out, 19 * (size_t)n)) != MCB_OK) return bail(st);
        float** scal[] = {&c->phi, &c->rho, &c->ux, &c->uy, &c->uz, &c->ffx, &c->ffy, &c->ffz, &c->normx,
                          &c->normy, &c->normz, &c->indicator, &c->curvature, &c->pxx, &c->pyy, &c->pzz,
                          &c->pxy, &c->pxz, &c->pyz};
        for (auto* q : scal)
            if ((st = alloc(c, q, (size_t)n)) != MCB_OK) return bail(st);
    }
    void* dp = nullptr;
    if (hipMalloc(&dp, (2 * kReduceBlocks + 2) * sizeof(double)) != hipSuccess) return bail(MCB_ERR_OUT_OF_MEMORY);
    c->allocs.push_back(dp);
    c->partial = static_cast<double*>(dp);
    c->dsum = c->partial + 2 * kReduceBlocks;
    *out = c;
    return MCB_OK;
}

mcb_status mcb_destroy(mcb_ctx* c) {
    if (!c) return MCB_ERR_INVALID_ARG;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (void* q : c->allocs) (void)hipFree(q);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : c->tev) (void)hipEventDestroy(e);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return MCB_OK;
}

mcb_status mcb_init_bubbles(mcb_ctx* c, int n, const mcb_bubble* bubbles, float width) {
    if (!c) return MCB_ERR_INVALID_ARG;
    std::vector<float> phi((size_t)c->geo.n);
    mcb_status st = mcb_host_init_phi(&c->p, n, bubbles, width, phi.data(), phi.size());
    if (st != MCB_OK) return fail(c, st, "invalid bubble set");
    return mcb_init_phi(c, phi.data(), phi.size());
}

mcb_status mcb_init_phi(mcb_ctx* c, const float* phi_host, size_t n) {
    if (!c || !phi_host || n != (size_t)c->geo.n) return fail(c, MCB_ERR_INVALID_ARG, "phi size mismatch");
    MCB_HIP(c, hipSetDevice(c->device));
    const size_t bytes = (size_t)c->geo.n * sizeof(float);
    const int NX = c->p.nx, NY = c->p.ny, NZ = c->p.nz;
    MCB_HIP(c, hipMemcpyAsync(c->phi, phi_host, bytes, hipMemcpyHostToDevice, c->stream));
    if (c->path == MCB_PATH_FUSED) {
        c->cur = 0;
        const size_t pbytes = (size_t)c->geo.plane * sizeof(float);
        float* zero[] = {c->rho, c->ux, c->uy, c->uz, c->ffx, c->ffy, c->ffz, c->normx, c->normy, c->normz};
        for (float* q : zero) MCB_HIP(c, hipMemsetAsync(q, 0, pbytes, c->stream));
        MCB_HIP(c, launch_init_pull(c->phi, c->f[0], c->g[0], c->geo, c->phys, c->stream));
    } else {
        // src/var.cu:86-101 clears these; initTensor/initDist set the rest.
        float* zero[] = {c->ux, c->uy, c->uz, c->normx, c->normy, c->normz, c->curvature, c->indicator,
                         c->ffx, c->ffy, c->ffz};
        for (float* q : zero) MCB_HIP(c, hipMemsetAsync(q, 0, bytes, c->stream));
        MCB_HIP(c, hipMemsetAsync(c->sg_out, 0, 19 * bytes, c->stream));
        MCB_OP(c, mcb_initTensor(c->pxx, c->pyy, c->pzz, c->pxy, c->pxz, c->pyz, c->rho, NX, NY, NZ, &c->oc, c->stream));
        MCB_OP(c, mcb_initDist(c->rho, c->phi, c->sf, c->sg, NX, NY, NZ, &c->oc, c->stream));
    }
    MCB_HIP(c, hipStreamSynchronize(c->stream));
    c->initialised = true;
    return MCB_OK;
}

mcb_status mcb_step(mcb_ctx* c, int nsteps) {
    mcb_status st = check_ready(c);
    if (st != MCB_OK) return st;
    if (nsteps < 0) return fail(c, MCB_ERR_INVALID_ARG, "negative step count");
    for (int t = 0; t < nsteps; ++t) {
        st = c->path == MCB_PATH_FUSED ? fused_step(c, false) : staged_step(c);
        if (st != MCB_OK) return st;
    }
    return MCB_OK;
}

mcb_status mcb_step_timed(mcb_ctx* c, int nsteps, double* ms) {
    mcb_status st = check_ready(c);
    if (st != MCB_OK) return st;
    if (nsteps < 0 || !ms) return fail(c, MCB_ERR_INVALID_ARG, "bad arguments");
    MCB_HIP(c, hipEventRecord(c->ev[0], c->stream));
    for (int t = 0; t < nsteps; ++t) {
        st = c->path == MCB_PATH_FUSED ? fused_step(c, false) : staged_step(c);
        if (st != MCB_OK) return st;
    }
    MCB_HIP(c, hipEventRecord(c->ev[3], c->stream));
    MCB_HIP(c, hipEventSynchronize(c->ev[3]));
    float t = 0.0f;
    MCB_HIP(c, hipEventElapsedTime(&t, c->ev[0], c->ev[3]));
    *ms = t;
    return MCB_OK;
}

mcb_status mcb_compute_macros(mcb_ctx* c) {
    mcb_status st = check_ready(c);
    if (st != MCB_OK) return st;
    if (c->path == MCB_PATH_FUSED) {
        MCB_HIP(c, launch_phase(c->g[c->cur], c->phi, c->geo, c->stream));
        MCB_HIP(c, launch_collide(collide_args(c), COLLIDE_MACROS, c->tiling, c->stream));
        return MCB_OK;
    }
    return staged_macros(c);
}

mcb_status mcb_read_field(mcb_ctx* c, int field, float* host, size_t n) {
    mcb_status st = check_ready(c);
    if (st != MCB_OK) return st;
    if (!host) return fail(c, MCB_ERR_INVALID_ARG, "null host buffer");
    const size_t cells = (size_t)c->geo.n;
    if (field == MCB_FIELD_F || field == MCB_FIELD_G) {
        if (n != 19 * cells) return fail(c, MCB_ERR_INVALID_ARG, "population buffer must hold 19*nx*ny*nz floats");
        if (c->path == MCB_PATH_STAGED) {
            const float* src = field == MCB_FIELD_F ? c->sf : c->sg;
            MCB_HIP(c, hipMemcpyAsync(host, src, 19 * cells * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        } else {
            // Gather into the idle A/B buffer, then copy plane by plane.
            const float* src = field == MCB_FIELD_F ? c->f[c->cur] : c->g[c->cur];
            float* tmp = field == MCB_FIELD_F ? c->f[c->cur ^ 1] : c->g[c->cur ^ 1];
            MCB_HIP(c, launch_link_shift(src, tmp, c->geo, +1, c->stream));
            for (int l = 0; l < 19; ++l)
                MCB_HIP(c, hipMemcpyAsync(host + l * cells, tmp + (size_t)l * c->geo.plane, cells * sizeof(float),
                                          hipMemcpyDeviceToHost, c->stream));
        }
    } else {
        const float* src = scalar_field(c, field);
        if (!src) return fail(c, MCB_ERR_INVALID_ARG, "unknown field");
        if (n != cells) return fail(c, MCB_ERR_INVALID_ARG, "field buffer must hold nx*ny*nz floats");
        MCB_HIP(c, hipMemcpyAsync(host, src, cells * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    }
    MCB_HIP(c, hipStreamSynchronize(c->stream));
    return MCB_OK;
}

mcb_status mcb_write_field(mcb_ctx* c, int field, const float* host, size_t n) {
    mcb_status st = check_ready(c);
    if (st != MCB_OK) return st;
    if (!host) return fail(c, MCB_ERR_INVALID_ARG, "null host buffer");
    const size_t cells = (size_t)c->geo.n;
    if (field == MCB_FIELD_F || field == MCB_FIELD_G) {
        if (n != 19 * cells) return fail(c, MCB_ERR_INVALID_ARG, "population buffer must hold 19*nx*ny*nz floats");
        if (c->path == MCB_PATH_STAGED) {
            float* dst = field == MCB_FIELD_F ? c->sf : c->sg;
            MCB_HIP(c, hipMemcpyAsync(dst, host, 19 * cells * sizeof(float), hipMemcpyHostToDevice, c->stream));
        } else {
            float* dst = field == MCB_FIELD_F ? c->f[c->cur] : c->g[c->cur];
            float* tmp = field == MCB_FIELD_F ? c->f[c->cur ^ 1] : c->g[c->cur ^ 1];
            for (int l = 0; l < 19; ++l)
                MCB_HIP(c, hipMemcpyAsync(tmp + (size_t)l * c->geo.plane, host + l * cells, cells * sizeof(float),
                                          hipMemcpyHostToDevice, c->stream));
            MCB_HIP(c, launch_link_shift(tmp, dst, c->geo, -1, c->stream));
        }
    } else {
        float* dst = scalar_field(c, field);
        if (!dst) return fail(c, MCB_ERR_INVALID_ARG, "unknown field");
        if (n != cells) return fail(c, MCB_ERR_INVALID_ARG, "field buffer must hold nx*ny*nz floats");
        MCB_HIP(c, hipMemcpyAsync(dst, host, cells * sizeof(float), hipMemcpyHostToDevice, c->stream));
    }
    MCB_HIP(c, hipStreamSynchronize(c->stream));
    return MCB_OK;
}

mcb_status mcb_reduce(mcb_ctx* c, double* sum_phi, double* kinetic) {
    mcb_status st = check_ready(c);
    if (st != MCB_OK) return st;
    MCB_HIP(c, launch_reduce(c->phi, c->rho, c->ux, c->uy, c->uz, c->geo.n, c->partial, c->dsum, c->stream));
    double h[2] = {0.0, 0.0};
    MCB_HIP(c, hipMemcpyAsync(h, c->dsum, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    MCB_HIP(c, hipStreamSynchronize(c->stream));
    if (sum_phi) *sum_phi = h[0];
    if (kinetic) *kinetic = h[1];
    if (!std::isfinite(h[0]) || !std::isfinite(h[1])) return fail(c, MCB_ERR_NONFINITE, "non-finite diagnostic");
    return MCB_OK;
}

mcb_status mcb_synchronize(mcb_ctx* c) {
    if (!c) return MCB_ERR_INVALID_ARG;
    MCB_HIP(c, hipSetDevice(c->device));
    MCB_HIP(c, hipStreamSynchronize(c->stream));
    return MCB_OK;
}

// ---- per-kernel timing and tiling control (used by bench.py and the tuner)

// Device time of each pass summed over nsteps fused steps:
// out_ms[0] = k_phase, out_ms[1] = k_collide.
mcb_status mcb_step_kernel_times(mcb_ctx* c, int nsteps, double* out_ms, int n_out) {
    mcb_status st = check_ready(c);
    if (st != MCB_OK) return st;
    if (c->path != MCB_PATH_FUSED) return fail(c, MCB_ERR_UNSUPPORTED, "kernel times need the fused path");
    if (!out_ms || n_out < 2 || nsteps < 1) return fail(c, MCB_ERR_INVALID_ARG, "bad arguments");
    // Three events per step, recorded back to back on the stream and read
    // once at the end, so the timed steps run without host round trips.
    const size_t need = 3 * (size_t)nsteps;
    while (c->tev.size() < need) {
        hipEvent_t e = nullptr;
        MCB_HIP(c, hipEventCreate(&e));
        c->tev.push_back(e);
    }
    for (int t = 0; t < nsteps; ++t) {
        hipEvent_t* e = &c->tev[3 * (size_t)t];
        MCB_HIP(c, hipEventRecord(e[0], c->stream));
        MCB_HIP(c, launch_phase(c->g[c->cur], c->phi, c->geo, c->stream));
        MCB_HIP(c, hipEventRecord(e[1], c->stream));
        MCB_HIP(c, launch_collide(collide_args(c), step_mode(c), c->tiling, c->stream));
        MCB_HIP(c, hipEventRecord(e[2], c->stream));
        c->cur ^= 1;
    }
    MCB_HIP(c, hipEventSynchronize(c->tev[need - 1]));
    double acc[2] = {0.0, 0.0};
    for (int t = 0; t < nsteps; ++t) {
        hipEvent_t* e = &c->tev[3 * (size_t)t];
        float a = 0.0f, b = 0.0f;
        MCB_HIP(c, hipEventElapsedTime(&a, e[0], e[1]));
        MCB_HIP(c, hipEventElapsedTime(&b, e[1], e[2]));
        acc[0] += a;
        acc[1] += b;
    }
    out_ms[0] = acc[0];
    out_ms[1] = acc[1];
    return MCB_OK;
}

mcb_status mcb_set_tiling(mcb_ctx* c, int tx, int ty, int zc) {
    if (!c) return MCB_ERR_INVALID_ARG;
    Tiling t{tx, ty, zc};
    if (!tiling_supported(t) || zc > c->p.nz) return fail(c, MCB_ERR_INVALID_ARG, "unsupported tiling");
    c->tiling = t;
    return MCB_OK;
}

mcb_status mcb_get_tiling(const mcb_ctx* c, int* tx, int* ty, int* zc) {
    if (!c || !tx || !ty || !zc) return MCB_ERR_INVALID_ARG;
    *tx = c->tiling.tx;
    *ty = c->tiling.ty;
    *zc = c->tiling.zc;
    return MCB_OK;
}

int mcb_tiling_count(void) { return tiling_count(); }

mcb_status mcb_tiling_at(int i, int* tx, int* ty) {
    if (i < 0 || i >= tiling_count() || !tx || !ty) return MCB_ERR_INVALID_ARG;
    Tiling t = tiling_at(i);
    *tx = t.tx;
    *ty = t.ty;
    return MCB_OK;
}

// ---- device helpers

mcb_status mcb_device_count(int* count) {
    if (!count) return MCB_ERR_INVALID_ARG;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    *count = e == hipSuccess ? n : 0;
    return e == hipSuccess ? MCB_OK : MCB_ERR_NO_DEVICE;
}

mcb_status mcb_device_alloc(int device, size_t bytes, void** dptr) {
    if (!dptr || bytes == 0) return MCB_ERR_INVALID_ARG;
    if (hipSetDevice(device) != hipSuccess) return MCB_ERR_NO_DEVICE;
    return hip_status(hipMalloc(dptr, bytes));
}

mcb_status mcb_device_free(void* dptr) { return hip_status(hipFree(dptr)); }

mcb_status mcb_memcpy_htod(void* dst, const void* src, size_t bytes) {
    if (!dst || !src) return MCB_ERR_INVALID_ARG;
    return hip_status(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
}

mcb_status mcb_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
    if (!dst || !src) return MCB_ERR_INVALID_ARG;
    return hip_status(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
}

mcb_status mcb_device_synchronize(void) { return hip_status(hipDeviceSynchronize()); }

mcb_status mcb_bench_copy(int device, size_t bytes, int iters, double* ms_per_copy) {
    if (!ms_per_copy || bytes < 16 || iters < 1) return MCB_ERR_INVALID_ARG;
    bytes &= ~(size_t)15;
    if (hipSetDevice(device) != hipSuccess) return MCB_ERR_NO_DEVICE;
    void *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    mcb_status st = MCB_OK;
    auto chk = [&](hipError_t e) {
        if (e != hipSuccess && st == MCB_OK) st = hip_status(e);
        return st == MCB_OK;
    };
    if (chk(hipMalloc(&a, bytes)) && chk(hipMalloc(&b, bytes)) && chk(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) &&
        chk(hipEventCreate(&e0)) && chk(hipEventCreate(&e1)) && chk(hipMemsetAsync(a, 0, bytes, s)) &&
        chk(launch_copy(a, b, bytes, s)) && chk(hipEventRecord(e0, s))) {
        for (int i = 0; i < iters && st == MCB_OK; ++i) chk(launch_copy(a, b, bytes, s));
        float t = 0.0f;
        if (chk(hipEventRecord(e1, s)) && chk(hipEventSynchronize(e1)) && chk(hipEventElapsedTime(&t, e0, e1)))
            *ms_per_copy = (double)t / iters;
    }
    if (s) (void)hipStreamSynchronize(s);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    return st;
}

}  // extern "C"
