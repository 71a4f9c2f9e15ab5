// Context and batch API of libmcb (include/mcb.h): device memory, initial
// state, the time step on one HIP stream, readout and diagnostics.
//
// A context runs one of three engines, all sharing the per-cell arithmetic of
// lattice.hpp:
//   FUSED FP32 : k_phase + k_collide per step, pull storage, A/B buffers
//                (fused.hip). Bit-identical to the FP32 transcription.
//   16-bit     : FP16/BF16 population storage; the case runs as a batch of
//                one (batch.hip, fused16.hip).
//   STAGED     : the reference's kernel sequence of src/main.cu:79-178 on the
//                reference's buffers (ops.hip), one stream, no host syncs.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "batch.hpp"
#include "state.hpp"
#include "common.hpp"
#include "fused.hpp"

using namespace mcb;

namespace mcb {
mcb_status host_validate(const mcb_params* p, std::string* why);
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
    Tiling tiling{64, 8, 16};
    std::string err;
    std::vector<void*> allocs;

    // FUSED FP32 state
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
    // 16-bit storage: a batch of one case
    std::unique_ptr<Batch> batch;
    int case_id = 0;
    bool initialised = false;
    // FUSED FP32 at tau_g = 1: g in moment form (GMode), phi double-buffered
    // (phis[cur] is written by the next phase pass, phis[cur ^ 1] is the state).
    int gm = GM_POPULATIONS;
    float* phis[2] = {nullptr, nullptr};
    float* gmu[3] = {nullptr, nullptr, nullptr};
    float* gmn[3] = {nullptr, nullptr, nullptr};   // null: the phase pass recomputes n
    // f in moment form (MCB_F_MOMENTS): f[0], f[1] hold 13 planes each, u is
    // their planes 1..3, and F/G readouts gather into fscratch (19 planes,
    // allocated on first use).
    bool fm = false;
    float* fscratch = nullptr;
    long long steps = 0;   // steps since the initial state (non-batch paths)
    // Fused path with REFERENCE boundaries: a written F put arbitrary values
    // into the face populations of f[ref_reset_f], which the next step reads
    // once; after it they are reset to w_l (-1: nothing pending).
    int ref_reset_f = -1;
    // Fused path: two steps captured in a hipGraph per A/B parity, with the
    // launch arguments they were captured with; a graph is replayed only
    // while the next two steps would launch exactly those (same views, tile
    // and collide mode), and recaptured otherwise.
    struct FusedGraph {
        hipGraphExec_t exec = nullptr;
        CaseView v[2];
        Tiling t{};
        int mode = -1;
    } fgraph[2];
    bool graphs_ok = true;
};

struct mcb_batch {
    std::unique_ptr<Batch> b;
    int device = 0;
    int storage = MCB_STORE_FP32;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    hipEvent_t ev[2] = {nullptr, nullptr};
    std::string err;
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

// boundaryConditions (src/lbm.cu:403-425, faces in the order z, y, x) on a
// host phi field: every face cell takes the phi of the interior cell at its
// clamped coordinates.
void reference_faces_from_interior(const mcb_params& p, float* phi) {
    const int NX = p.nx, NY = p.ny, NZ = p.nz;
    auto idx = [&](int i, int j, int k) { return (size_t)i + (size_t)NX * ((size_t)j + (size_t)NY * k); };
    auto cl = [](int a, int n) { return a < 1 ? 1 : (a > n - 2 ? n - 2 : a); };
    for (int k = 0; k < NZ; ++k)
        for (int j = 0; j < NY; ++j) {
            const bool edge = k == 0 || k == NZ - 1 || j == 0 || j == NY - 1;
            for (int i = 0; i < NX; i += (edge || i == NX - 1) ? 1 : NX - 1)
                phi[idx(i, j, k)] = phi[idx(cl(i, NX), cl(j, NY), cl(k, NZ))];
        }
}

// Plane stride: cells rounded up to 256 elements, so every population
// plane starts 1 KiB-aligned.
long long plane_stride(long long n) { return plane_stride_cells(n); }

// The fused kernels address all 19 planes of a buffer through one 32-bit
// buffer resource (FP32 lattices up to ~56.5M cells, 16-bit up to ~113M).
bool fused_fits(const mcb_params& p) {
    const long long P = plane_stride((long long)p.nx * p.ny * p.nz);
    return P * 19 * (p.storage == MCB_STORE_FP32 ? 4 : 2) <= 0xFFFFFFFFLL;
}

// The fused kernels take the reference's ghost-layer boundaries in FP32 with g
// in moment form (tau_g = 1, the reference's own phase collision) and without
// the dead movingWallBC; other REFERENCE cases run the staged kernels.
bool fused_reference_ok(const mcb_params& p) {
    return p.bc_mode == MCB_BC_REFERENCE && p.storage == MCB_STORE_FP32 && p.tau_g == 1.0f && !p.moving_wall;
}

int resolve_path(const mcb_params& p) {
    if (p.path == MCB_PATH_STAGED) return MCB_PATH_STAGED;
    if (p.path == MCB_PATH_FUSED) return MCB_PATH_FUSED;
    if (p.storage != MCB_STORE_FP32 || p.f_form == MCB_F_MOMENTS) return MCB_PATH_FUSED;
    return (p.bc_mode == MCB_BC_PERIODIC || fused_reference_ok(p)) && fused_fits(p) ? MCB_PATH_FUSED
                                                                                    : MCB_PATH_STAGED;
}

// Whether a single-case FP32 moment-form context recomputes n in the phase
// pass: as asked, or (AUTO) where its tiling gives the pass long z-chunks.
bool ctx_recomputes_n(const mcb_params& p) {
    if (p.normals != MCB_NORMALS_AUTO) return p.normals == MCB_NORMALS_RECOMPUTED;
    const Geo g{p.nx, p.ny, p.nz, p.nx * p.ny, (long long)p.nx * p.ny * p.nz, 0};
    return phase_recomputes_n(g, choose_tiling_single(p));
}

size_t footprint(const mcb_params& p, int path) {
    // 16-bit: the case plus the batch's ten macro fields.
    if (p.storage != MCB_STORE_FP32) return case_bytes(p) + 10 * (size_t)plane_stride((long long)p.nx * p.ny * p.nz) * 4;
    const long long n = (long long)p.nx * p.ny * p.nz;
    const long long P = plane_stride(n);
    size_t floats = 0;
    if (path == MCB_PATH_FUSED && p.f_form == MCB_F_MOMENTS)   // 13 moments A/B (u among them), phi twice, n
        floats = 2 * 13 * (size_t)P + (ctx_recomputes_n(p) ? 12 : 15) * (size_t)P;
    else if (path == MCB_PATH_FUSED && p.tau_g == 1.0f)   // g in moment form: phi twice, u, and n unless recomputed
        floats = 2 * 19 * (size_t)P + (ctx_recomputes_n(p) ? 15 : 18) * (size_t)P;
    else if (path == MCB_PATH_FUSED) floats = 4 * 19 * (size_t)P + 11 * (size_t)P;
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

CaseView fused_view(mcb_ctx* c) {
    CaseView v{};
    v.geo = c->geo;
    v.phys = c->phys;
    v.zc = c->tiling.zc;
    v.fsrc = c->f[c->cur];
    v.fdst = c->f[c->cur ^ 1];
    v.gsrc = c->g[c->cur];
    v.gdst = c->g[c->cur ^ 1];
    v.phi = c->phi;
    v.gm = c->gm;
    v.bc = c->p.bc_mode;
    if (c->gm != GM_POPULATIONS) {
        v.gsrc = v.gdst = nullptr;
        v.phi = c->phis[c->cur];
        v.phi_prev = c->phis[c->cur ^ 1];
        for (int k = 0; k < 3; ++k) {
            v.gm_u[k] = c->fm ? c->f[c->cur] + (size_t)(1 + k) * c->geo.plane : c->gmu[k];
            v.gm_n[k] = c->gmn[k];
        }
    }
    v.fm = c->fm ? 1 : 0;
    v.rho = c->rho; v.ux = c->ux; v.uy = c->uy; v.uz = c->uz;
    v.fx = c->ffx; v.fy = c->ffy; v.fz = c->ffz;
    v.nx = c->normx; v.ny = c->normy; v.nz = c->normz;
    return v;
}

int step_mode(const mcb_ctx* c) { return c->p.tau_g == 1.0f ? COLLIDE_STEP : COLLIDE_STEP_BGK; }

// Fused path, before the initial phi is written: A/B parity 0, and in moment
// form phi0 goes to the state buffer phis[1].
void fused_init_target(mcb_ctx* c) {
    c->cur = 0;
    c->ref_reset_f = -1;
    if (c->g[0] == nullptr) {
        c->gm = GM_INITIAL;
        c->phi = c->phis[1];
    }
}

// Fused path, after the initial phi is in c->phi: f*_l = w_l, and g* =
// w_l phi0(y + c_l) in pull form or nothing (initDist is implied by
// GM_INITIAL) in moment form; the macro fields start at zero.
hipError_t fused_init_state(mcb_ctx* c) {
    const size_t pbytes = (size_t)c->geo.plane * sizeof(float);
    float* zero[] = {c->rho, c->ux, c->uy, c->uz, c->ffx, c->ffy, c->ffz, c->normx, c->normy, c->normz};
    for (float* q : zero) {
        hipError_t e = hipMemsetAsync(q, 0, pbytes, c->stream);
        if (e != hipSuccess) return e;
    }
    if (c->fm) return launch_fm_init(c->f[0], c->geo, c->stream);
    return launch_init_pull(c->phi, c->f[0], c->g[0], c->geo, c->phys, c->stream);
}

// After a fused step: flip the A/B buffers; in moment form the phi just
// written becomes the state's phi.
hipError_t after_fused_step(mcb_ctx* c) {
    hipError_t e = hipSuccess;
    if (c->ref_reset_f == c->cur) {
        // The step just read the written faces; the next one writes this
        // buffer's interior, so w_l everywhere restores the face invariant.
        e = launch_init_pull(c->phi, c->f[c->cur], nullptr, c->geo, c->phys, c->stream);
        c->ref_reset_f = -1;
    }
    c->cur ^= 1;
    if (c->gm != GM_POPULATIONS) {
        c->gm = GM_MOMENTS;
        c->phi = c->phis[c->cur ^ 1];
    }
    return e;
}

mcb_status fused_step(mcb_ctx* c) {
    const CaseView v = fused_view(c);
    MCB_HIP(c, launch_phase(v, c->tiling, c->stream));
    MCB_HIP(c, launch_collide(v, step_mode(c), c->tiling, c->stream));
    MCB_HIP(c, after_fused_step(c));
    return MCB_OK;
}

// The views of the next two fused steps, without advancing the context.
void next_two_views(mcb_ctx* c, CaseView (&v)[2]) {
    const int cur = c->cur, gm = c->gm;
    float* const phi = c->phi;
    v[0] = fused_view(c);
    c->cur ^= 1;
    if (c->gm != GM_POPULATIONS) {
        c->gm = GM_MOMENTS;
        c->phi = c->phis[c->cur ^ 1];
    }
    v[1] = fused_view(c);
    c->cur = cur;
    c->gm = gm;
    c->phi = phi;
}

void drop_fused_graphs(mcb_ctx* c) {
    for (auto& g : c->fgraph) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g.exec = nullptr;
    }
}

// Two fused steps from a hipGraph (captured on first use for this parity and
// these launch arguments); *done is false, and nothing is launched, when the
// runtime cannot capture, after which the context steps eagerly.
mcb_status fused_two_steps_graph(mcb_ctx* c, bool* done) {
    *done = false;
    if (!c->graphs_ok) return MCB_OK;
    // A caller capturing the context's stream into a graph of its own gets
    // the eager launches recorded there.
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    MCB_HIP(c, hipStreamIsCapturing(c->stream, &cap));
    if (cap != hipStreamCaptureStatusNone) return MCB_OK;
    CaseView v[2];
    next_two_views(c, v);
    auto& g = c->fgraph[c->cur];
    const int mode = step_mode(c);
    const bool same = g.exec && g.mode == mode && g.t.tx == c->tiling.tx && g.t.ty == c->tiling.ty &&
                      g.t.zc == c->tiling.zc && std::memcmp(g.v, v, sizeof(v)) == 0;
    if (!same) {
        if (g.exec) MCB_HIP(c, hipGraphExecDestroy(g.exec));
        g.exec = nullptr;
        hipGraph_t gr = nullptr;
        if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            (void)hipGetLastError();
            c->graphs_ok = false;
            return MCB_OK;
        }
        hipError_t e = hipSuccess;
        for (int k = 0; k < 2 && e == hipSuccess; ++k) {
            e = launch_phase(v[k], c->tiling, c->stream);
            if (e == hipSuccess) e = launch_collide(v[k], mode, c->tiling, c->stream);
        }
        const hipError_t ee = hipStreamEndCapture(c->stream, &gr);
        if (e == hipSuccess && ee == hipSuccess) e = hipGraphInstantiate(&g.exec, gr, nullptr, nullptr, 0);
        if (gr) (void)hipGraphDestroy(gr);
        if (e != hipSuccess || ee != hipSuccess) {
            (void)hipGetLastError();
            g.exec = nullptr;
            c->graphs_ok = false;
            return MCB_OK;
        }
        std::memcpy(g.v, v, sizeof(v));
        g.t = c->tiling;
        g.mode = mode;
    }
    MCB_HIP(c, hipGraphLaunch(g.exec, c->stream));
    for (int k = 0; k < 2; ++k) MCB_HIP(c, after_fused_step(c));
    *done = true;
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

mcb_status batch_status(mcb_ctx* c, mcb_status st) {
    if (st != MCB_OK) c->err = c->batch->error();
    return st;
}

mcb_status check_ready(mcb_ctx* c) {
    if (!c) return MCB_ERR_INVALID_ARG;
    if (!c->initialised) return fail(c, MCB_ERR_INVALID_ARG, "context has no initial state (call mcb_init_*)");
    MCB_HIP(c, hipSetDevice(c->device));
    return MCB_OK;
}

mcb_status any_step(mcb_ctx* c, int nsteps) {
    if (c->batch) return batch_status(c, c->batch->step(nsteps));
    for (int t = 0; t < nsteps;) {
        // Fused path: pairs of steps in the steady state (g in moment form
        // past its initial state, or as populations; no face reset pending)
        // replay a captured graph.
        if (c->path == MCB_PATH_FUSED && nsteps - t >= 2 && c->gm != GM_INITIAL && c->ref_reset_f < 0) {
            bool done = false;
            mcb_status st = fused_two_steps_graph(c, &done);
            if (st != MCB_OK) return st;
            if (done) {
                t += 2;
                c->steps += 2;
                continue;
            }
        }
        mcb_status st = c->path == MCB_PATH_FUSED ? fused_step(c) : staged_step(c);
        if (st != MCB_OK) return st;
        ++c->steps;
        ++t;
    }
    return MCB_OK;
}

// What the next step of a non-batch context reads: staged path, every
// reference variable (src/var.cu:13-18): with ghost-layer boundaries the
// kernels leave the face values of phi and the macroscopic fields from one
// iteration to the next, so they are state too; fused FP32, f of the current
// buffer and g as populations (with the phi the last step reported) or in
// moment form (the state's phi and u; n is gradCalc of that phi, rebuilt on
// loading where the context stores it).
std::vector<StatePart> ctx_state_parts(mcb_ctx* c) {
    const size_t N = (size_t)c->geo.n, P = (size_t)c->geo.plane;
    if (c->path == MCB_PATH_STAGED) {
        std::vector<StatePart> parts{{"f", c->sf, 19, N, N, 4, false}, {"g", c->sg, 19, N, N, 4, false}};
        const std::pair<const char*, float*> fields[] = {
            {"phi", c->phi},     {"rho", c->rho},       {"ux", c->ux},   {"uy", c->uy},   {"uz", c->uz},
            {"fx", c->ffx},      {"fy", c->ffy},        {"fz", c->ffz},  {"nx", c->normx}, {"ny", c->normy},
            {"nz", c->normz},    {"ind", c->indicator}, {"curv", c->curvature}, {"pxx", c->pxx}, {"pyy", c->pyy},
            {"pzz", c->pzz},     {"pxy", c->pxy},       {"pxz", c->pxz}, {"pyz", c->pyz}};
        for (const auto& f : fields) parts.push_back({f.first, f.second, 1, N, N, 4, false});
        return parts;
    }
    if (c->fm)   // the 13 moments (u among them) and the state's phi
        return {{"fm", c->f[c->cur], kFm, N, P, 4, false}, {"phi", c->phis[c->cur ^ 1], 1, N, N, 4, false}};
    // Images keep populations in planes: with blocked populations they pass
    // through the idle A/B buffers (pops_image_copy / pops_image_done).
    const int side = c->geo.blocked ? c->cur ^ 1 : c->cur;
    std::vector<StatePart> parts{{"f", c->f[side], 19, N, P, 4, false}};
    if (c->gm == GM_POPULATIONS) {
        // With g as populations phi is what the last step reported (phiCalc
        // of the populations before that step's collision): not given back
        // by the state's populations, so kept for a read after loading.
        parts.push_back({"g", c->g[side], 19, N, P, 4, false});
        parts.push_back({"phi", c->phi, 1, N, N, 4, false});
    } else {
        static const char* const un[3] = {"u0", "u1", "u2"};
        parts.push_back({"phi", c->phis[c->cur ^ 1], 1, N, N, 4, false});
        for (int k = 0; k < 3; ++k) parts.push_back({un[k], c->gmu[k], 1, N, N, 4, false});
    }
    return parts;
}

// Blocked populations around a state image (fused path): the current f (and
// g populations) into the idle buffers in planes, or back. With ghost-layer
// faces the idle f's face populations then go back to w_l (a step reads the
// current buffer and writes every interior cell of the idle one).
hipError_t pops_image_copy(mcb_ctx* c, bool to_image) {
    const int cur = c->cur, idle = c->cur ^ 1;
    float* const* bufs[2] = {c->f, c->gm == GM_POPULATIONS ? c->g : nullptr};
    for (float* const* b : bufs) {
        if (!b) continue;
        const hipError_t e = to_image ? launch_pop_relayout(b[cur], 1, b[idle], 0, c->geo, 4, c->stream)
                                      : launch_pop_relayout(b[idle], 0, b[cur], 1, c->geo, 4, c->stream);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}
hipError_t pops_image_done(mcb_ctx* c) {
    if (c->p.bc_mode != MCB_BC_REFERENCE) return hipSuccess;
    return launch_init_pull(c->phi, c->f[c->cur ^ 1], nullptr, c->geo, c->phys, c->stream);
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
    if (path == MCB_PATH_FUSED && ((p->bc_mode != MCB_BC_PERIODIC && !fused_reference_ok(*p)) || !fused_fits(*p)))
        return MCB_ERR_UNSUPPORTED;
    if (p->storage != MCB_STORE_FP32 && path != MCB_PATH_FUSED) return MCB_ERR_UNSUPPORTED;
    if (p->f_form == MCB_F_MOMENTS && (path != MCB_PATH_FUSED || p->storage != MCB_STORE_FP32 || p->tau_g != 1.0f ||
                                       p->bc_mode != MCB_BC_PERIODIC))
        return MCB_ERR_UNSUPPORTED;

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
    if (p->storage != MCB_STORE_FP32) {
        c->geo.plane = P;
        c->batch.reset(new (std::nothrow) Batch(c->device, c->stream, true));
        if (!c->batch) return bail(MCB_ERR_OUT_OF_MEMORY);
        c->tiling = choose_tiling(*p);
    } else if (path == MCB_PATH_FUSED) {
        c->geo.plane = P;
        // At tau_g = 1 g is kept in moment form: phi (two buffers), u and n
        // (or n recomputed by the phase pass) instead of two sets of 19 planes.
        const bool moments = p->tau_g == 1.0f;
        c->fm = p->f_form == MCB_F_MOMENTS;
        // Populations in 64-cell blocks where a z plane holds whole blocks.
        c->geo.blocked = MCB_POP_BLOCKS && !c->fm && c->geo.nxy % 64 == 0;
        for (int b = 0; b < 2; ++b) {
            if ((st = alloc(c, &c->f[b], (c->fm ? kFm : 19) * (size_t)P)) != MCB_OK) return bail(st);
            if (!moments && (st = alloc(c, &c->g[b], 19 * (size_t)P)) != MCB_OK) return bail(st);
        }
        float** scal[] = {&c->phi, &c->rho, &c->ux, &c->uy, &c->uz, &c->ffx, &c->ffy, &c->ffz,
                          &c->normx, &c->normy, &c->normz};
        for (auto* q : scal)
            if ((st = alloc(c, q, (size_t)P)) != MCB_OK) return bail(st);
        if (moments) {
            c->gm = GM_INITIAL;
            c->phis[0] = c->phi;
            float** gmf[] = {&c->phis[1], &c->gmu[0], &c->gmu[1], &c->gmu[2]};
            for (auto* q : gmf)   // with f in moment form u is among its planes
                if ((q == &c->phis[1] || !c->fm) && (st = alloc(c, q, (size_t)P)) != MCB_OK) return bail(st);
            // n is stored beside u unless the phase pass recomputes it.
            if (!ctx_recomputes_n(*p))
                for (int k = 0; k < 3; ++k)
                    if ((st = alloc(c, &c->gmn[k], (size_t)P)) != MCB_OK) return bail(st);
        }
        if (p->bc_mode == MCB_BC_REFERENCE) {
            // Face cells never collide: their f*_l stay w_l in both buffers
            // (fgBoundary's rho(e) w_l with the face rho of 1) and their u and
            // n stay 0, so the moment form gives w_l phi(e) for their g.
            for (int b = 0; b < 2; ++b)
                if (launch_init_pull(c->phi, c->f[b], nullptr, c->geo, c->phys, c->stream) != hipSuccess)
                    return bail(MCB_ERR_HIP);
            for (int k = 0; k < 3; ++k)
                if (hipMemsetAsync(c->gmu[k], 0, (size_t)P * sizeof(float), c->stream) != hipSuccess ||
                    (c->gmn[k] && hipMemsetAsync(c->gmn[k], 0, (size_t)P * sizeof(float), c->stream) != hipSuccess))
                    return bail(MCB_ERR_HIP);
            if (hipStreamSynchronize(c->stream) != hipSuccess) return bail(MCB_ERR_HIP);
        }
        c->tiling = choose_tiling_single(*p);
        if (c->fm) {
            if (!fm_tiling_supported(c->tiling))
                c->tiling = Tiling{p->nx >= 64 ? 64 : (p->nx >= 32 ? 32 : 16), p->nx >= 32 ? 8 : 16, c->tiling.zc};
            // Its collide pass starts each chunk with two source planes:
            // 32-plane chunks where the lattice keeps two workgroups per CU
            // (256^3: 0.595 against 0.609 ms with 16).
            const Tiling& t = c->tiling;
            const long long layer = (long long)((p->nx + t.tx - 1) / t.tx) * ((p->ny + t.ty - 1) / t.ty);
            if (2 * t.zc <= p->nz && t.zc < 32 && layer * ((p->nz + 2 * t.zc - 1) / (2 * t.zc)) >= 512)
                c->tiling.zc = 2 * t.zc;
        }
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
    c->batch.reset();
    drop_fused_graphs(c);
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

extern "C++" {
namespace mcb {
// Initial state from bubbles computed on the device (periodic fused or
// 16-bit contexts; used by the serving loop, where a host-side phi of a large
// lattice would cost more than its steps). Not part of the C ABI.
mcb_status ctx_init_bubbles_device(mcb_ctx* c, const BubbleSet& s) {
    if (!c) return MCB_ERR_INVALID_ARG;
    MCB_HIP(c, hipSetDevice(c->device));
    if (c->batch) {
        if (c->case_id) {
            mcb_status st = batch_status(c, c->batch->remove(c->case_id));
            if (st != MCB_OK) return st;
            c->case_id = 0;
        }
        mcb_status st = batch_status(c, c->batch->add_bubbles(c->p, s.b, s.n, s.width, &c->case_id));
        if (st != MCB_OK) return st;
        if ((st = batch_status(c, c->batch->set_tiling(c->case_id, c->tiling))) != MCB_OK) return st;
        c->steps = 0;
        c->initialised = true;
        return MCB_OK;
    }
    if (c->path != MCB_PATH_FUSED || c->p.bc_mode != MCB_BC_PERIODIC)
        return fail(c, MCB_ERR_UNSUPPORTED, "device initialisation needs the periodic fused path");
    fused_init_target(c);
    MCB_HIP(c, launch_bubble_phi(c->phi, c->geo, s, c->stream));
    MCB_HIP(c, fused_init_state(c));
    MCB_HIP(c, hipStreamSynchronize(c->stream));
    c->steps = 0;
    c->initialised = true;
    return MCB_OK;
}

// Re-target a context at another case of the same lattice sides, storage,
// boundary mode and g form, then initialise it from bubbles on the device.
// The serving loop's sequential baseline keeps one context per lattice shape
// this way instead of allocating and freeing one per request. Not part of
// the C ABI.
mcb_status ctx_reset_case(mcb_ctx* c, const mcb_params& p, const BubbleSet& s) {
    if (!c) return MCB_ERR_INVALID_ARG;
    if (host_validate(&p, &c->err) != MCB_OK) return MCB_ERR_INVALID_ARG;
    if (p.nx != c->p.nx || p.ny != c->p.ny || p.nz != c->p.nz || p.storage != c->p.storage ||
        p.bc_mode != c->p.bc_mode || (p.tau_g == 1.0f) != (c->p.tau_g == 1.0f) || p.f_form != c->p.f_form)
        return fail(c, MCB_ERR_INVALID_ARG, "a context is re-targeted only at a case of the same shape and form");
    c->p.tau_f = p.tau_f;
    c->p.tau_g = p.tau_g;
    c->p.sigma = p.sigma;
    c->p.sharp_c = p.sharp_c;
    c->p.weight_set = p.weight_set;
    c->p.moving_wall = p.moving_wall;
    c->phys = host_phys(c->p);
    mcb_op_consts_from_params(&c->p, &c->oc);
    return ctx_init_bubbles_device(c, s);
}
}  // namespace mcb
}  // extern "C++"

mcb_status mcb_init_phi(mcb_ctx* c, const float* phi_host, size_t n) {
    if (!c || !phi_host || n != (size_t)c->geo.n) return fail(c, MCB_ERR_INVALID_ARG, "phi size mismatch");
    MCB_HIP(c, hipSetDevice(c->device));
    if (c->batch) {
        if (c->case_id) {
            mcb_status st = batch_status(c, c->batch->remove(c->case_id));
            if (st != MCB_OK) return st;
            c->case_id = 0;
        }
        mcb_status st = batch_status(c, c->batch->add(c->p, phi_host, &c->case_id));
        if (st != MCB_OK) return st;
        Tiling t = c->tiling;
        if ((st = batch_status(c, c->batch->set_tiling(c->case_id, t))) != MCB_OK) return st;
        c->steps = 0;
        c->initialised = true;
        return MCB_OK;
    }
    const size_t bytes = (size_t)c->geo.n * sizeof(float);
    const int NX = c->p.nx, NY = c->p.ny, NZ = c->p.nz;
    if (c->path == MCB_PATH_FUSED) {
        fused_init_target(c);
        MCB_HIP(c, hipMemcpyAsync(c->phi, phi_host, bytes, hipMemcpyHostToDevice, c->stream));
        MCB_HIP(c, fused_init_state(c));
    } else {
        MCB_HIP(c, hipMemcpyAsync(c->phi, phi_host, bytes, hipMemcpyHostToDevice, c->stream));
        // src/var.cu:86-101 clears these; initTensor/initDist set the rest.
        float* zero[] = {c->ux, c->uy, c->uz, c->normx, c->normy, c->normz, c->curvature, c->indicator,
                         c->ffx, c->ffy, c->ffz};
        for (float* q : zero) MCB_HIP(c, hipMemsetAsync(q, 0, bytes, c->stream));
        MCB_HIP(c, hipMemsetAsync(c->sg_out, 0, 19 * bytes, c->stream));
        MCB_OP(c, mcb_initTensor(c->pxx, c->pyy, c->pzz, c->pxy, c->pxz, c->pyz, c->rho, NX, NY, NZ, &c->oc, c->stream));
        MCB_OP(c, mcb_initDist(c->rho, c->phi, c->sf, c->sg, NX, NY, NZ, &c->oc, c->stream));
    }
    MCB_HIP(c, hipStreamSynchronize(c->stream));
    c->steps = 0;
    c->initialised = true;
    return MCB_OK;
}

mcb_status mcb_step(mcb_ctx* c, int nsteps) {
    mcb_status st = check_ready(c);
    if (st != MCB_OK) return st;
    if (nsteps < 0) return fail(c, MCB_ERR_INVALID_ARG, "negative step count");
    return any_step(c, nsteps);
}

mcb_status mcb_step_timed(mcb_ctx* c, int nsteps, double* ms) {
    mcb_status st = check_ready(c);
    if (st != MCB_OK) return st;
    if (nsteps < 0 || !ms) return fail(c, MCB_ERR_INVALID_ARG, "bad arguments");
    MCB_HIP(c, hipEventRecord(c->ev[0], c->stream));
    if ((st = any_step(c, nsteps)) != MCB_OK) return st;
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
    if (c->batch) return batch_status(c, c->batch->macros(c->case_id));
    if (c->path == MCB_PATH_FUSED) {
        const CaseView v = fused_view(c);
        MCB_HIP(c, launch_phase(v, c->tiling, c->stream));
        MCB_HIP(c, launch_collide(v, COLLIDE_MACROS, c->tiling, c->stream));
        if (c->gm != GM_POPULATIONS) c->phi = v.phi;   // the phi just computed (not the state's)
        return MCB_OK;
    }
    return staged_macros(c);
}

mcb_status mcb_read_field(mcb_ctx* c, int field, float* host, size_t n) {
    mcb_status st = check_ready(c);
    if (st != MCB_OK) return st;
    if (!host) return fail(c, MCB_ERR_INVALID_ARG, "null host buffer");
    if (c->batch) return batch_status(c, c->batch->read(c->case_id, field, host, n));
    const size_t cells = (size_t)c->geo.n;
    if (field == MCB_FIELD_F || field == MCB_FIELD_G) {
        if (n != 19 * cells) return fail(c, MCB_ERR_INVALID_ARG, "population buffer must hold 19*nx*ny*nz floats");
        if (c->path == MCB_PATH_STAGED) {
            const float* src = field == MCB_FIELD_F ? c->sf : c->sg;
            MCB_HIP(c, hipMemcpyAsync(host, src, 19 * cells * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        } else {
            if (field == MCB_FIELD_G && c->p.bc_mode == MCB_BC_REFERENCE)
                return fail(c, MCB_ERR_UNSUPPORTED,
                            "g in the reference layout with ghost-layer boundaries: use the staged path");
            // Gather into the idle A/B buffer (f in moment form: a scratch
            // buffer of 19 planes), then copy plane by plane.
            if (c->fm && !c->fscratch) {
                mcb_status as = alloc(c, &c->fscratch, 19 * (size_t)c->geo.plane);
                if (as != MCB_OK) return as;
            }
            float* idle_f = c->fm ? c->fscratch : c->f[c->cur ^ 1];
            if (field == MCB_FIELD_F && c->fm) {
                MCB_HIP(c, launch_fm_to_ref(fused_view(c), idle_f, c->stream));
            } else if (field == MCB_FIELD_G && c->gm != GM_POPULATIONS) {
                MCB_HIP(c, launch_gm_to_ref(fused_view(c), idle_f, c->stream));
            } else {
                const float* src = field == MCB_FIELD_F ? c->f[c->cur] : c->g[c->cur];
                MCB_HIP(c, launch_link_shift(src, c->geo.blocked, field == MCB_FIELD_F ? idle_f : c->g[c->cur ^ 1],
                                             0, c->geo, +1, c->stream));
            }
            const float* tmp = field == MCB_FIELD_F || c->gm != GM_POPULATIONS ? idle_f : c->g[c->cur ^ 1];
            for (int l = 0; l < 19; ++l)
                MCB_HIP(c, hipMemcpyAsync(host + l * cells, tmp + (size_t)l * c->geo.plane, cells * sizeof(float),
                                          hipMemcpyDeviceToHost, c->stream));
            if (c->p.bc_mode == MCB_BC_REFERENCE)   // restore the idle buffer's face populations (w_l)
                MCB_HIP(c, launch_init_pull(c->phi, c->f[c->cur ^ 1], nullptr, c->geo, c->phys, c->stream));
        }
    } else {
        const float* src = scalar_field(c, field);
        if (!src) return fail(c, MCB_ERR_INVALID_ARG, "unknown field");
        if (n != cells) return fail(c, MCB_ERR_INVALID_ARG, "field buffer must hold nx*ny*nz floats");
        MCB_HIP(c, hipMemcpyAsync(host, src, cells * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        MCB_HIP(c, hipStreamSynchronize(c->stream));
        if (field == MCB_FIELD_PHI && c->path == MCB_PATH_FUSED && c->p.bc_mode == MCB_BC_REFERENCE &&
            c->gm == GM_MOMENTS && src == c->phis[c->cur ^ 1])
            reference_faces_from_interior(c->p, host);
    }
    MCB_HIP(c, hipStreamSynchronize(c->stream));
    return MCB_OK;
}

mcb_status mcb_write_field(mcb_ctx* c, int field, const float* host, size_t n) {
    mcb_status st = check_ready(c);
    if (st != MCB_OK) return st;
    if (!host) return fail(c, MCB_ERR_INVALID_ARG, "null host buffer");
    if (c->batch) return batch_status(c, c->batch->write(c->case_id, field, host, n));
    const size_t cells = (size_t)c->geo.n;
    if (field == MCB_FIELD_F || field == MCB_FIELD_G) {
        if (n != 19 * cells) return fail(c, MCB_ERR_INVALID_ARG, "population buffer must hold 19*nx*ny*nz floats");
        if (c->path == MCB_PATH_STAGED) {
            float* dst = field == MCB_FIELD_F ? c->sf : c->sg;
            MCB_HIP(c, hipMemcpyAsync(dst, host, 19 * cells * sizeof(float), hipMemcpyHostToDevice, c->stream));
        } else {
            if (c->fm)
                return fail(c, MCB_ERR_UNSUPPORTED, "populations cannot be written with f in moment form");
            if (field == MCB_FIELD_G && c->p.bc_mode == MCB_BC_REFERENCE)
                return fail(c, MCB_ERR_UNSUPPORTED, "writing g with ghost-layer boundaries: use the staged path");
            if (field == MCB_FIELD_G && c->gm != GM_POPULATIONS) {
                // Arbitrary g populations have no moment form: from here on
                // the context keeps g as populations (phi as its one buffer).
                // (A state image of the moment form loaded after an earlier
                // write keeps these buffers: allocate only what is missing.)
                for (int b = 0; b < 2; ++b) {
                    if (c->g[b]) continue;
                    mcb_status as = alloc(c, &c->g[b], 19 * (size_t)c->geo.plane);
                    if (as != MCB_OK) return as;
                }
                c->phi = c->phis[c->cur ^ 1];
                c->gm = GM_POPULATIONS;
            }
            float* dst = field == MCB_FIELD_F ? c->f[c->cur] : c->g[c->cur];
            float* tmp = field == MCB_FIELD_F ? c->f[c->cur ^ 1] : c->g[c->cur ^ 1];
            for (int l = 0; l < 19; ++l)
                MCB_HIP(c, hipMemcpyAsync(tmp + (size_t)l * c->geo.plane, host + l * cells, cells * sizeof(float),
                                          hipMemcpyHostToDevice, c->stream));
            MCB_HIP(c, launch_link_shift(tmp, 0, dst, c->geo.blocked, c->geo, -1, c->stream));
            if (c->p.bc_mode == MCB_BC_REFERENCE) {
                // The written face populations feed the next step once (as
                // the reference's first momentiCalc reads them); after it they
                // go back to w_l. The idle buffer was the staging area.
                MCB_HIP(c, launch_init_pull(c->phi, tmp, nullptr, c->geo, c->phys, c->stream));
                c->ref_reset_f = c->cur;
            }
        }
    } else {
        if (c->path == MCB_PATH_FUSED && c->p.bc_mode == MCB_BC_REFERENCE && field == MCB_FIELD_PHI)
            return fail(c, MCB_ERR_UNSUPPORTED, "writing phi with ghost-layer boundaries: use the staged path");
        if (field == MCB_FIELD_PHI && c->gm != GM_POPULATIONS && n == cells) {
            // phi is part of the moment-form state: write the buffer the next
            // phase pass recomputes, as writing phi does with populations.
            c->phi = c->phis[c->cur];
        }
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
    if (c->batch) return batch_status(c, c->batch->reduce(c->case_id, sum_phi, kinetic));
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

// Device time of each pass summed over nsteps fused FP32 steps:
// out_ms[0] = k_phase, out_ms[1] = k_collide.
mcb_status mcb_step_kernel_times(mcb_ctx* c, int nsteps, double* out_ms, int n_out) {
    mcb_status st = check_ready(c);
    if (st != MCB_OK) return st;
    if (c->path != MCB_PATH_FUSED || c->batch)
        return fail(c, MCB_ERR_UNSUPPORTED, "kernel times need the fused FP32 path");
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
        const CaseView v = fused_view(c);
        MCB_HIP(c, hipEventRecord(e[0], c->stream));
        MCB_HIP(c, launch_phase(v, c->tiling, c->stream));
        MCB_HIP(c, hipEventRecord(e[1], c->stream));
        MCB_HIP(c, launch_collide(v, step_mode(c), c->tiling, c->stream));
        MCB_HIP(c, hipEventRecord(e[2], c->stream));
        MCB_HIP(c, after_fused_step(c));
        ++c->steps;
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
    const bool ok = c->fm ? fm_tiling_supported(t)
                          : (c->p.storage == MCB_STORE_FP32 ? tiling_supported(t) : tiling_supported16(t));
    if (!ok || zc > c->p.nz) return fail(c, MCB_ERR_INVALID_ARG, "unsupported tiling");
    if (c->batch && c->case_id) {
        mcb_status st = batch_status(c, c->batch->set_tiling(c->case_id, t));
        if (st != MCB_OK) return st;
    }
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

mcb_status mcb_autotune(const mcb_params* p, int* tx, int* ty, int* zc, double* ms_per_step) {
    if (!p || host_validate(p, nullptr) != MCB_OK) return MCB_ERR_INVALID_ARG;
    if (!fused_fits(*p)) return MCB_ERR_UNSUPPORTED;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= p->device) return MCB_ERR_NO_DEVICE;
    Tiling t{};
    double ms = 0.0;
    mcb_status st = autotune(*p, p->device, &t, &ms);
    if (st != MCB_OK) return st;
    if (tx) *tx = t.tx;
    if (ty) *ty = t.ty;
    if (zc) *zc = t.zc;
    if (ms_per_step) *ms_per_step = ms;
    return MCB_OK;
}

// ---- batches

size_t mcb_case_bytes(const mcb_params* p) {
    if (!p || host_validate(p, nullptr) != MCB_OK) return 0;
    return case_bytes(*p);
}

mcb_status mcb_batch_create(int device, int storage, void* stream, mcb_batch** out) {
    if (!out) return MCB_ERR_INVALID_ARG;
    *out = nullptr;
    if (storage < MCB_STORE_FP32 || storage > MCB_STORE_BF16) return MCB_ERR_INVALID_ARG;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0 || device < 0 || device >= count)
        return MCB_ERR_NO_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return MCB_ERR_NO_DEVICE;
    mcb_batch* b = new (std::nothrow) mcb_batch();
    if (!b) return MCB_ERR_OUT_OF_MEMORY;
    b->device = device;
    b->storage = storage;
    if (stream) {
        b->stream = static_cast<hipStream_t>(stream);
    } else if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess) {
        b->own_stream = true;
    } else {
        delete b;
        return MCB_ERR_HIP;
    }
    if (hipEventCreate(&b->ev[0]) != hipSuccess || hipEventCreate(&b->ev[1]) != hipSuccess) {
        mcb_batch_destroy(b);
        return MCB_ERR_HIP;
    }
    b->b.reset(new (std::nothrow) Batch(device, b->stream, true));
    if (!b->b) {
        mcb_batch_destroy(b);
        return MCB_ERR_OUT_OF_MEMORY;
    }
    *out = b;
    return MCB_OK;
}

mcb_status mcb_batch_destroy(mcb_batch* b) {
    if (!b) return MCB_ERR_INVALID_ARG;
    (void)hipSetDevice(b->device);
    b->b.reset();
    for (auto& e : b->ev)
        if (e) (void)hipEventDestroy(e);
    if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return MCB_OK;
}

static mcb_status bstat(mcb_batch* b, mcb_status st) {
    if (st != MCB_OK) b->err = b->b->error();
    return st;
}

mcb_status mcb_batch_add_phi(mcb_batch* b, const mcb_params* p, const float* phi, size_t n, int* case_id) {
    if (!b || !p || !phi || !case_id) return MCB_ERR_INVALID_ARG;
    if (host_validate(p, nullptr) != MCB_OK) return MCB_ERR_INVALID_ARG;
    if (n != (size_t)p->nx * p->ny * p->nz) return MCB_ERR_INVALID_ARG;
    if (p->storage != b->storage || p->bc_mode != MCB_BC_PERIODIC) return MCB_ERR_UNSUPPORTED;
    return bstat(b, b->b->add(*p, phi, case_id));
}

mcb_status mcb_batch_add(mcb_batch* b, const mcb_params* p, int n, const mcb_bubble* bubbles, float width,
                         int* case_id) {
    if (!b || !p || !case_id) return MCB_ERR_INVALID_ARG;
    if (host_validate(p, nullptr) != MCB_OK) return MCB_ERR_INVALID_ARG;
    std::vector<float> phi((size_t)p->nx * p->ny * p->nz);
    mcb_status st = mcb_host_init_phi(p, n, bubbles, width, phi.data(), phi.size());
    if (st != MCB_OK) return st;
    return mcb_batch_add_phi(b, p, phi.data(), phi.size(), case_id);
}

mcb_status mcb_batch_remove(mcb_batch* b, int case_id) {
    if (!b) return MCB_ERR_INVALID_ARG;
    return bstat(b, b->b->remove(case_id));
}

mcb_status mcb_batch_step(mcb_batch* b, int nsteps) {
    if (!b) return MCB_ERR_INVALID_ARG;
    return bstat(b, b->b->step(nsteps));
}

mcb_status mcb_batch_step_timed(mcb_batch* b, int nsteps, double* ms) {
    if (!b || !ms) return MCB_ERR_INVALID_ARG;
    if (hipSetDevice(b->device) != hipSuccess) return MCB_ERR_NO_DEVICE;
    if (b->b->count() > 0) {
        // Build views, tables and graphs outside the timed region.
        mcb_status st = bstat(b, b->b->step(0));
        if (st != MCB_OK) return st;
    }
    if (hipEventRecord(b->ev[0], b->stream) != hipSuccess) return MCB_ERR_HIP;
    mcb_status st = bstat(b, b->b->step(nsteps));
    if (st != MCB_OK) return st;
    float t = 0.0f;
    if (hipEventRecord(b->ev[1], b->stream) != hipSuccess || hipEventSynchronize(b->ev[1]) != hipSuccess ||
        hipEventElapsedTime(&t, b->ev[0], b->ev[1]) != hipSuccess)
        return MCB_ERR_HIP;
    *ms = t;
    return MCB_OK;
}

mcb_status mcb_batch_compute_macros(mcb_batch* b, int case_id) {
    if (!b) return MCB_ERR_INVALID_ARG;
    return bstat(b, b->b->macros(case_id));
}

mcb_status mcb_batch_read_field(mcb_batch* b, int case_id, int field, float* host, size_t n) {
    if (!b) return MCB_ERR_INVALID_ARG;
    return bstat(b, b->b->read(case_id, field, host, n));
}

mcb_status mcb_batch_write_field(mcb_batch* b, int case_id, int field, const float* host, size_t n) {
    if (!b) return MCB_ERR_INVALID_ARG;
    return bstat(b, b->b->write(case_id, field, host, n));
}

mcb_status mcb_batch_reduce(mcb_batch* b, int case_id, double* sum_phi, double* kinetic) {
    if (!b) return MCB_ERR_INVALID_ARG;
    return bstat(b, b->b->reduce(case_id, sum_phi, kinetic));
}

mcb_status mcb_batch_case_steps(mcb_batch* b, int case_id, long long* steps) {
    if (!b || !steps) return MCB_ERR_INVALID_ARG;
    CaseState* c = b->b->get(case_id);
    if (!c) return MCB_ERR_INVALID_ARG;
    *steps = c->steps;
    return MCB_OK;
}

mcb_status mcb_batch_set_tiling(mcb_batch* b, int case_id, int tx, int ty, int zc) {
    if (!b) return MCB_ERR_INVALID_ARG;
    return bstat(b, b->b->set_tiling(case_id, Tiling{tx, ty, zc}));
}

mcb_status mcb_batch_synchronize(mcb_batch* b) {
    if (!b) return MCB_ERR_INVALID_ARG;
    return bstat(b, b->b->synchronize());
}

int mcb_batch_count(const mcb_batch* b) { return b ? b->b->count() : 0; }

size_t mcb_batch_bytes(const mcb_batch* b) { return b ? b->b->bytes() : 0; }

const char* mcb_batch_last_error(const mcb_batch* b) { return b ? b->err.c_str() : "null batch"; }

mcb_status mcb_state_size(mcb_ctx* c, size_t* bytes) {
    mcb_status st = check_ready(c);
    if (st != MCB_OK) return st;
    if (!bytes) return fail(c, MCB_ERR_INVALID_ARG, "null size pointer");
    if (c->batch) return batch_status(c, c->batch->state_size(c->case_id, bytes));
    *bytes = state_image_bytes(ctx_state_parts(c));
    return MCB_OK;
}

mcb_status mcb_save_state(mcb_ctx* c, void* buf, size_t bytes) {
    mcb_status st = check_ready(c);
    if (st != MCB_OK) return st;
    if (c->batch) return batch_status(c, c->batch->save_state(c->case_id, buf, bytes));
    MCB_HIP(c, hipStreamSynchronize(c->stream));
    StateHeader h{};
    h.params = c->p;
    h.path = c->path;
    h.gm = c->path == MCB_PATH_STAGED ? GM_POPULATIONS : c->gm;
    h.steps = c->steps;
    const bool staged = c->path == MCB_PATH_FUSED && c->geo.blocked;
    if (staged) MCB_HIP(c, pops_image_copy(c, true));
    st = state_image_save(c->stream, h, ctx_state_parts(c), buf, bytes, &c->err);
    if (staged) MCB_HIP(c, pops_image_done(c));
    MCB_HIP(c, hipStreamSynchronize(c->stream));
    return st;
}

mcb_status mcb_load_state(mcb_ctx* c, const void* buf, size_t bytes) {
    if (!c) return MCB_ERR_INVALID_ARG;
    StateHeader h;
    mcb_status st = state_image_header(buf, bytes, &h, &c->err);
    if (st != MCB_OK) return st;
    if (!same_state_params(h.params, c->p) || h.path != c->path)
        return fail(c, MCB_ERR_INVALID_ARG, "state image of another lattice, parameter set or path");
    if (h.gm < GM_POPULATIONS || h.gm > GM_INITIAL || (h.gm != GM_POPULATIONS && c->p.tau_g != 1.0f) ||
        (c->path == MCB_PATH_STAGED && h.gm != GM_POPULATIONS))
        return fail(c, MCB_ERR_INVALID_ARG, "state image with an invalid g form");
    if (!c->initialised) {
        // Any initial state gives the context its buffers (and, for 16 bits,
        // its batch case); the image then replaces it.
        const std::vector<float> zero((size_t)c->geo.n, 0.0f);
        if ((st = mcb_init_phi(c, zero.data(), zero.size())) != MCB_OK) return st;
    }
    if ((st = check_ready(c)) != MCB_OK) return st;
    if (c->batch) return batch_status(c, c->batch->load_state(c->case_id, buf, bytes));
    MCB_HIP(c, hipStreamSynchronize(c->stream));
    if (c->path == MCB_PATH_FUSED) {
        if (h.gm == GM_POPULATIONS && c->gm != GM_POPULATIONS) {
            // As writing G: from here on g is kept as populations.
            for (int b = 0; b < 2; ++b)
                if (!c->g[b] && (st = alloc(c, &c->g[b], 19 * (size_t)c->geo.plane)) != MCB_OK) return st;
            c->phi = c->phis[c->cur ^ 1];
        }
        c->gm = h.gm;
        c->ref_reset_f = -1;   // the image's face populations are w_l
    }
    if ((st = state_image_load(c->stream, buf, bytes, ctx_state_parts(c), &c->err)) != MCB_OK) return st;
    if (c->path == MCB_PATH_FUSED && c->geo.blocked) {
        MCB_HIP(c, pops_image_copy(c, false));
        MCB_HIP(c, pops_image_done(c));
        MCB_HIP(c, hipStreamSynchronize(c->stream));
    }
    if (c->path == MCB_PATH_FUSED && c->gm != GM_POPULATIONS) {
        c->phi = c->phis[c->cur ^ 1];
        if (c->gmn[0]) {   // n is not in the image: gradCalc of the state's phi
            MCB_HIP(c, launch_normals(fused_view(c), c->stream));
            MCB_HIP(c, hipStreamSynchronize(c->stream));
        }
    }
    c->steps = h.steps;
    return MCB_OK;
}

mcb_status mcb_state_info(const void* buf, size_t bytes, mcb_params* params, long long* steps) {
    StateHeader h;
    mcb_status st = state_image_header(buf, bytes, &h, nullptr);
    if (st != MCB_OK) return st;
    if (params) *params = h.params;
    if (steps) *steps = h.steps;
    return MCB_OK;
}

mcb_status mcb_batch_state_size(mcb_batch* b, int case_id, size_t* bytes) {
    if (!b) return MCB_ERR_INVALID_ARG;
    return bstat(b, b->b->state_size(case_id, bytes));
}

mcb_status mcb_batch_save_state(mcb_batch* b, int case_id, void* buf, size_t bytes) {
    if (!b) return MCB_ERR_INVALID_ARG;
    return bstat(b, b->b->save_state(case_id, buf, bytes));
}

mcb_status mcb_batch_load_state(mcb_batch* b, int case_id, const void* buf, size_t bytes) {
    if (!b) return MCB_ERR_INVALID_ARG;
    return bstat(b, b->b->load_state(case_id, buf, bytes));
}

mcb_status mcb_batch_add_state(mcb_batch* b, const void* buf, size_t bytes, int* case_id) {
    if (!b || !case_id) return MCB_ERR_INVALID_ARG;
    StateHeader h;
    mcb_status st = state_image_header(buf, bytes, &h, &b->err);
    if (st != MCB_OK) return st;
    if (h.params.storage != b->storage) {
        b->err = "state image of another storage form than the batch's";
        return MCB_ERR_INVALID_ARG;
    }
    // The header's parameters reach the kernels: hold them to the same checks
    // as parameters passed by a caller.
    if ((st = host_validate(&h.params, &b->err)) != MCB_OK) return st;
    return bstat(b, b->b->add_state(buf, bytes, case_id));
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
