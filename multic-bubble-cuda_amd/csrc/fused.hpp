// Fused two-pass periodic step of libmcb, internal interface.
//
// Storage ("pull" form): after a step, buffer f[cur] holds the
// post-collision populations f*_l(x) at their source cell x, unstreamed, and
// likewise g[cur]. The pre-collision population the reference holds at x
// after streaming is f*_l(x - c_l), so every kernel gathers with a shift and
// every store is aligned. This carries exactly the reference's numbers
// (collisionFluid pushes f*_l(x) to x + c_l, src/lbm.cu:285-286;
// streamingCalc pulls g_l(x - c_l), src/lbm.cu:350-355).
//
// Pass 1, phase   : phi(x) = sum_l g*_l(x - c_l)              (phiCalc)
// Pass 2, collide : gradCalc + curvatureCalc on LDS-staged phi and n,
//                   momentiCalc on gathered f, collisionFluid and
//                   collisionPhase, aligned stores of f*, g* to the other
//                   buffer.
#pragma once
#include "common.hpp"

namespace mcb {

enum CollideMode { COLLIDE_STEP = 0, COLLIDE_STEP_BGK = 1, COLLIDE_MACROS = 2 };

// How a case's state holds g (CaseView::gm).
//   GM_POPULATIONS : 19 planes of g* (pull form), any tau_g.
//   GM_MOMENTS     : phi, u and n of the last collision (tau_g = 1 only); n
//                    is stored or recomputed from phi (CaseView::gm_n).
//   GM_INITIAL     : the initial state g_l(x) = w_l phi0(x) (initDist), before
//                    the first step of the moment form; phi_prev holds phi0.
enum GMode { GM_POPULATIONS = 0, GM_MOMENTS = 1, GM_INITIAL = 2 };

// Planes of f in moment form (CaseView::fm): rho, ux, uy, uz, pxx, pyy,
// pzz, pxy, pxz, pyz, fx, fy, fz.
constexpr int kFm = 13;

struct Tiling {
    int tx, ty, zc;   // x/y tile of a workgroup, z cells marched per workgroup
};

// Mass bookkeeping of a case with 16-bit populations (device memory). The
// phase pass adds delta[q] to every phi it computes at case parity q and
// writes one FP64 partial sum of phi per work item; after the pass
// launch_mass_finish sums the partials in a fixed order and stores
// delta[q ^ 1] = (target - sum phi) / cells for the next step.
struct MassCell {
    double target;          // sum of phi the case must keep
    double delta[2];
};

// Mass-drift correction of the 16-bit storage forms in moment form (tau_g =
// 1): the phase pass sums phi as 64-bit fixed-point integers (kMassQScale
// per unit), an exact and hence order-independent sum, so the total and the
// correction do not depend on how cells fall into workgroups (the step stays
// the same bits under shifts and tilings). Each workgroup stores its partial;
// a small kernel after the pass (launch_mass_finish) sums them and records
// the first moment-form pass's total as the target, or sets the correction
// delta that the next pass adds to every phi. (No workgroup waits on others:
// an agent-scope release per workgroup would write back its XCD's L2.)
struct MassQ {
    long long target;   // fixed-point sum of phi to keep
    int set;            // target recorded
    int reserved;
    double delta[2];    // per-cell correction the next phase pass adds, by state parity
};
// 2^31 per unit. A cell whose phi is below half the resolution counts as
// zero, so what the correction adds there is invisible to the sum, and the
// loop keeps adding until such cells reach half the resolution: at most
// 2^-32 per cell (4e-3 on 256^3, ~3e-8 of the benchmark's total phase); at
// 2^-20 this was 8 and ran away. The sum stays below 2^63 up to 3.6e9 cells
// at phi <= 1.2.
constexpr double kMassQScale = 2147483648.0;

// Everything a kernel needs about one case (one lattice).
struct CaseView {
    Geo geo;
    Phys phys;
    int zc;                 // z planes per collide workgroup
    int tx, ty;             // tile of the case's work items
    int mslot;              // parity of the case's state: delta slot the phase pass reads
    MassCell* mass;         // 16-bit storage only
    MassQ* massq;           // 16-bit storage in moment form only
    double* partial;        // one per work item of the case's tiling
    const void* fsrc;       // populations of the previous step (pull form)
    void* fdst;             // populations of this step
    const void* gsrc;
    void* gdst;
    float* phi;             // phi of this step: written by the phase pass
    const float* phi_prev;  // 16-bit g populations: phi of the previous step (anchor of g);
                            // moment form: phi of the previous step's collision
    // Moment form of g (FP32, tau_g = 1): g*_l(y) is collisionPhase of
    // (phi, u, n) at y, so the state keeps u (gm_u) instead of 19 g planes
    // and the phase pass recomputes g*_l(x - c_l). n is gradCalc of phi
    // around y: with gm_n set the collide pass stores it, with gm_n null the
    // phase pass recomputes it (phase_tile_gm_rn; 24 fewer bytes per cell
    // and step, more arithmetic per plane, so it pays on long z-chunks).
    // 16-bit forms: gm_u holds B = 3 phi u + phi_n n as FP16 (2 bytes per
    // cell; ld_b), gm_n is null.
    float* gm_u[3];
    float* gm_n[3];
    int gm;                 // GM_POPULATIONS, GM_MOMENTS or GM_INITIAL
    // f in moment form (FP32, g in moment form, periodic; collide_tile_fm):
    // fsrc/fdst hold 13 planes (rho, u, Pi^neq, F of a cell's last collision)
    // instead of 19 populations, and gm_u points at the state's u planes.
    int fm;
    // MCB_BC_PERIODIC, or MCB_BC_REFERENCE (FP32 moment form of a single
    // case): the reference's ghost-layer boundaries (src/lbm.cu:365-425) in
    // pull form. Face cells never collide; their f*_l stays w_l, their u
    // stays 0 and their n is 0 (gradCalc skips them), so an interior cell that gathers from a face e receives
    // rho(e) w_l = w_l (fgBoundary's f; face rho stays 1) and w_l phi(e)
    // (fgBoundary's g). The phase pass writes the state phi of a face cell as
    // boundaryConditions does: the previous state phi at the clamped interior
    // cell (at the first step: its own initial value), so the phi buffer a
    // collide pass reads carries the faces of the previous iteration, as the
    // reference's d_phi does when gradCalc runs.
    int bc;
    // COLLIDE_MACROS outputs
    float* rho;
    float* ux;
    float* uy;
    float* uz;
    float* fx;
    float* fy;
    float* fz;
    float* nx;
    float* ny;
    float* nz;
};

// One work item of a batched launch: case index, tile origin x0, y0, z0.
struct WorkItem {
    int c, x0, y0, z0;
};

// Tile shapes (TX, TY) the kernels are instantiated for, as X(TX, TY) lists.
// The FP32 list is also the autotuner's candidate list. The TX x 4 shapes
// with TX = 48 ... 176 span a whole row of lattices with that x side: one
// contiguous run per tile row and no padded lanes, which measured 1.1-1.9x
// the 64- and 32-wide tiles for the 16-bit forms on such sides.
#define MCB_TILES_FP32(X) \
    X(64, 8) X(64, 4) X(32, 8) X(32, 4) X(32, 16) X(16, 16) X(16, 8) X(48, 4) X(80, 4) X(96, 4) X(112, 4) X(144, 4) \
    X(176, 4)
#define MCB_TILES_16(X) \
    X(64, 8) X(64, 4) X(32, 8) X(16, 16) X(48, 4) X(80, 4) X(96, 4) X(112, 4) X(144, 4) X(176, 4)

bool tiling_supported(const Tiling& t);
int tiling_count();
Tiling tiling_at(int i);

// ---- single case, FP32 populations (bit-exact form)
// Phase pass; in moment form (gm == GM_MOMENTS) it runs on the tiles t.
hipError_t launch_phase(const CaseView& c, const Tiling& t, hipStream_t s);
// Tile and z-chunk of the moment-form phase pass for collide tiles t, rn =
// n recomputed: the tile is t's, or the recomputing pass's own where the
// lattice spans it; the chunk is as long as the lattice allows with two
// workgroups per CU (64 planes, halved down to 8: each chunk also reads the
// planes around it, and the recomputing pass starts each chunk with two
// planes of normals).
Tiling phase_tiling(const Geo& geo, const Tiling& t, bool rn);
// Whether a single-case FP32 moment-form context recomputes n in the phase
// pass (CaseView::gm_n null) rather than storing it: on chunks of 32 planes
// or more, where the 24 bytes per cell saved outweigh the per-chunk start.
bool phase_recomputes_n(const Geo& geo, const Tiling& t);
// n from the state phi into gm_n (a stored-normal case loaded from an image).
hipError_t launch_normals(const CaseView& c, hipStream_t s);
hipError_t launch_collide(const CaseView& c, int mode, const Tiling& t, hipStream_t s);
// f*_l(y) = w_l * 1, g*_l(y) = w_l * phi(y + c_l): the pull-form image of
// initTensor + initDist (src/init.cu:21-24, 69-79), in geo's population
// layout. g may be null (moment form).
hipError_t launch_init_pull(const float* phi, float* f, float* g, const Geo& geo, const Phys& p, hipStream_t s);
// Moment form of g -> reference layout (post-stream g_l(x), 19 planes):
// FP32 (phi, u) form (n recomputed), and the (phi, B) form of the 16-bit storage forms.
hipError_t launch_gm_to_ref(const CaseView& c, float* dst, hipStream_t s);
// f in moment form: the initial state (rho = 1, every other moment 0, whose
// rebuilt f*_l is initTensor's w_l) into the 13 planes of fm, and the
// reference's post-stream f layout (19 planes) rebuilt from the state c.fsrc.
hipError_t launch_fm_init(float* fm, const Geo& geo, hipStream_t s);
hipError_t launch_fm_to_ref(const CaseView& c, float* dst, hipStream_t s);
// Tile shapes instantiated for the f-moment collide pass.
bool fm_tiling_supported(const Tiling& t);
hipError_t launch_ab_to_ref(const CaseView& c, float* dst, hipStream_t s);
// dst_l(x) = src_l(x - sign*c_l) for all 19 links (sign = +1: pull form ->
// reference order; sign = -1: reference order -> pull form), src in
// population layout sblk and dst in dblk (Geo::blocked values; the reference
// order is always in planes).
hipError_t launch_link_shift(const float* src, int sblk, float* dst, int dblk, const Geo& geo, int sign,
                             hipStream_t s);
// The same populations in another layout (raw elements of elem_bytes, 4 or 2).
hipError_t launch_pop_relayout(const void* src, int sblk, void* dst, int dblk, const Geo& geo, unsigned elem_bytes,
                               hipStream_t s);

// ---- batched (many cases, one launch per pass), any storage form
// storage: MCB_STORE_FP32 / FP16 / BF16; work: device array of n_work items.
// After a batch's phase pass: the mass-drift corrections of its 16-bit
// cases (MassCell for stored g, MassQ for the moment form), one workgroup
// per case of the view array.
hipError_t launch_mass_finish(const CaseView* cases, unsigned n_cases, hipStream_t s);
hipError_t launch_phase_batch(const CaseView* cases, const WorkItem* work, unsigned n_work, const Tiling& t,
                              int storage, hipStream_t s);
hipError_t launch_collide_batch(const CaseView* cases, const WorkItem* work, unsigned n_work, const Tiling& t,
                                int mode, int storage, hipStream_t s);

// ---- 16-bit populations (a single case runs as a batch of one)
bool tiling_supported16(const Tiling& t);
// 16-bit initial state: f*_l = w_l (offset 0); g*_l(y) = w_l phi(y + c_l)
// against the anchor phi(y) when g is stored (g may be null: moment form).
hipError_t launch_init16(const float* phi, void* f, void* g, const Geo& geo, const Phys& p, int storage,
                         hipStream_t s);
// Decode 16-bit pull-form populations to FP32 reference layout:
// dst_l(x) = w_l anchor(x - c_l) + t_l(x - c_l) / K_l (anchor null: 1).
hipError_t launch_decode16(const void* src, const float* anchor, float* dst, const Geo& geo, const Phys& p,
                           int storage, hipStream_t s);
// Encode FP32 reference-layout populations to 16-bit pull form against the
// anchor field: t_l(x) = (src_l(x + c_l) - w_l anchor(x)) K_l (anchor null: 1).
hipError_t launch_encode16(const float* src, const float* anchor, void* dst, const Geo& geo, const Phys& p,
                           int storage, hipStream_t s);
// field[i] += delta for i < n (uniform phase correction of the drift).
hipError_t launch_add_scalar(float* field, long long n, const double* delta, hipStream_t s);

// Bubbles of a periodic initial state, passed by value to the init kernel.
constexpr int kMaxBubbleSet = 8;
struct BubbleSet {
    mcb_bubble b[kMaxBubbleSet];
    int n;
    float width;
};
// Periodic initial phi of a bubble set on the device (bubble_phi_cell: the
// bits of the host initialisation).
hipError_t launch_bubble_phi(float* phi, const Geo& geo, const BubbleSet& bubbles, hipStream_t s);
// *out = sum of v[0..n) in FP64 (partial: kReduceBlocks doubles of scratch).
hipError_t launch_sum(const float* v, long long n, double* partial, double* out, hipStream_t s);

// dst = src for bytes (a multiple of 16) with a plain float4 copy kernel.
hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s);
// Deterministic FP64 sums: out[0] = sum phi, out[1] = sum 0.5*rho*|u|^2.
constexpr int kReduceBlocks = 1024;
hipError_t launch_reduce(const float* phi, const float* rho, const float* ux, const float* uy, const float* uz,
                         long long n, double* partial, double* out, hipStream_t s);

}  // namespace mcb
