// Fused two-pass periodic step, FP32 populations (the bit-exact form), for
// one case or a batch of cases; plus init, layout shift, copy and reduction
// kernels. See fused.hpp for the storage convention and tile.hpp for the
// tile bodies.
//
// k_phase   : one wave64 per 64 consecutive x cells; 19 gathered loads (a
//             row shifted by at most one cell) and one aligned store.
// k_collide : a workgroup owns a TX x TY column of cells and marches it
//             through zc planes (collide_tile).
#include "bubble.hpp"
#include "tile.hpp"

namespace mcb {

// 16-bit storage launchers (fused16.hip).
hipError_t launch_phase_batch16(const CaseView* cases, const WorkItem* work, unsigned n_work, const Tiling& t,
                                int storage, hipStream_t s);
hipError_t launch_collide_batch16(const CaseView* cases, const WorkItem* work, unsigned n_work, const Tiling& t,
                                  int mode, int storage, hipStream_t s);

namespace {

// ----------------------------------------------------------------- k_phase

constexpr int PBX = 64, PBY = 4;

// Tiles per XCD run of the collide pass with FP32 populations in blocks
// (xcd_runs).
constexpr unsigned kCollideRun = 8;

// 1-D grid of gx*gy*nz workgroups of 64x4 cells, renumbered by xcd_order.
__global__ __launch_bounds__(PBX * PBY) void k_phase(CaseView c, unsigned gx, unsigned gy) {
    const Geo& G = c.geo;
    unsigned t = xcd_order(blockIdx.x, gridDim.x);
    const unsigned bx = t % gx;
    t /= gx;
    const unsigned by = t % gy;
    const int z = (int)(t / gy);
    const int x = (int)(bx * PBX + threadIdx.x);
    const int y = (int)(by * PBY + threadIdx.y);
    if (x >= G.nx || y >= G.ny) return;
    // Offsets of x-1, x, x+1 / rows y-1, y, y+1 / planes z-1, z, z+1.
    const unsigned xs[3] = {(unsigned)shift<-1>(x, G.nx), (unsigned)x, (unsigned)shift<1>(x, G.nx)};
    const unsigned ys[3] = {(unsigned)(shift<-1>(y, G.ny) * G.nx), (unsigned)(y * G.nx),
                            (unsigned)(shift<1>(y, G.ny) * G.nx)};
    const unsigned zs[3] = {(unsigned)(shift<-1>(z, G.nz) * G.nxy), (unsigned)(z * G.nxy),
                            (unsigned)(shift<1>(z, G.nz) * G.nxy)};
    const PopPlanes<StoreF32> gp(c.gsrc, G);
    const unsigned me = zs[1] + ys[1] + xs[1];
    float v = phase_sum<StoreF32>(c, gp, Columns(xs, ys, gp), xs, ys, zs);
    // Ghost-layer faces (moment form): before the first step a face keeps its
    // initial phi, later it takes boundaryConditions' copy.
    if (c.bc == MCB_BC_REFERENCE && ref_face(G, x, y, z))
        v = ld32(c.phi_prev, c.gm == GM_INITIAL ? me : ref_clamped(G, x, y, z));
    st32(c.phi, me, v);
}

// Moment-form phase pass (phase_tile_gm, or phase_tile_gm_rn when n is
// recomputed): the collide pass's tiles and grid. The recomputing form at
// 512-thread tiles is held to 80 VGPRs: three workgroups per CU (LDS ~49 KB
// each).
template <int TX, int TY, bool RN> constexpr int phase_gm_waves_per_eu() { return RN && TX * TY == 512 ? 6 : 1; }

// Tile of the phase pass when it recomputes n (nx, ny at least its sides):
// its ring cells each compute a normal, so the taller 32x16 tile (100 ring
// cells per 512, in two waves) does less than the collide pass's 64x8 (148
// per 512, three waves).
#ifndef MCB_PHASE_RN_TX
#define MCB_PHASE_RN_TX 32
#endif
#ifndef MCB_PHASE_RN_TY
#define MCB_PHASE_RN_TY 16
#endif
template <int TX, int TY, bool RN>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(phase_gm_waves_per_eu<TX, TY, RN>())))
void k_phase_gm(CaseView c) {
    const Geo& G = c.geo;
    const unsigned gx = (unsigned)(G.nx + TX - 1) / TX, gy = (unsigned)(G.ny + TY - 1) / TY;
    unsigned t = xcd_order(blockIdx.x, gridDim.x);
    const int x0 = (int)(t % gx) * TX;
    t /= gx;
    const int y0 = (int)(t % gy) * TY;
    const int zb = (int)(t / gy) * c.zc;
    if constexpr (RN) phase_tile_gm_rn<TX, TY>(c, x0, y0, zb, min(zb + c.zc, G.nz));
    else phase_tile_gm<TX, TY>(c, x0, y0, zb, min(zb + c.zc, G.nz));
}

// n of every cell from the state phi (normal_at), into gm_n: a stored-normal
// case resumed from a state image, which keeps phi and u only.
__global__ __launch_bounds__(PBX * PBY) void k_normals(CaseView c) {
    const Geo& G = c.geo;
    const int x = blockIdx.x * PBX + threadIdx.x;
    const int y = blockIdx.y * PBY + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= G.nx || y >= G.ny) return;
    float nx, ny, nz;
    normal_at(c, x, y, z, nx, ny, nz);
    const unsigned me = (unsigned)(z * G.nxy + y * G.nx + x);
    st32(c.gm_n[0], me, nx);
    st32(c.gm_n[1], me, ny);
    st32(c.gm_n[2], me, nz);
}

// --------------------------------------------------------------- k_collide

// f in moment form (collide_tile_fm): the collide pass's grid, held to 128
// VGPRs (two 8-wave workgroups per CU).
template <int TX, int TY, int MODE>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(4))) void k_collide_fm(CaseView c) {
    const Geo& G = c.geo;
    const unsigned gx = (unsigned)(G.nx + TX - 1) / TX, gy = (unsigned)(G.ny + TY - 1) / TY;
    unsigned t = xcd_order(blockIdx.x, gridDim.x);
    const int x0 = (int)(t % gx) * TX;
    t /= gx;
    const int y0 = (int)(t % gy) * TY;
    const int zb = (int)(t / gy) * c.zc;
    collide_tile_fm<TX, TY, MODE>(c, x0, y0, zb, min(zb + c.zc, G.nz));
}

template <int TX, int TY, int MODE>
__global__ __launch_bounds__(TX * TY) void k_collide(CaseView c) {
    const Geo& G = c.geo;
    // 1-D grid of gx*gy*gz tiles, x fastest, renumbered by xcd_order, or by
    // xcd_runs for populations in 64-cell blocks: with blocks, keeping the
    // whole device on one contiguous stretch of the lattice matters more than
    // the halo lines each XCD's own range shares (collide at 256^3: 0.545 ms
    // renumbered by xcd_order, 0.521 in dispatch order, 0.517 in runs of 8;
    // at 128^3, where every tile is resident at once, dispatch order was 4%
    // slower than xcd_order and runs of 8 as fast).
    const unsigned gx = (unsigned)(G.nx + TX - 1) / TX, gy = (unsigned)(G.ny + TY - 1) / TY;
    unsigned t = G.blocked ? xcd_runs(blockIdx.x, gridDim.x, kCollideRun) : xcd_order(blockIdx.x, gridDim.x);
    const int x0 = (int)(t % gx) * TX;
    t /= gx;
    const int y0 = (int)(t % gy) * TY;
    const int zb = (int)(t / gy) * c.zc;
    collide_tile<TX, TY, MODE, StoreF32>(c, x0, y0, zb, min(zb + c.zc, G.nz));
}

// ------------------------------------------------------------ init / shift

__global__ __launch_bounds__(PBX * PBY) void k_init_pull(const float* __restrict__ phi, float* __restrict__ f,
                                                         float* __restrict__ g, Geo G, Phys p) {
    const int x = blockIdx.x * PBX + threadIdx.x;
    const int y = blockIdx.y * PBY + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= G.nx || y >= G.ny) return;
    const int xs[3] = {shift<-1>(x, G.nx), x, shift<1>(x, G.nx)};
    const int ys[3] = {shift<-1>(y, G.ny) * G.nx, y * G.nx, shift<1>(y, G.ny) * G.nx};
    const int zs[3] = {shift<-1>(z, G.nz) * G.nxy, z * G.nxy, shift<1>(z, G.nz) * G.nxy};
    const int me = zs[1] + ys[1] + xs[1];
    for_links([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        const float w = p.w[WC[l]];
        f[pop_index(G, G.blocked, l, me)] = w * 1.0f;
        if (g) g[pop_index(G, G.blocked, l, me)] = w * phi[zs[1 + CZ[l]] + ys[1 + CY[l]] + xs[1 + CX[l]]];
    });
}

// Moment form -> the reference's post-stream g layout:
// dst_l(x) = g*_l(x - c_l) recomputed from phi and u, with n recomputed from
// phi (or w_l phi0(x) before the first step).
__global__ __launch_bounds__(PBX * PBY) void k_gm_to_ref(CaseView c, float* __restrict__ dst) {
    const Geo& G = c.geo;
    const int x = blockIdx.x * PBX + threadIdx.x;
    const int y = blockIdx.y * PBY + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= G.nx || y >= G.ny) return;
    const unsigned xs[3] = {(unsigned)shift<-1>(x, G.nx), (unsigned)x, (unsigned)shift<1>(x, G.nx)};
    const unsigned ys[3] = {(unsigned)(shift<-1>(y, G.ny) * G.nx), (unsigned)(y * G.nx),
                            (unsigned)(shift<1>(y, G.ny) * G.nx)};
    const unsigned zs[3] = {(unsigned)(shift<-1>(z, G.nz) * G.nxy), (unsigned)(z * G.nxy),
                            (unsigned)(shift<1>(z, G.nz) * G.nxy)};
    const unsigned me = zs[1] + ys[1] + xs[1];
    for_links([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        float v;
        if (c.gm == GM_INITIAL) v = c.phys.w[WC[l]] * ld32(c.phi_prev, me);
        else v = g_from_moments<l>(c, x - CX[l], y - CY[l], z - CZ[l]);
        dst[(size_t)l * G.plane + me] = v;
    });
}

// f in moment form: the initial state, rho = 1 and the other 12 moments 0.
__global__ __launch_bounds__(256) void k_fm_init(float* __restrict__ fm, long long plane) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < kFm * plane; i += (long long)gridDim.x * 256)
        fm[i] = i < plane ? 1.0f : 0.0f;
}

// f in moment form -> the reference's post-stream f layout:
// dst_l(x) = f*_l(x - c_l), collisionFluid of the source cell's moments.
__global__ __launch_bounds__(PBX * PBY) void k_fm_to_ref(CaseView c, float* __restrict__ dst) {
    const Geo& G = c.geo;
    const int x = blockIdx.x * PBX + threadIdx.x;
    const int y = blockIdx.y * PBY + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= G.nx || y >= G.ny) return;
    const Derived d(c.phys);
    const float* m = static_cast<const float*>(c.fsrc);
    const unsigned me = (unsigned)(z * G.nxy + y * G.nx + x);
    for_links([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        const unsigned src = (unsigned)(wrapmod(z - CZ[l], G.nz) * G.nxy + wrapmod(y - CY[l], G.ny) * G.nx +
                                        wrapmod(x - CX[l], G.nx));
        float v[kFm];
#pragma unroll
        for (int k = 0; k < kFm; ++k) v[k] = m[(size_t)k * G.plane + src];
        dst[(size_t)l * G.plane + me] = fluid_post<l>(fluid_cell_of(v), c.phys, d);
    });
}

// dst_l(x) = src_l(x - SIGN c_l), src in layout sblk and dst in dblk.
template <int SIGN>
__global__ __launch_bounds__(PBX * PBY) void k_link_shift(const float* __restrict__ src, int sblk,
                                                          float* __restrict__ dst, int dblk, Geo G) {
    const int x = blockIdx.x * PBX + threadIdx.x;
    const int y = blockIdx.y * PBY + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= G.nx || y >= G.ny) return;
    const int xs[3] = {shift<-1>(x, G.nx), x, shift<1>(x, G.nx)};
    const int ys[3] = {shift<-1>(y, G.ny) * G.nx, y * G.nx, shift<1>(y, G.ny) * G.nx};
    const int zs[3] = {shift<-1>(z, G.nz) * G.nxy, z * G.nxy, shift<1>(z, G.nz) * G.nxy};
    for_links([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        dst[pop_index(G, dblk, l, zs[1] + ys[1] + xs[1])] =
            src[pop_index(G, sblk, l, zs[1 - SIGN * CZ[l]] + ys[1 - SIGN * CY[l]] + xs[1 - SIGN * CX[l]])];
    });
}

// dst (layout dblk) = src (layout sblk): the 19 links of every cell, as raw
// elements of T (4-byte FP32 or 2-byte 16-bit populations).
template <class T>
__global__ __launch_bounds__(256) void k_pop_relayout(const T* __restrict__ src, int sblk, T* __restrict__ dst,
                                                      int dblk, Geo G) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < G.n; i += (long long)gridDim.x * 256)
        for_links([&](auto lc) {
            constexpr int l = decltype(lc)::value;
            dst[pop_index(G, dblk, l, i)] = src[pop_index(G, sblk, l, i)];
        });
}

// --------------------------------------------------------------- reductions

constexpr int RB = 256;

__global__ __launch_bounds__(RB) void k_reduce_partial(const float* __restrict__ phi, const float* __restrict__ rho,
                                                       const float* __restrict__ ux, const float* __restrict__ uy,
                                                       const float* __restrict__ uz, long long n,
                                                       double* __restrict__ partial) {
    double sp = 0.0, ke = 0.0;
    for (long long i = (long long)blockIdx.x * RB + threadIdx.x; i < n; i += (long long)gridDim.x * RB) {
        sp += (double)phi[i];
        const double vx = ux[i], vy = uy[i], vz = uz[i];
        ke += 0.5 * (double)rho[i] * (vx * vx + vy * vy + vz * vz);
    }
    // Butterfly within the wave64, then the four waves in order.
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        sp += __shfl_xor(sp, off, 64);
        ke += __shfl_xor(ke, off, 64);
    }
    __shared__ double sh[2][RB / 64];
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0) {
        sh[0][wave] = sp;
        sh[1][wave] = ke;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int w = 0; w < RB / 64; ++w) {
            a += sh[0][w];
            b += sh[1][w];
        }
        partial[2 * blockIdx.x] = a;
        partial[2 * blockIdx.x + 1] = b;
    }
}

__global__ void k_reduce_final(const double* __restrict__ partial, int nb, double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double a = 0.0, b = 0.0;
    for (int i = 0; i < nb; ++i) {
        a += partial[2 * i];
        b += partial[2 * i + 1];
    }
    out[0] = a;
    out[1] = b;
}

// Plain streaming copy, 16 bytes per lane, grid-stride. Used to measure the
// HBM bandwidth a simple kernel reaches and to calibrate byte counters.
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, long long n4) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n4; i += (long long)gridDim.x * 256) dst[i] = src[i];
}

// Periodic initial phase field of a set of bubbles: bubble_phi_cell, the
// same bits as the host initialisation (src/init.cu:27-49 with minimum-image
// distances).
__global__ __launch_bounds__(PBX * PBY) void k_bubble_phi(float* __restrict__ phi, Geo G, BubbleSet s) {
    const int i = blockIdx.x * PBX + threadIdx.x;
    const int j = blockIdx.y * PBY + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= G.nx || j >= G.ny) return;
    phi[(size_t)k * G.nxy + (size_t)j * G.nx + i] = bubble_phi_cell(i, j, k, G.nx, G.ny, G.nz, true, s.b, s.n, s.width);
}

// FP64 sum of a float field: per-block partial sums in a fixed order, then
// one thread adds the partials in order into *out.
__global__ __launch_bounds__(256) void k_sum_partial(const float* __restrict__ v, long long n,
                                                     double* __restrict__ partial) {
    double s = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        s += (double)v[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    __shared__ double sh[4];
    if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = s;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

__global__ void k_sum_final(const double* __restrict__ partial, int nb, double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (int i = 0; i < nb; ++i) s += partial[i];
    *out = s;
}

__global__ __launch_bounds__(256) void k_add_scalar(float* __restrict__ f, long long n, const double* __restrict__ d) {
    const float v = (float)d[0];
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) f[i] += v;
}

dim3 cell_grid(const Geo& G) { return dim3((G.nx + PBX - 1) / PBX, (G.ny + PBY - 1) / PBY, G.nz); }

#define MCB_TILE_ENTRY(TX, TY) {TX, TY, 0},
const Tiling kTilings[] = {MCB_TILES_FP32(MCB_TILE_ENTRY)};
#undef MCB_TILE_ENTRY

// Tile shapes of the f-moment collide pass (the FP32 tiles choose_tiling
// picks: 64x8, 32x8, 16x16; and 16x8).
#define MCB_TILES_FM(X) X(64, 8) X(32, 8) X(32, 16) X(16, 16) X(16, 8)

template <int TX, int TY>
hipError_t launch_collide_t(const CaseView& c, int mode, hipStream_t s) {
    const unsigned tiles = (unsigned)((c.geo.nx + TX - 1) / TX) * (unsigned)((c.geo.ny + TY - 1) / TY) *
                           (unsigned)((c.geo.nz + c.zc - 1) / c.zc);
    dim3 grid(tiles, 1, 1), block(TX, TY, 1);
    if (c.fm) {
#define MCB_FM_CASE(FX, FY)                                                                                   \
    if constexpr (TX == FX && TY == FY) {                                                                     \
        if (mode == COLLIDE_STEP) hipLaunchKernelGGL((k_collide_fm<TX, TY, COLLIDE_STEP>), grid, block, 0, s, c); \
        else if (mode == COLLIDE_MACROS)                                                                      \
            hipLaunchKernelGGL((k_collide_fm<TX, TY, COLLIDE_MACROS>), grid, block, 0, s, c);                 \
        else return hipErrorInvalidValue;                                                                     \
        return hipGetLastError();                                                                             \
    }
        MCB_TILES_FM(MCB_FM_CASE)
#undef MCB_FM_CASE
        return hipErrorInvalidValue;
    }
    switch (mode) {
        case COLLIDE_STEP: hipLaunchKernelGGL((k_collide<TX, TY, COLLIDE_STEP>), grid, block, 0, s, c); break;
        case COLLIDE_STEP_BGK: hipLaunchKernelGGL((k_collide<TX, TY, COLLIDE_STEP_BGK>), grid, block, 0, s, c); break;
        case COLLIDE_MACROS: hipLaunchKernelGGL((k_collide<TX, TY, COLLIDE_MACROS>), grid, block, 0, s, c); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int TX, int TY>
hipError_t launch_batch_t(const CaseView* cases, const WorkItem* work, unsigned n, int mode, bool phase,
                          hipStream_t s) {
    dim3 grid(n, 1, 1), block(TX, TY, 1);
    if (phase) {
        hipLaunchKernelGGL((k_phase_batch<TX, TY, StoreF32>), grid, block, 0, s, cases, work);
        return hipGetLastError();
    }
    switch (mode) {
        case COLLIDE_STEP:
            hipLaunchKernelGGL((k_collide_batch<TX, TY, COLLIDE_STEP, StoreF32>), grid, block, 0, s, cases, work);
            break;
        case COLLIDE_STEP_BGK:
            hipLaunchKernelGGL((k_collide_batch<TX, TY, COLLIDE_STEP_BGK, StoreF32>), grid, block, 0, s, cases, work);
            break;
        case COLLIDE_MACROS:
            hipLaunchKernelGGL((k_collide_batch<TX, TY, COLLIDE_MACROS, StoreF32>), grid, block, 0, s, cases, work);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_batch(const CaseView* cases, const WorkItem* work, unsigned n, const Tiling& t, int mode,
                        bool phase, hipStream_t s) {
    if (n == 0) return hipSuccess;
#define MCB_TILE_CASE(TX, TY) \
    if (t.tx == TX && t.ty == TY) return launch_batch_t<TX, TY>(cases, work, n, mode, phase, s);
    MCB_TILES_FP32(MCB_TILE_CASE)
#undef MCB_TILE_CASE
    return hipErrorInvalidValue;
}

}  // namespace

bool fm_tiling_supported(const Tiling& t) {
#define MCB_FM_ENTRY(FX, FY) if (t.tx == FX && t.ty == FY) return t.zc >= 1;
    MCB_TILES_FM(MCB_FM_ENTRY)
#undef MCB_FM_ENTRY
    return false;
}

bool tiling_supported(const Tiling& t) {
    for (const auto& k : kTilings)
        if (k.tx == t.tx && k.ty == t.ty) return t.zc >= 1;
    return false;
}

int tiling_count() { return (int)(sizeof(kTilings) / sizeof(kTilings[0])); }
Tiling tiling_at(int i) { return kTilings[i]; }

Tiling phase_tiling(const Geo& geo, const Tiling& t, bool rn) {
    Tiling p = t;
    if (rn && geo.nx >= MCB_PHASE_RN_TX && geo.ny >= MCB_PHASE_RN_TY) {
        p.tx = MCB_PHASE_RN_TX;
        p.ty = MCB_PHASE_RN_TY;
    }
    // Chunks of 64 planes, halved (down to 8) until the pass has two
    // workgroups per CU.
    const unsigned layer = (unsigned)((geo.nx + p.tx - 1) / p.tx) * (unsigned)((geo.ny + p.ty - 1) / p.ty);
    p.zc = std::min(64, geo.nz);
    while (p.zc > 8 && layer * (unsigned)((geo.nz + p.zc - 1) / p.zc) < 512u) p.zc = (p.zc + 1) / 2;
    return p;
}

bool phase_recomputes_n(const Geo& geo, const Tiling& t) { return phase_tiling(geo, t, true).zc >= 32; }

hipError_t launch_phase(const CaseView& c0, const Tiling& t, hipStream_t s) {
    if (c0.gm == GM_MOMENTS) {
        CaseView c = c0;
        const bool rn = c.gm_n[0] == nullptr;
        const Tiling p = phase_tiling(c.geo, t, rn);
        c.zc = p.zc;
        const unsigned tiles = (unsigned)((c.geo.nx + p.tx - 1) / p.tx) * (unsigned)((c.geo.ny + p.ty - 1) / p.ty) *
                               (unsigned)((c.geo.nz + c.zc - 1) / c.zc);
        if (rn && p.tx == MCB_PHASE_RN_TX && p.ty == MCB_PHASE_RN_TY) {
            hipLaunchKernelGGL((k_phase_gm<MCB_PHASE_RN_TX, MCB_PHASE_RN_TY, true>), dim3(tiles),
                               dim3(MCB_PHASE_RN_TX, MCB_PHASE_RN_TY, 1), 0, s, c);
            return hipGetLastError();
        }
#define MCB_PHASE_GM(TX, TY)                                                                                     \
    if (p.tx == TX && p.ty == TY) {                                                                              \
        if (rn) hipLaunchKernelGGL((k_phase_gm<TX, TY, true>), dim3(tiles), dim3(TX, TY, 1), 0, s, c);          \
        else hipLaunchKernelGGL((k_phase_gm<TX, TY, false>), dim3(tiles), dim3(TX, TY, 1), 0, s, c);            \
        return hipGetLastError();                                                                                \
    }
        MCB_TILES_FP32(MCB_PHASE_GM)
#undef MCB_PHASE_GM
        return hipErrorInvalidValue;
    }
    const CaseView& c = c0;
    const unsigned gx = (unsigned)(c.geo.nx + PBX - 1) / PBX, gy = (unsigned)(c.geo.ny + PBY - 1) / PBY;
    hipLaunchKernelGGL(k_phase, dim3(gx * gy * (unsigned)c.geo.nz), dim3(PBX, PBY, 1), 0, s, c, gx, gy);
    return hipGetLastError();
}

hipError_t launch_collide(const CaseView& c0, int mode, const Tiling& t, hipStream_t s) {
    CaseView c = c0;
    c.zc = t.zc;
#define MCB_TILE_CASE(TX, TY) \
    if (t.tx == TX && t.ty == TY) return launch_collide_t<TX, TY>(c, mode, s);
    MCB_TILES_FP32(MCB_TILE_CASE)
#undef MCB_TILE_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_phase_batch(const CaseView* cases, const WorkItem* work, unsigned n_work, const Tiling& t,
                              int storage, hipStream_t s) {
    if (storage != MCB_STORE_FP32) return launch_phase_batch16(cases, work, n_work, t, storage, s);
    return launch_batch(cases, work, n_work, t, 0, true, s);
}

hipError_t launch_collide_batch(const CaseView* cases, const WorkItem* work, unsigned n_work, const Tiling& t,
                                int mode, int storage, hipStream_t s) {
    if (storage != MCB_STORE_FP32) return launch_collide_batch16(cases, work, n_work, t, mode, storage, s);
    return launch_batch(cases, work, n_work, t, mode, false, s);
}

hipError_t launch_init_pull(const float* phi, float* f, float* g, const Geo& geo, const Phys& p, hipStream_t s) {
    hipLaunchKernelGGL(k_init_pull, cell_grid(geo), dim3(PBX, PBY, 1), 0, s, phi, f, g, geo, p);
    return hipGetLastError();
}

hipError_t launch_fm_init(float* fm, const Geo& geo, hipStream_t s) {
    const long long blocks = std::min<long long>((kFm * geo.plane + 255) / 256, 256LL * 16);
    hipLaunchKernelGGL(k_fm_init, dim3((unsigned)blocks), dim3(256), 0, s, fm, geo.plane);
    return hipGetLastError();
}

hipError_t launch_fm_to_ref(const CaseView& c, float* dst, hipStream_t s) {
    if (!c.fm) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_fm_to_ref, cell_grid(c.geo), dim3(PBX, PBY, 1), 0, s, c, dst);
    return hipGetLastError();
}

hipError_t launch_normals(const CaseView& c, hipStream_t s) {
    if (!c.gm_n[0]) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_normals, cell_grid(c.geo), dim3(PBX, PBY, 1), 0, s, c);
    return hipGetLastError();
}

hipError_t launch_gm_to_ref(const CaseView& c, float* dst, hipStream_t s) {
    if (c.gm == GM_POPULATIONS) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_gm_to_ref, cell_grid(c.geo), dim3(PBX, PBY, 1), 0, s, c, dst);
    return hipGetLastError();
}

hipError_t launch_link_shift(const float* src, int sblk, float* dst, int dblk, const Geo& geo, int sign,
                             hipStream_t s) {
    const dim3 grid = cell_grid(geo), block(PBX, PBY, 1);
    if (sign > 0) hipLaunchKernelGGL(k_link_shift<1>, grid, block, 0, s, src, sblk, dst, dblk, geo);
    else hipLaunchKernelGGL(k_link_shift<-1>, grid, block, 0, s, src, sblk, dst, dblk, geo);
    return hipGetLastError();
}

hipError_t launch_pop_relayout(const void* src, int sblk, void* dst, int dblk, const Geo& geo, unsigned elem_bytes,
                               hipStream_t s) {
    const unsigned blocks = (unsigned)std::min<long long>((geo.n + 255) / 256, 256LL * 16);
    if (elem_bytes == 4)
        hipLaunchKernelGGL(k_pop_relayout<unsigned>, dim3(blocks), dim3(256), 0, s, static_cast<const unsigned*>(src),
                           sblk, static_cast<unsigned*>(dst), dblk, geo);
    else
        hipLaunchKernelGGL(k_pop_relayout<unsigned short>, dim3(blocks), dim3(256), 0, s,
                           static_cast<const unsigned short*>(src), sblk, static_cast<unsigned short*>(dst), dblk, geo);
    return hipGetLastError();
}

hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s) {
    const long long n4 = (long long)(bytes / 16);
    const long long blocks = std::min<long long>((n4 + 255) / 256, 256LL * 32);
    hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, s, static_cast<const float4*>(src),
                       static_cast<float4*>(dst), n4);
    return hipGetLastError();
}

hipError_t launch_bubble_phi(float* phi, const Geo& geo, const BubbleSet& bubbles, hipStream_t s) {
    if (bubbles.n < 1 || bubbles.n > kMaxBubbleSet) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_bubble_phi, cell_grid(geo), dim3(PBX, PBY, 1), 0, s, phi, geo, bubbles);
    return hipGetLastError();
}

hipError_t launch_sum(const float* v, long long n, double* partial, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_sum_partial, dim3(kReduceBlocks), dim3(256), 0, s, v, n, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(64), 0, s, partial, kReduceBlocks, out);
    return hipGetLastError();
}

hipError_t launch_add_scalar(float* field, long long n, const double* delta, hipStream_t s) {
    const long long blocks = std::min<long long>((n + 255) / 256, 256LL * 16);
    hipLaunchKernelGGL(k_add_scalar, dim3((unsigned)blocks), dim3(256), 0, s, field, n, delta);
    return hipGetLastError();
}

hipError_t launch_reduce(const float* phi, const float* rho, const float* ux, const float* uy, const float* uz,
                         long long n, double* partial, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_partial, dim3(kReduceBlocks), dim3(RB), 0, s, phi, rho, ux, uy, uz, n, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(64), 0, s, partial, kReduceBlocks, out);
    return hipGetLastError();
}

}  // namespace mcb
