// Operator face of libmcb (include/mcb_ops.h): one HIP kernel per reference
// kernel of src/kernels.cuh, on the reference's SoA layout. Thread per cell,
// 64-wide rows along x (one wave64 per row segment) instead of the reference's
// 8x8x8 blocks, so every population plane is read and written in 256-byte
// runs. The per-cell arithmetic lives in lattice.hpp and is shared with the
// fused step kernels.
#include "bubble.hpp"
#include "common.hpp"

namespace mcb {
namespace {

constexpr int OBX = 64, OBY = 4;

struct OpArgs {
    Geo geo;
    Phys phys;
    int mode;        // mcb_bc_mode
    int tau_g_one;
};

__device__ __forceinline__ bool is_interior(int i, int j, int k, const Geo& g) {
    return !(i == 0 || i == g.nx - 1 || j == 0 || j == g.ny - 1 || k == 0 || k == g.nz - 1);
}

// Cell coordinates of this thread; false when outside the lattice or, in
// REFERENCE mode with interior_only, on a face (lbm.cu:14 and its copies).
__device__ __forceinline__ bool cell_of(const OpArgs& a, bool interior_only, int& i, int& j, int& k) {
    i = blockIdx.x * OBX + threadIdx.x;
    j = blockIdx.y * OBY + threadIdx.y;
    k = blockIdx.z;
    if (i >= a.geo.nx || j >= a.geo.ny || k >= a.geo.nz) return false;
    if (interior_only && a.mode == MCB_BC_REFERENCE && !is_interior(i, j, k, a.geo)) return false;
    return true;
}

__device__ __forceinline__ long long lin(int i, int j, int k, const Geo& g) {
    return (long long)k * g.nxy + (long long)j * g.nx + i;
}

template <int L> __device__ __forceinline__ long long nbr(int i, int j, int k, const Geo& g) {
    return lin(shift<CX[L]>(i, g.nx), shift<CY[L]>(j, g.ny), shift<CZ[L]>(k, g.nz), g);
}

// ------------------------------------------------------------------ init.cu

__global__ __launch_bounds__(256) void k_initTensor(float* pxx, float* pyy, float* pzz, float* pxy, float* pxz,
                                                    float* pyz, float* rho, OpArgs a) {
    int i, j, k;
    if (!cell_of(a, false, i, j, k)) return;
    const long long x = lin(i, j, k, a.geo);
    const float val = 1.0f;
    pxx[x] = val; pyy[x] = val; pzz[x] = val;
    pxy[x] = val; pxz[x] = val; pyz[x] = val;
    rho[x] = val;
}

struct BubbleArgs {
    int n;
    float b[4 * MCB_OPS_MAX_BUBBLES];
    float width;
};

// bubble_phi_cell: the same bits as the host initialisation.
__global__ __launch_bounds__(256) void k_initPhase(float* phi, OpArgs a, BubbleArgs bb) {
    int i, j, k;
    if (!cell_of(a, true, i, j, k)) return;
    static_assert(sizeof(mcb_bubble) == 4 * sizeof(float), "bubble record = cx, cy, cz, r");
    phi[lin(i, j, k, a.geo)] = bubble_phi_cell(i, j, k, a.geo.nx, a.geo.ny, a.geo.nz, a.mode == MCB_BC_PERIODIC,
                                               reinterpret_cast<const mcb_bubble*>(bb.b), bb.n, bb.width);
}

__global__ __launch_bounds__(256) void k_initDist(const float* rho, const float* phi, float* f, float* g, OpArgs a) {
    int i, j, k;
    if (!cell_of(a, false, i, j, k)) return;
    const long long x = lin(i, j, k, a.geo);
    const float r = rho[x], ph = phi[x];
    for_links([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        f[x + l * a.geo.plane] = a.phys.w[WC[l]] * r;
    });
    for_links([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        g[x + l * a.geo.plane] = a.phys.w[WC[l]] * ph;
    });
}

// ------------------------------------------------------------------- lbm.cu

__global__ __launch_bounds__(256) void k_phiCalc(float* phi, const float* g, OpArgs a) {
    int i, j, k;
    if (!cell_of(a, true, i, j, k)) return;
    const long long x = lin(i, j, k, a.geo);
    float sum = 0.0f;
    for_links([&](auto lc) { sum += g[x + decltype(lc)::value * a.geo.plane]; });
    phi[x] = sum;
}

__global__ __launch_bounds__(256) void k_gradCalc(const float* phi, float* normx, float* normy, float* normz,
                                                  float* indicator, OpArgs a) {
    int i, j, k;
    if (!cell_of(a, true, i, j, k)) return;
    const Derived d(a.phys);
    float nx, ny, nz, ind;
    grad_normal([&](auto lc) { return phi[nbr<decltype(lc)::value>(i, j, k, a.geo)]; }, d, nx, ny, nz, ind);
    const long long x = lin(i, j, k, a.geo);
    normx[x] = nx; normy[x] = ny; normz[x] = nz; indicator[x] = ind;
}

__global__ __launch_bounds__(256) void k_curvatureCalc(float* curvature, const float* indicator, const float* normx,
                                                       const float* normy, const float* normz, float* ffx, float* ffy,
                                                       float* ffz, OpArgs a) {
    int i, j, k;
    if (!cell_of(a, true, i, j, k)) return;
    const Derived d(a.phys);
    const long long x = lin(i, j, k, a.geo);
    float curv, fx, fy, fz;
    curvature_force(
        [&](auto lc, float& ax, float& ay, float& az) {
            constexpr int l = decltype(lc)::value;
            const long long o = nbr<l>(i, j, k, a.geo);
            if constexpr (CX[l] != 0) ax = normx[o];
            if constexpr (CY[l] != 0) ay = normy[o];
            if constexpr (CZ[l] != 0) az = normz[o];
        },
        a.phys, d, normx[x], normy[x], normz[x], indicator[x], curv, fx, fy, fz);
    curvature[x] = curv;
    ffx[x] = fx; ffy[x] = fy; ffz[x] = fz;
}

__global__ __launch_bounds__(256) void k_momentiCalc(float* ux, float* uy, float* uz, float* rho, const float* ffx,
                                                     const float* ffy, const float* ffz, const float* f, float* pxx,
                                                     float* pyy, float* pzz, float* pxy, float* pxz, float* pyz,
                                                     OpArgs a) {
    int i, j, k;
    if (!cell_of(a, true, i, j, k)) return;
    const long long x = lin(i, j, k, a.geo);
    float fv[Q];
    for_links([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        fv[l] = f[x + l * a.geo.plane];
    });
    const Moments m = moments(fv, a.phys, ffx[x], ffy[x], ffz[x]);
    pxx[x] = m.pxx; pyy[x] = m.pyy; pzz[x] = m.pzz;
    pxy[x] = m.pxy; pxz[x] = m.pxz; pyz[x] = m.pyz;
    ux[x] = m.ux; uy[x] = m.uy; uz[x] = m.uz;
    rho[x] = m.rho;
}

__global__ __launch_bounds__(256) void k_collisionFluid(float* f, const float* ux, const float* uy, const float* uz,
                                                        const float* ffx, const float* ffy, const float* ffz,
                                                        const float* rho, const float* pxx, const float* pyy,
                                                        const float* pzz, const float* pxy, const float* pxz,
                                                        const float* pyz, OpArgs a) {
    int i, j, k;
    if (!cell_of(a, true, i, j, k)) return;
    const Derived d(a.phys);
    const long long x = lin(i, j, k, a.geo);
    Moments m;
    m.rho = rho[x]; m.ux = ux[x]; m.uy = uy[x]; m.uz = uz[x];
    m.pxx = pxx[x]; m.pyy = pyy[x]; m.pzz = pzz[x];
    m.pxy = pxy[x]; m.pxz = pxz[x]; m.pyz = pyz[x];
    const FluidCell c = fluid_cell(m, ffx[x], ffy[x], ffz[x]);
    for_links([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        f[nbr<l>(i, j, k, a.geo) + l * a.geo.plane] = fluid_post<l>(c, a.phys, d);
    });
}

__global__ __launch_bounds__(256) void k_collisionPhase(float* g, const float* ux, const float* uy, const float* uz,
                                                        const float* phi, const float* normx, const float* normy,
                                                        const float* normz, OpArgs a) {
    int i, j, k;
    if (!cell_of(a, true, i, j, k)) return;
    const Derived d(a.phys);
    const long long x = lin(i, j, k, a.geo);
    const PhaseCell c = phase_cell(a.phys, phi[x], ux[x], uy[x], uz[x], normx[x], normy[x], normz[x]);
    if (a.tau_g_one) {
        for_links([&](auto lc) {
            constexpr int l = decltype(lc)::value;
            g[x + l * a.geo.plane] = phase_post<l>(c, a.phys);
        });
    } else {
        for_links([&](auto lc) {
            constexpr int l = decltype(lc)::value;
            const long long o = x + l * a.geo.plane;
            g[o] = phase_post_bgk<l>(c, a.phys, d, g[o]);
        });
    }
}

__global__ __launch_bounds__(256) void k_streamingCalc(const float* g, float* g_out, OpArgs a) {
    int i, j, k;
    if (!cell_of(a, false, i, j, k)) return;
    const long long x = lin(i, j, k, a.geo);
    for_links([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        const long long src = lin(shift<-CX[l]>(i, a.geo.nx), shift<-CY[l]>(j, a.geo.ny), shift<-CZ[l]>(k, a.geo.nz), a.geo);
        g_out[x + l * a.geo.plane] = g[src + l * a.geo.plane];
    });
}

__global__ __launch_bounds__(256) void k_fgBoundary(float* f, const float* rho, float* g, const float* phi, OpArgs a) {
    int i, j, k;
    if (!cell_of(a, false, i, j, k)) return;
    if (is_interior(i, j, k, a.geo)) return;
    const long long x = lin(i, j, k, a.geo);
    const float r = rho[x], ph = phi[x];
    auto inside = [&](int ni, int nj, int nk) {
        return ni >= 0 && ni < a.geo.nx && nj >= 0 && nj < a.geo.ny && nk >= 0 && nk < a.geo.nz;
    };
    for_links([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        const int ni = i + CX[l], nj = j + CY[l], nk = k + CZ[l];
        if (inside(ni, nj, nk)) f[lin(ni, nj, nk, a.geo) + l * a.geo.plane] = r * a.phys.w[WC[l]];
    });
    for_links([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        const int ni = i + CX[l], nj = j + CY[l], nk = k + CZ[l];
        if (inside(ni, nj, nk)) g[lin(ni, nj, nk, a.geo) + l * a.geo.plane] = ph * a.phys.w[WC[l]];
    });
}

// Zero-gradient phi on the six faces. Serial order z, y, x (each reading the
// previous result) leaves every face node equal to the interior node at the
// clamped coordinates, which a single pass can write without a race: face
// nodes only read interior nodes.
__global__ __launch_bounds__(256) void k_boundaryConditions(float* phi, OpArgs a) {
    int i, j, k;
    if (!cell_of(a, false, i, j, k)) return;
    if (is_interior(i, j, k, a.geo)) return;
    const int ci = min(max(i, 1), a.geo.nx - 2);
    const int cj = min(max(j, 1), a.geo.ny - 2);
    const int ck = min(max(k, 1), a.geo.nz - 2);
    phi[lin(i, j, k, a.geo)] = phi[lin(ci, cj, ck, a.geo)];
}

__global__ __launch_bounds__(256) void k_movingWallBC(float* f, const float* rho, OpArgs a) {
    // The reference maps (i, k) onto a 3-D launch and covers k < 8*ceil(NY/8);
    // one thread per (i, k) here, same stores.
    const int i = blockIdx.x * OBX + threadIdx.x;
    const int k = blockIdx.y * OBY + threadIdx.y;
    const int kmax = min(a.geo.nz, ((a.geo.ny + 7) / 8) * 8);
    if (i >= a.geo.nx || k >= kmax) return;
    const float U_wall = 0.03f;
    const long long xb = lin(i, 0, k, a.geo), xt = lin(i, a.geo.ny - 1, k, a.geo);
    const float rb = rho[xb], rt = rho[xt];
    for_links([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        const float cu = ((float)CX[l] * (-U_wall)) / CSSQ;
        f[xb + l * a.geo.plane] = rb * a.phys.w[WC[l]] * (1.0f + cu);
    });
    for_links([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        const float cu = ((float)CX[l] * U_wall) / CSSQ;
        f[xt + l * a.geo.plane] = rt * a.phys.w[WC[l]] * (1.0f + cu);
    });
}

// ------------------------------------------------------------ launch helpers

mcb_status check_common(int NX, int NY, int NZ, const mcb_op_consts* c) {
    if (!c) return MCB_ERR_INVALID_ARG;
    if (NX < 3 || NY < 3 || NZ < 3) return MCB_ERR_INVALID_ARG;
    if ((long long)NX * NY * NZ >= (1LL << 31)) return MCB_ERR_INVALID_ARG;
    if (NZ > 65535) return MCB_ERR_INVALID_ARG;
    if (c->bc_mode != MCB_BC_PERIODIC && c->bc_mode != MCB_BC_REFERENCE) return MCB_ERR_INVALID_ARG;
    return MCB_OK;
}

OpArgs make_args(int NX, int NY, int NZ, const mcb_op_consts* c) {
    OpArgs a;
    a.geo.nx = NX; a.geo.ny = NY; a.geo.nz = NZ;
    a.geo.nxy = NX * NY;
    a.geo.n = (long long)NX * NY * NZ;
    a.geo.plane = a.geo.n;   // the reference layout has no padding
    a.phys = phys_from_consts(*c);
    a.mode = c->bc_mode;
    a.tau_g_one = c->tau_g_is_one;
    return a;
}

dim3 grid_of(const Geo& g) { return dim3((g.nx + OBX - 1) / OBX, (g.ny + OBY - 1) / OBY, g.nz); }
const dim3 kBlock(OBX, OBY, 1);

mcb_status launched() { return hip_status(hipGetLastError()); }

template <class... P> bool any_null(P... p) { return ((p == nullptr) || ...); }

}  // namespace
}  // namespace mcb

using namespace mcb;

#define MCB_OP_PROLOGUE(...)                                          \
    mcb_status st_ = check_common(NX, NY, NZ, c);                     \
    if (st_ != MCB_OK) return st_;                                    \
    if (any_null(__VA_ARGS__)) return MCB_ERR_INVALID_ARG;            \
    const OpArgs a = make_args(NX, NY, NZ, c);                        \
    hipStream_t s = static_cast<hipStream_t>(stream)

extern "C" {

mcb_status mcb_initTensor(float* pxx, float* pyy, float* pzz, float* pxy, float* pxz, float* pyz, float* rho,
                          int NX, int NY, int NZ, const mcb_op_consts* c, void* stream) {
    MCB_OP_PROLOGUE(pxx, pyy, pzz, pxy, pxz, pyz, rho);
    hipLaunchKernelGGL(k_initTensor, grid_of(a.geo), kBlock, 0, s, pxx, pyy, pzz, pxy, pxz, pyz, rho, a);
    return launched();
}

mcb_status mcb_initPhase(float* phi, int NX, int NY, int NZ, const mcb_op_consts* c, void* stream) {
    MCB_OP_PROLOGUE(phi);
    if (c->n_bubbles < 1 || c->n_bubbles > MCB_OPS_MAX_BUBBLES || !(c->width > 0.0f)) return MCB_ERR_INVALID_ARG;
    BubbleArgs bb;
    bb.n = c->n_bubbles;
    for (int q = 0; q < 4 * MCB_OPS_MAX_BUBBLES; ++q) bb.b[q] = c->bubbles[q];
    bb.width = c->width;
    hipLaunchKernelGGL(k_initPhase, grid_of(a.geo), kBlock, 0, s, phi, a, bb);
    return launched();
}

mcb_status mcb_initDist(const float* rho, const float* phi, float* f, float* g, int NX, int NY, int NZ,
                        const mcb_op_consts* c, void* stream) {
    MCB_OP_PROLOGUE(rho, phi, f, g);
    hipLaunchKernelGGL(k_initDist, grid_of(a.geo), kBlock, 0, s, rho, phi, f, g, a);
    return launched();
}

mcb_status mcb_phiCalc(float* phi, const float* g, int NX, int NY, int NZ, const mcb_op_consts* c, void* stream) {
    MCB_OP_PROLOGUE(phi, g);
    hipLaunchKernelGGL(k_phiCalc, grid_of(a.geo), kBlock, 0, s, phi, g, a);
    return launched();
}

mcb_status mcb_gradCalc(const float* phi, float* normx, float* normy, float* normz, float* indicator, int NX, int NY,
                        int NZ, const mcb_op_consts* c, void* stream) {
    MCB_OP_PROLOGUE(phi, normx, normy, normz, indicator);
    hipLaunchKernelGGL(k_gradCalc, grid_of(a.geo), kBlock, 0, s, phi, normx, normy, normz, indicator, a);
    return launched();
}

mcb_status mcb_curvatureCalc(float* curvature, const float* indicator, const float* normx, const float* normy,
                             const float* normz, float* ffx, float* ffy, float* ffz, int NX, int NY, int NZ,
                             const mcb_op_consts* c, void* stream) {
    MCB_OP_PROLOGUE(curvature, indicator, normx, normy, normz, ffx, ffy, ffz);
    hipLaunchKernelGGL(k_curvatureCalc, grid_of(a.geo), kBlock, 0, s, curvature, indicator, normx, normy, normz, ffx,
                       ffy, ffz, a);
    return launched();
}

mcb_status mcb_momentiCalc(float* ux, float* uy, float* uz, float* rho, const float* ffx, const float* ffy,
                           const float* ffz, const float* f, float* pxx, float* pyy, float* pzz, float* pxy,
                           float* pxz, float* pyz, int NX, int NY, int NZ, const mcb_op_consts* c, void* stream) {
    MCB_OP_PROLOGUE(ux, uy, uz, rho, ffx, ffy, ffz, f, pxx, pyy, pzz, pxy, pxz, pyz);
    hipLaunchKernelGGL(k_momentiCalc, grid_of(a.geo), kBlock, 0, s, ux, uy, uz, rho, ffx, ffy, ffz, f, pxx, pyy, pzz,
                       pxy, pxz, pyz, a);
    return launched();
}

mcb_status mcb_collisionFluid(float* f, const float* ux, const float* uy, const float* uz, const float* ffx,
                              const float* ffy, const float* ffz, const float* rho, const float* pxx, const float* pyy,
                              const float* pzz, const float* pxy, const float* pxz, const float* pyz, int NX, int NY,
                              int NZ, const mcb_op_consts* c, void* stream) {
    MCB_OP_PROLOGUE(f, ux, uy, uz, ffx, ffy, ffz, rho, pxx, pyy, pzz, pxy, pxz, pyz);
    hipLaunchKernelGGL(k_collisionFluid, grid_of(a.geo), kBlock, 0, s, f, ux, uy, uz, ffx, ffy, ffz, rho, pxx, pyy,
                       pzz, pxy, pxz, pyz, a);
    return launched();
}

mcb_status mcb_collisionPhase(float* g, const float* ux, const float* uy, const float* uz, const float* phi,
                              const float* normx, const float* normy, const float* normz, int NX, int NY, int NZ,
                              const mcb_op_consts* c, void* stream) {
    MCB_OP_PROLOGUE(g, ux, uy, uz, phi, normx, normy, normz);
    hipLaunchKernelGGL(k_collisionPhase, grid_of(a.geo), kBlock, 0, s, g, ux, uy, uz, phi, normx, normy, normz, a);
    return launched();
}

mcb_status mcb_streamingCalc(const float* g, float* g_out, int NX, int NY, int NZ, const mcb_op_consts* c,
                             void* stream) {
    MCB_OP_PROLOGUE(g, g_out);
    if (g == g_out) return MCB_ERR_INVALID_ARG;
    hipLaunchKernelGGL(k_streamingCalc, grid_of(a.geo), kBlock, 0, s, g, g_out, a);
    return launched();
}

mcb_status mcb_fgBoundary(float* f, const float* rho, float* g, const float* phi, int NX, int NY, int NZ,
                          const mcb_op_consts* c, void* stream) {
    MCB_OP_PROLOGUE(f, rho, g, phi);
    hipLaunchKernelGGL(k_fgBoundary, grid_of(a.geo), kBlock, 0, s, f, rho, g, phi, a);
    return launched();
}

mcb_status mcb_boundaryConditions(float* phi, int NX, int NY, int NZ, const mcb_op_consts* c, void* stream) {
    MCB_OP_PROLOGUE(phi);
    hipLaunchKernelGGL(k_boundaryConditions, grid_of(a.geo), kBlock, 0, s, phi, a);
    return launched();
}

mcb_status mcb_movingWallBC(float* f, const float* rho, int NX, int NY, int NZ, const mcb_op_consts* c,
                            void* stream) {
    MCB_OP_PROLOGUE(f, rho);
    const int kmax = std::min(NZ, ((NY + 7) / 8) * 8);
    dim3 grid((NX + OBX - 1) / OBX, (kmax + OBY - 1) / OBY, 1);
    hipLaunchKernelGGL(k_movingWallBC, grid, kBlock, 0, s, f, rho, a);
    return launched();
}

}  // extern "C"
