// Fused two-pass periodic step (see fused.hpp for the storage convention).
//
// k_phase   : one wave64 per 64 consecutive x cells; 19 gathered loads (a
//             row shifted by at most one cell) and one aligned store.
// k_collide : a workgroup owns a TX x TY column of cells and marches it
//             through zc planes. phi planes (with a 2-cell halo) live in a
//             4-slot LDS ring, the normals (1-cell halo) in a 3-slot ring, so
//             each phi value is fetched from HBM about (1 + 4/zc) times
//             and the gradient is recomputed only on the halo ring.
//             Populations are gathered straight from HBM (each read once)
//             and stored aligned.
#include "fused.hpp"

namespace mcb {
namespace {

__device__ __forceinline__ int wrapmod(int v, int n) {
    v %= n;
    return v < 0 ? v + n : v;
}

template <int C, class T> __device__ __forceinline__ T pick(T m, T z, T p) {
    if constexpr (C < 0) return m;
    else if constexpr (C > 0) return p;
    else return z;
}

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup i runs on
// XCD i % 8), and each XCD has its own L2. xcd_order renumbers them so every
// XCD walks one contiguous range of tiles: x-neighbouring tiles, which share
// the cache lines that the x-shifted gathers only partly use, then meet in one
// L2 instead of fetching those lines from HBM twice. Bijective on [0, total).
constexpr unsigned kXcds = 8;
__device__ __forceinline__ unsigned xcd_order(unsigned lid, unsigned total) {
    const unsigned q = total / kXcds, r = total % kXcds;
    const unsigned xcd = lid % kXcds, slot = lid / kXcds;
    return (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + slot;
}

// ----------------------------------------------------------------- k_phase

constexpr int PBX = 64, PBY = 4;

// 1-D grid of gx*gy*nz workgroups of 64x4 cells, renumbered by xcd_order.
__global__ __launch_bounds__(PBX * PBY) void k_phase(const float* __restrict__ g, float* __restrict__ phi, Geo G,
                                                     unsigned gx, unsigned gy) {
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
    const Planes gp(g, (unsigned)G.plane);
    float sum = 0.0f;
    for_links([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        sum += gp.ld<l>(zs[1 - CZ[l]] + ys[1 - CY[l]] + xs[1 - CX[l]]);
    });
    st32(phi, zs[1] + ys[1] + xs[1], sum);
}

// --------------------------------------------------------------- k_collide

template <int TX, int TY, int MODE>
__global__ __launch_bounds__(TX * TY) void k_collide(CollideArgs a) {
    constexpr int NT = TX * TY;
    constexpr int PX = TX + 4, PY = TY + 4, PN = PX * PY;  // phi plane + 2-halo
    constexpr int HX = TX + 2, HY = TY + 2, HN = HX * HY;  // normal plane + 1-halo
    constexpr int KP = (PN + NT - 1) / NT;
    constexpr int KH = (HN + NT - 1) / NT;
    __shared__ float s_phi[4][PN];
    __shared__ float s_n[3][3][HN];
    __shared__ float s_ind[2][NT];

    const Geo& G = a.geo;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
    // 1-D grid of gx*gy*gz tiles, x fastest, renumbered by xcd_order.
    const unsigned gx = (unsigned)(G.nx + TX - 1) / TX, gy = (unsigned)(G.ny + TY - 1) / TY;
    unsigned t = xcd_order(blockIdx.x, gridDim.x);
    const int x0 = (int)(t % gx) * TX;
    t /= gx;
    const int y0 = (int)(t % gy) * TY;
    const int zb = (int)(t / gy) * a.zc;
    const int ze = min(zb + a.zc, G.nz);
    const Derived d(a.phys);
    const Planes fsrc(a.fsrc, (unsigned)G.plane), fdst(a.fdst, (unsigned)G.plane);
    const Planes gsrc(a.gsrc, (unsigned)G.plane), gdst(a.gdst, (unsigned)G.plane);

    // In-plane offsets of the halo phi elements this thread stages.
    unsigned poff[KP];
#pragma unroll
    for (int q = 0; q < KP; ++q) {
        const int e = tid + q * NT;
        unsigned off = 0;
        if (e < PN) {
            const int lx = e % PX, ly = e / PX;
            off = (unsigned)(wrapmod(y0 - 2 + ly, G.ny) * G.nx + wrapmod(x0 - 2 + lx, G.nx));
        }
        poff[q] = off;
    }

    auto load_phi = [&](int z) {
        const float* __restrict__ src = a.phi + (size_t)wrapmod(z, G.nz) * G.nxy;
        float* dst = s_phi[z & 3];
#pragma unroll
        for (int q = 0; q < KP; ++q) {
            const int e = tid + q * NT;
            if (q < KP - 1 || e < PN) dst[e] = ld32(src, poff[q]);
        }
    };
    auto nslot = [](int z) { return (z + 3) % 3; };  // z >= -3
    auto comp_n = [&](int z) {
        const float* pm = s_phi[(z - 1) & 3];
        const float* p0 = s_phi[z & 3];
        const float* pp = s_phi[(z + 1) & 3];
        const int sl = nslot(z);
        float* ind_dst = s_ind[z & 1];
#pragma unroll
        for (int q = 0; q < KH; ++q) {
            const int e = tid + q * NT;
            if (q < KH - 1 || e < HN) {
                const int hx = e % HX, hy = e / HX;
                const int c = (hy + 1) * PX + (hx + 1);
                float nx, ny, nz, ind;
                grad_normal(
                    [&](auto lc) {
                        constexpr int l = decltype(lc)::value;
                        return pick<CZ[l]>(pm, p0, pp)[c + CY[l] * PX + CX[l]];
                    },
                    d, nx, ny, nz, ind);
                s_n[sl][0][e] = nx;
                s_n[sl][1][e] = ny;
                s_n[sl][2][e] = nz;
                if (hx >= 1 && hx <= TX && hy >= 1 && hy <= TY) ind_dst[(hy - 1) * TX + (hx - 1)] = ind;
            }
        }
    };

    const int x = x0 + tx, y = y0 + ty;
    const bool valid = x < G.nx && y < G.ny;
    // Unsigned 32-bit offsets: the gathers then use a scalar base + 32-bit
    // vector offset instead of 64-bit per-lane addresses.
    const unsigned xs[3] = {valid ? (unsigned)shift<-1>(x, G.nx) : 0u, (unsigned)x,
                            valid ? (unsigned)shift<1>(x, G.nx) : 0u};
    const unsigned ys[3] = {valid ? (unsigned)(shift<-1>(y, G.ny) * G.nx) : 0u, (unsigned)(y * G.nx),
                            valid ? (unsigned)(shift<1>(y, G.ny) * G.nx) : 0u};

    // Prologue: phi planes zb-2..zb+1, normals of planes zb-1 and zb.
    load_phi(zb - 2);
    load_phi(zb - 1);
    load_phi(zb);
    load_phi(zb + 1);
    __syncthreads();
    comp_n(zb - 1);
    comp_n(zb);
    __syncthreads();

    for (int z = zb; z < ze; ++z) {
        // Slot (z+2)&3 held plane z-2, last read by comp_n(z-1) before the
        // previous barrier.
        load_phi(z + 2);
        __syncthreads();
        // n slot of z+1 held plane z-2 (read last in iteration z-1).
        comp_n(z + 1);
        __syncthreads();
        if (!valid) continue;

        const unsigned zs[3] = {(unsigned)(shift<-1>(z, G.nz) * G.nxy), (unsigned)(z * G.nxy),
                                (unsigned)(shift<1>(z, G.nz) * G.nxy)};
        const unsigned me = zs[1] + ys[1] + xs[1];
        const int sm = nslot(z - 1), s0 = nslot(z), sp = nslot(z + 1);
        const int e0 = (ty + 1) * HX + (tx + 1);
        const float ncx = s_n[s0][0][e0], ncy = s_n[s0][1][e0], ncz = s_n[s0][2][e0];
        const float ind = s_ind[z & 1][tid];

        float curv, Fx, Fy, Fz;
        curvature_force(
            [&](auto lc, float& ax, float& ay, float& az) {
                constexpr int l = decltype(lc)::value;
                const int slot = pick<CZ[l]>(sm, s0, sp);
                const int e = e0 + CY[l] * HX + CX[l];
                if constexpr (CX[l] != 0) ax = s_n[slot][0][e];
                if constexpr (CY[l] != 0) ay = s_n[slot][1][e];
                if constexpr (CZ[l] != 0) az = s_n[slot][2][e];
            },
            a.phys, d, ncx, ncy, ncz, ind, curv, Fx, Fy, Fz);

        float fv[Q];
        for_links([&](auto lc) {
            constexpr int l = decltype(lc)::value;
            fv[l] = fsrc.ld<l>(zs[1 - CZ[l]] + ys[1 - CY[l]] + xs[1 - CX[l]]);
        });
        const Moments m = moments(fv, a.phys, Fx, Fy, Fz);

        if constexpr (MODE == COLLIDE_MACROS) {
            st32(a.rho, me, m.rho);
            st32(a.ux, me, m.ux);
            st32(a.uy, me, m.uy);
            st32(a.uz, me, m.uz);
            st32(a.fx, me, Fx);
            st32(a.fy, me, Fy);
            st32(a.fz, me, Fz);
            st32(a.nx, me, ncx);
            st32(a.ny, me, ncy);
            st32(a.nz, me, ncz);
        } else {
            const FluidCell fc = fluid_cell(m, Fx, Fy, Fz);
            for_links([&](auto lc) {
                constexpr int l = decltype(lc)::value;
                fdst.st<l>(me, fluid_post<l>(fc, a.phys, d));
            });
            const float phic = s_phi[z & 3][(ty + 2) * PX + (tx + 2)];
            const PhaseCell pc = phase_cell(a.phys, phic, m.ux, m.uy, m.uz, ncx, ncy, ncz);
            if constexpr (MODE == COLLIDE_STEP) {
                for_links([&](auto lc) {
                    constexpr int l = decltype(lc)::value;
                    gdst.st<l>(me, phase_post<l>(pc, a.phys));
                });
            } else {
                for_links([&](auto lc) {
                    constexpr int l = decltype(lc)::value;
                    const float gpre = gsrc.ld<l>(zs[1 - CZ[l]] + ys[1 - CY[l]] + xs[1 - CX[l]]);
                    gdst.st<l>(me, phase_post_bgk<l>(pc, a.phys, d, gpre));
                });
            }
        }
    }
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
        f[(size_t)l * G.plane + me] = w * 1.0f;
        g[(size_t)l * G.plane + me] = w * phi[zs[1 + CZ[l]] + ys[1 + CY[l]] + xs[1 + CX[l]]];
    });
}

template <int SIGN>
__global__ __launch_bounds__(PBX * PBY) void k_link_shift(const float* __restrict__ src, float* __restrict__ dst,
                                                          Geo G) {
    const int x = blockIdx.x * PBX + threadIdx.x;
    const int y = blockIdx.y * PBY + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= G.nx || y >= G.ny) return;
    const int xs[3] = {shift<-1>(x, G.nx), x, shift<1>(x, G.nx)};
    const int ys[3] = {shift<-1>(y, G.ny) * G.nx, y * G.nx, shift<1>(y, G.ny) * G.nx};
    const int zs[3] = {shift<-1>(z, G.nz) * G.nxy, z * G.nxy, shift<1>(z, G.nz) * G.nxy};
    for_links([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        const size_t pl = (size_t)l * G.plane;
        dst[pl + zs[1] + ys[1] + xs[1]] = src[pl + zs[1 - SIGN * CZ[l]] + ys[1 - SIGN * CY[l]] + xs[1 - SIGN * CX[l]]];
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

dim3 cell_grid(const Geo& G) { return dim3((G.nx + PBX - 1) / PBX, (G.ny + PBY - 1) / PBY, G.nz); }

// Tile shapes the collide pass is instantiated for.
const Tiling kTilings[] = {
    {64, 4, 0}, {64, 8, 0}, {32, 8, 0}, {32, 4, 0}, {16, 16, 0}, {16, 8, 0},
};

template <int TX, int TY>
hipError_t launch_collide_t(const CollideArgs& a, int mode, hipStream_t s) {
    const unsigned tiles = (unsigned)((a.geo.nx + TX - 1) / TX) * (unsigned)((a.geo.ny + TY - 1) / TY) *
                           (unsigned)((a.geo.nz + a.zc - 1) / a.zc);
    dim3 grid(tiles, 1, 1);
    dim3 block(TX, TY, 1);
    switch (mode) {
        case COLLIDE_STEP: hipLaunchKernelGGL((k_collide<TX, TY, COLLIDE_STEP>), grid, block, 0, s, a); break;
        case COLLIDE_STEP_BGK: hipLaunchKernelGGL((k_collide<TX, TY, COLLIDE_STEP_BGK>), grid, block, 0, s, a); break;
        case COLLIDE_MACROS: hipLaunchKernelGGL((k_collide<TX, TY, COLLIDE_MACROS>), grid, block, 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

bool tiling_supported(const Tiling& t) {
    for (const auto& k : kTilings)
        if (k.tx == t.tx && k.ty == t.ty) return t.zc >= 1;
    return false;
}

int tiling_count() { return (int)(sizeof(kTilings) / sizeof(kTilings[0])); }
Tiling tiling_at(int i) { return kTilings[i]; }

hipError_t launch_phase(const float* g, float* phi, const Geo& geo, hipStream_t s) {
    const unsigned gx = (unsigned)(geo.nx + PBX - 1) / PBX, gy = (unsigned)(geo.ny + PBY - 1) / PBY;
    hipLaunchKernelGGL(k_phase, dim3(gx * gy * (unsigned)geo.nz), dim3(PBX, PBY, 1), 0, s, g, phi, geo, gx, gy);
    return hipGetLastError();
}

hipError_t launch_collide(const CollideArgs& a0, int mode, const Tiling& t, hipStream_t s) {
    CollideArgs a = a0;
    a.zc = t.zc;
    if (t.tx == 64 && t.ty == 4) return launch_collide_t<64, 4>(a, mode, s);
    if (t.tx == 64 && t.ty == 8) return launch_collide_t<64, 8>(a, mode, s);
    if (t.tx == 32 && t.ty == 8) return launch_collide_t<32, 8>(a, mode, s);
    if (t.tx == 32 && t.ty == 4) return launch_collide_t<32, 4>(a, mode, s);
    if (t.tx == 16 && t.ty == 16) return launch_collide_t<16, 16>(a, mode, s);
    if (t.tx == 16 && t.ty == 8) return launch_collide_t<16, 8>(a, mode, s);
    return hipErrorInvalidValue;
}

hipError_t launch_init_pull(const float* phi, float* f, float* g, const Geo& geo, const Phys& p, hipStream_t s) {
    hipLaunchKernelGGL(k_init_pull, cell_grid(geo), dim3(PBX, PBY, 1), 0, s, phi, f, g, geo, p);
    return hipGetLastError();
}

hipError_t launch_link_shift(const float* src, float* dst, const Geo& geo, int sign, hipStream_t s) {
    if (sign > 0) hipLaunchKernelGGL(k_link_shift<1>, cell_grid(geo), dim3(PBX, PBY, 1), 0, s, src, dst, geo);
    else hipLaunchKernelGGL(k_link_shift<-1>, cell_grid(geo), dim3(PBX, PBY, 1), 0, s, src, dst, geo);
    return hipGetLastError();
}

hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s) {
    const long long n4 = (long long)(bytes / 16);
    const long long blocks = std::min<long long>((n4 + 255) / 256, 256LL * 32);
    hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, s, static_cast<const float4*>(src),
                       static_cast<float4*>(dst), n4);
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
