// 16-bit population storage (FP16 or BF16 offsets from FP32 anchors, see
// tile.hpp): batched phase and collide kernels (a single case is a batch of
// one), the initial state, and conversions to and from the FP32 reference
// layout. Arithmetic is FP32. This translation unit is compiled with FMA
// contraction allowed: the 16-bit form is compared with an FP64 run, not
// with the FP32 transcription bit for bit.
#include "tile.hpp"

namespace mcb {
namespace {

constexpr int PBX = 64, PBY = 4;

dim3 cell_grid(const Geo& G) { return dim3((G.nx + PBX - 1) / PBX, (G.ny + PBY - 1) / PBY, G.nz); }

struct Offs {
    unsigned xs[3], ys[3], zs[3];
    __device__ Offs(int x, int y, int z, const Geo& G) {
        xs[0] = (unsigned)shift<-1>(x, G.nx); xs[1] = (unsigned)x; xs[2] = (unsigned)shift<1>(x, G.nx);
        ys[0] = (unsigned)(shift<-1>(y, G.ny) * G.nx); ys[1] = (unsigned)(y * G.nx);
        ys[2] = (unsigned)(shift<1>(y, G.ny) * G.nx);
        zs[0] = (unsigned)(shift<-1>(z, G.nz) * G.nxy); zs[1] = (unsigned)(z * G.nxy);
        zs[2] = (unsigned)(shift<1>(z, G.nz) * G.nxy);
    }
    __device__ unsigned me() const { return zs[1] + ys[1] + xs[1]; }
    template <int L, int SIGN> __device__ unsigned nb() const {
        return zs[1 + SIGN * CZ[L]] + ys[1 + SIGN * CY[L]] + xs[1 + SIGN * CX[L]];
    }
};

// Initial pull-form state: f*_l = w_l (offset 0); g*_l(y) = w_l * phi(y + c_l)
// with anchor phi(y) when g is stored as populations (g null: moment form).
template <class S>
__global__ __launch_bounds__(PBX * PBY) void k_init16(const float* __restrict__ phi, void* f, void* g, Geo G,
                                                      Phys p) {
    const int x = blockIdx.x * PBX + threadIdx.x;
    const int y = blockIdx.y * PBY + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= G.nx || y >= G.ny) return;
    const Offs o(x, y, z, G);
    const unsigned me = o.me();
    const PopPlanes<S> fp(f, G), gp(g, G);
    for_links([&](auto lc) { fp.template st<decltype(lc)::value>(me, 0.0f); });
    if (!g) return;
    const float ph = phi[me];
    float rg = 0.0f;
    for_links([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        if constexpr (l != 0) {
            const float s = offset_of_pop<l>(p, ph, p.w[WC[l]] * phi[o.template nb<l, 1>()]);
            rg += (s - gp.template st<l>(me, s)) * off_unscale(WC[l]);
        }
    });
    gp.template st<0>(me, offset_of_pop<0>(p, ph, p.w[0] * ph + rg));
}

// 16-bit pull form -> FP32 reference layout (post-stream populations).
template <class S>
__global__ __launch_bounds__(PBX * PBY) void k_decode16(const void* src, const float* __restrict__ anchor,
                                                        float* __restrict__ dst, Geo G, Phys p) {
    const int x = blockIdx.x * PBX + threadIdx.x;
    const int y = blockIdx.y * PBY + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= G.nx || y >= G.ny) return;
    const Offs o(x, y, z, G);
    const PopPlanes<S> sp(src, G);
    for_links([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        const unsigned s = o.template nb<l, -1>();
        dst[(size_t)l * G.plane + o.me()] = pop_from_offset<l>(p, anchor ? anchor[s] : 1.0f, sp.template ld<l>(s));
    });
}

// FP32 reference layout -> 16-bit pull form against the anchor field.
template <class S>
__global__ __launch_bounds__(PBX * PBY) void k_encode16(const float* __restrict__ src,
                                                        const float* __restrict__ anchor, void* dst, Geo G,
                                                        Phys p) {
    const int x = blockIdx.x * PBX + threadIdx.x;
    const int y = blockIdx.y * PBY + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= G.nx || y >= G.ny) return;
    const Offs o(x, y, z, G);
    const unsigned me = o.me();
    const PopPlanes<S> dp(dst, G);
    const float a = anchor ? anchor[me] : 1.0f;
    float r = 0.0f;
    for_links([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        if constexpr (l != 0) {
            const float s = offset_of_pop<l>(p, a, src[(size_t)l * G.plane + o.template nb<l, 1>()]);
            r += (s - dp.template st<l>(me, s)) * off_unscale(WC[l]);
        }
    });
    dp.template st<0>(me, offset_of_pop<0>(p, a, src[me] + r));
}

// (phi, B) moment form -> reference layout: dst_l(x) = w_l (phi + c_l.B)(x - c_l),
// or w_l phi0(x) before the first step.
__global__ __launch_bounds__(PBX * PBY) void k_ab_to_ref(CaseView c, float* __restrict__ dst) {
    const Geo& G = c.geo;
    const int x = blockIdx.x * PBX + threadIdx.x;
    const int y = blockIdx.y * PBY + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= G.nx || y >= G.ny) return;
    const Offs o(x, y, z, G);
    const unsigned me = o.me();
    for_links([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        float v;
        if (c.gm == GM_INITIAL) {
            v = c.phys.w[WC[l]] * ld32(c.phi_prev, me);
        } else {
            const unsigned s = o.template nb<l, -1>();
            float a = ld32(c.phi_prev, s);
            if constexpr (CX[l] != 0) a = CX[l] > 0 ? a + ld_b(c.gm_u[0], s) : a - ld_b(c.gm_u[0], s);
            if constexpr (CY[l] != 0) a = CY[l] > 0 ? a + ld_b(c.gm_u[1], s) : a - ld_b(c.gm_u[1], s);
            if constexpr (CZ[l] != 0) a = CZ[l] > 0 ? a + ld_b(c.gm_u[2], s) : a - ld_b(c.gm_u[2], s);
            v = c.phys.w[WC[l]] * a;
        }
        dst[(size_t)l * G.plane + me] = v;
    });
}

template <int TX, int TY, class S>
hipError_t batch_t(const CaseView* cases, const WorkItem* work, unsigned n, int mode, bool phase, hipStream_t s) {
    dim3 grid(n, 1, 1), block(TX, TY, 1);
    if (phase) {
        hipLaunchKernelGGL((k_phase_batch<TX, TY, S>), grid, block, 0, s, cases, work);
        return hipGetLastError();
    }
    switch (mode) {
        case COLLIDE_STEP:
            hipLaunchKernelGGL((k_collide_batch<TX, TY, COLLIDE_STEP, S>), grid, block, 0, s, cases, work);
            break;
        case COLLIDE_STEP_BGK:
            hipLaunchKernelGGL((k_collide_batch<TX, TY, COLLIDE_STEP_BGK, S>), grid, block, 0, s, cases, work);
            break;
        case COLLIDE_MACROS:
            hipLaunchKernelGGL((k_collide_batch<TX, TY, COLLIDE_MACROS, S>), grid, block, 0, s, cases, work);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class S>
hipError_t batch_s(const CaseView* cases, const WorkItem* work, unsigned n, const Tiling& t, int mode, bool phase,
                   hipStream_t s) {
    if (n == 0) return hipSuccess;
#define MCB_TILE_CASE(TX, TY) \
    if (t.tx == TX && t.ty == TY) return batch_t<TX, TY, S>(cases, work, n, mode, phase, s);
    MCB_TILES_16(MCB_TILE_CASE)
#undef MCB_TILE_CASE
    return hipErrorInvalidValue;
}

hipError_t batch16(const CaseView* cases, const WorkItem* work, unsigned n, const Tiling& t, int mode, bool phase,
                   int storage, hipStream_t s) {
    if (storage == MCB_STORE_FP16) return batch_s<StoreF16>(cases, work, n, t, mode, phase, s);
    if (storage == MCB_STORE_BF16) return batch_s<StoreBF16>(cases, work, n, t, mode, phase, s);
    return hipErrorInvalidValue;
}

}  // namespace

bool tiling_supported16(const Tiling& t) {
#define MCB_TILE_CASE(TX, TY) \
    if (t.tx == TX && t.ty == TY) return t.zc >= 1;
    MCB_TILES_16(MCB_TILE_CASE)
#undef MCB_TILE_CASE
    return false;
}

namespace {
// launch_mass_finish: one workgroup per case of the view array.
constexpr int kFinishThreads = 256;
__global__ __launch_bounds__(kFinishThreads) void k_mass_finish(const CaseView* __restrict__ cases) {
    const CaseView c = cases[blockIdx.x];
    const Geo& G = c.geo;
    const unsigned n_tiles = (unsigned)((G.nx + c.tx - 1) / c.tx) * (unsigned)((G.ny + c.ty - 1) / c.ty) *
                             (unsigned)((G.nz + c.zc - 1) / c.zc);
    if (c.mass && c.gm == GM_POPULATIONS) {   // stored 16-bit g: FP64 partials, fixed order
        double acc = 0.0;
        for (unsigned i = threadIdx.x; i < n_tiles; i += kFinishThreads) acc += c.partial[i];
        const double total = block_sum<kFinishThreads>(acc);
        if (threadIdx.x == 0) c.mass->delta[c.mslot ^ 1] = (c.mass->target - total) / (double)G.n;
        return;
    }
    if (!c.massq || c.gm != GM_MOMENTS) return;   // uniform per workgroup
    const long long* partial = reinterpret_cast<const long long*>(c.partial);
    long long acc = 0;
    for (unsigned i = threadIdx.x; i < n_tiles; i += kFinishThreads) acc += partial[i];
    const long long total = block_sum_q<kFinishThreads>(acc);
    if (threadIdx.x == 0) {
        MassQ* m = c.massq;
        if (!m->set) {
            m->target = total;
            m->set = 1;
            m->delta[c.mslot ^ 1] = 0.0;
        } else {
            m->delta[c.mslot ^ 1] = (double)(m->target - total) / kMassQScale / (double)G.n;
        }
    }
}
}  // namespace

hipError_t launch_mass_finish(const CaseView* cases, unsigned n_cases, hipStream_t s) {
    if (n_cases == 0) return hipSuccess;
    hipLaunchKernelGGL(k_mass_finish, dim3(n_cases), dim3(kFinishThreads), 0, s, cases);
    return hipGetLastError();
}

hipError_t launch_phase_batch16(const CaseView* cases, const WorkItem* work, unsigned n_work, const Tiling& t,
                                int storage, hipStream_t s) {
    return batch16(cases, work, n_work, t, 0, true, storage, s);
}

hipError_t launch_collide_batch16(const CaseView* cases, const WorkItem* work, unsigned n_work, const Tiling& t,
                                  int mode, int storage, hipStream_t s) {
    return batch16(cases, work, n_work, t, mode, false, storage, s);
}

hipError_t launch_init16(const float* phi, void* f, void* g, const Geo& geo, const Phys& p, int storage,
                         hipStream_t s) {
    if (storage == MCB_STORE_FP16)
        hipLaunchKernelGGL(k_init16<StoreF16>, cell_grid(geo), dim3(PBX, PBY, 1), 0, s, phi, f, g, geo, p);
    else if (storage == MCB_STORE_BF16)
        hipLaunchKernelGGL(k_init16<StoreBF16>, cell_grid(geo), dim3(PBX, PBY, 1), 0, s, phi, f, g, geo, p);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_ab_to_ref(const CaseView& c, float* dst, hipStream_t s) {
    if (c.gm == GM_POPULATIONS) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_ab_to_ref, cell_grid(c.geo), dim3(PBX, PBY, 1), 0, s, c, dst);
    return hipGetLastError();
}

hipError_t launch_decode16(const void* src, const float* anchor, float* dst, const Geo& geo, const Phys& p,
                           int storage, hipStream_t s) {
    if (storage == MCB_STORE_FP16)
        hipLaunchKernelGGL(k_decode16<StoreF16>, cell_grid(geo), dim3(PBX, PBY, 1), 0, s, src, anchor, dst, geo, p);
    else if (storage == MCB_STORE_BF16)
        hipLaunchKernelGGL(k_decode16<StoreBF16>, cell_grid(geo), dim3(PBX, PBY, 1), 0, s, src, anchor, dst, geo, p);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_encode16(const float* src, const float* anchor, void* dst, const Geo& geo, const Phys& p,
                           int storage, hipStream_t s) {
    if (storage == MCB_STORE_FP16)
        hipLaunchKernelGGL(k_encode16<StoreF16>, cell_grid(geo), dim3(PBX, PBY, 1), 0, s, src, anchor, dst, geo, p);
    else if (storage == MCB_STORE_BF16)
        hipLaunchKernelGGL(k_encode16<StoreBF16>, cell_grid(geo), dim3(PBX, PBY, 1), 0, s, src, anchor, dst, geo, p);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace mcb
