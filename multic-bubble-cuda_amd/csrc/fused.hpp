// Fused two-pass step of libmcb (PERIODIC mode), internal interface.
//
// Storage ("pull" form): after a step, buffer f[cur] holds the
// post-collision populations f*_l(x) at their source cell x, unstreamed, and
// likewise g[cur]. The pre-collision population the reference holds at x
// after streaming is f*_l(x - c_l), so every kernel gathers with a shift and
// every store is aligned. This carries exactly the reference's numbers
// (collisionFluid pushes f*_l(x) to x + c_l, src/lbm.cu:285-286;
// streamingCalc pulls g_l(x - c_l), src/lbm.cu:350-355).
//
// Pass 1, k_phase   : phi(x) = sum_l g*_l(x - c_l)              (phiCalc)
// Pass 2, k_collide : gradCalc + curvatureCalc on LDS-staged phi and n,
//                     momentiCalc on gathered f, collisionFluid and
//                     collisionPhase, aligned stores of f*, g* to the other
//                     buffer.
#pragma once
#include "common.hpp"

namespace mcb {

enum CollideMode { COLLIDE_STEP = 0, COLLIDE_STEP_BGK = 1, COLLIDE_MACROS = 2 };

struct Tiling {
    int tx, ty, zc;   // x/y tile of a workgroup, z cells marched per workgroup
};

struct CollideArgs {
    Geo geo;
    Phys phys;
    const float* phi;
    const float* fsrc;
    const float* gsrc;
    float* fdst;
    float* gdst;
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
    int zc;
};

bool tiling_supported(const Tiling& t);
int tiling_count();
Tiling tiling_at(int i);

hipError_t launch_phase(const float* g, float* phi, const Geo& geo, hipStream_t s);
hipError_t launch_collide(const CollideArgs& a, int mode, const Tiling& t, hipStream_t s);
// f*_l(y) = w_l * 1, g*_l(y) = w_l * phi(y + c_l): the pull-form image of
// initTensor + initDist (src/init.cu:21-24, 69-79).
hipError_t launch_init_pull(const float* phi, float* f, float* g, const Geo& geo, const Phys& p, hipStream_t s);
// dst_l(x) = src_l(x - sign*c_l) for all 19 planes (sign = +1: pull form ->
// reference layout; sign = -1: reference layout -> pull form).
hipError_t launch_link_shift(const float* src, float* dst, const Geo& geo, int sign, hipStream_t s);
// dst = src for bytes (a multiple of 16) with a plain float4 copy kernel.
hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s);
// Deterministic FP64 sums: out[0] = sum phi, out[1] = sum 0.5*rho*|u|^2.
constexpr int kReduceBlocks = 1024;
hipError_t launch_reduce(const float* phi, const float* rho, const float* ux, const float* uy, const float* uz,
                         long long n, double* partial, double* out, hipStream_t s);

}  // namespace mcb
