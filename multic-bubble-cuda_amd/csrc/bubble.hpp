// Initial bubble phase field of initPhase (src/init.cu:27-49), one definition
// for the host initialisation (host.cpp) and the device kernels (k_bubble_phi,
// k_initPhase), so that both produce the same bits.
//
// tanh: the reference calls CUDA's tanhf (src/init.cu:45), whose last bit is
// the toolkit's own choice. Here tanh of a float argument is evaluated in
// double with IEEE add, subtract, multiply and divide and exact power-of-two
// scaling only. Each of those is correctly rounded on x86-64 (SSE2) and on
// gfx950, and with contraction off no fused multiply-add changes a rounding,
// so the host and the GPU compute identical bits. The double result has a
// relative error below 1e-15 and is rounded once to float: this is the
// correctly rounded tanh except where the exact value lies within that
// distance of a float rounding boundary.
//
// Method (textbook argument reduction, e.g. J.-M. Muller, "Elementary
// Functions", ch. 11): tanh|x| = m / (m + 2) with m = expm1(2|x|);
// 2|x| = k ln2 + r with |r| <= ln2/2 (Cody-Waite: ln2 split in a high part
// with 21 trailing zero bits, so k * ln2_hi is exact, and a low part);
// expm1(r) by its Taylor series to r^13/13! (truncation < 2e-17 relative);
// expm1(2|x|) = 2^k expm1(r) + (2^k - 1), both terms exact but the sum.
#pragma once
#include <cmath>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MCB_BUBBLE_HD __host__ __device__ __forceinline__
#else
#define MCB_BUBBLE_HD inline
#endif

namespace mcb {

MCB_BUBBLE_HD float tanh_f32(float xf) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    const double x = (double)xf;
    const double a = x < 0.0 ? -x : x;
    // |x| >= 20: tanh rounds to +-1 (1 - tanh(9.1) < 2^-25 already). NaN
    // fails the comparison and is returned unchanged; so is +-0.
    if (!(a < 20.0)) return x < 0.0 ? -1.0f : (x > 0.0 ? 1.0f : xf);
    if (a == 0.0) return xf;
    const double y = a + a;
    const int k = (int)(y * 1.4426950408889634 + 0.5);   // round(y / ln2), y >= 0
    const double kd = (double)k;
    const double ln2_hi = 6.93147180369123816490e-01;    // 0x3FE62E42FEE00000
    const double ln2_lo = 1.90821492927058770002e-10;    // ln2 - ln2_hi
    const double r = (y - kd * ln2_hi) - kd * ln2_lo;
    // expm1(r) = r + r^2 (1/2! + r (1/3! + ... + r / 13!)).
    double s = 1.6059043836821613e-10;                   // 1/13!
    s = s * r + 2.08767569878681e-09;                    // 1/12!
    s = s * r + 2.505210838544172e-08;                   // 1/11!
    s = s * r + 2.755731922398589e-07;                   // 1/10!
    s = s * r + 2.7557319223985893e-06;                  // 1/9!
    s = s * r + 2.48015873015873e-05;                    // 1/8!
    s = s * r + 0.0001984126984126984;                   // 1/7!
    s = s * r + 0.001388888888888889;                    // 1/6!
    s = s * r + 0.008333333333333333;                    // 1/5!
    s = s * r + 0.041666666666666664;                    // 1/4!
    s = s * r + 0.16666666666666666;                     // 1/3!
    s = s * r + 0.5;                                     // 1/2!
    const double em1r = r + (r * r) * s;
    const double p2k = ldexp(1.0, k);
    const double m = p2k * em1r + (p2k - 1.0);
    const double t = m / (m + 2.0);
    return (float)(x < 0.0 ? -t : t);
}

// phi of one cell for a set of bubbles (centres, radii in cell units):
// 0.5 + 0.5 tanh(2 (R - r) / width) (src/init.cu:39-46), combined by maximum
// over the bubbles; periodic: minimum-image distances (no reference
// counterpart). B is any type with float members cx, cy, cz, r.
template <class B>
MCB_BUBBLE_HD float bubble_phi_cell(int i, int j, int k, int nx, int ny, int nz, bool periodic, const B* bubbles,
                                    int n, float width) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    const float hx = (float)nx * 0.5f, hy = (float)ny * 0.5f, hz = (float)nz * 0.5f;
    float v = 0.0f;
    for (int b = 0; b < n; ++b) {
        float dx = (float)i - bubbles[b].cx;
        float dy = (float)j - bubbles[b].cy;
        float dz = (float)k - bubbles[b].cz;
        if (periodic) {
            if (dx > hx) dx -= (float)nx; else if (dx < -hx) dx += (float)nx;
            if (dy > hy) dy -= (float)ny; else if (dy < -hy) dy += (float)ny;
            if (dz > hz) dz -= (float)nz; else if (dz < -hz) dz += (float)nz;
        }
        const float r = sqrtf(dx * dx + dy * dy + dz * dz);
        const float pb = 0.5f + 0.5f * tanh_f32(2.0f * (bubbles[b].r - r) / width);
        v = b == 0 ? pb : (pb > v ? pb : v);
    }
    return v;
}

}  // namespace mcb
