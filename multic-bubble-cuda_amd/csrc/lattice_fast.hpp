// Per-cell arithmetic of the 16-bit population forms: the same discrete
// model as lattice.hpp (gradCalc, curvatureCalc, momentiCalc,
// collisionFluid, collisionPhase of src/lbm.cu), evaluated with fewer
// operations. Results agree with lattice.hpp up to FP32 rounding, not bit
// for bit, so this is only used where populations are stored in 16 bits.
//
// D3Q19 links come in opposite pairs (2k-1, 2k), k = 1..9, with c_{2k} =
// -c_{2k-1}. With a = c.u / cs^2, b = 1.5 (c.F) / rho, beta = 1.5 (u.F) / rho,
// e = a^2 / 2 - 1.5 |u|^2 and o = 1 + e, the reference's quantities of the
// pair are
//   feq_+- - HeF_+- = w rho (o +- a)(1 + beta -+ b)      (momentiCalc, lbm.cu:200-212)
//   feq_+- + HeF_+- = w rho (o +- a)(1 - beta +- b)      (collisionFluid, lbm.cu:273-286)
// so each pair needs one even part E = w rho (o (1 -+ beta) +- a b) and one
// odd part O, and the regularised non-equilibrium term (Q_l : Pi, even in l)
// once. Pi only needs the even parts: fneq_+ + fneq_- = f_+ + f_- - 2 E.
// The phase collision (lbm.cu:319-326) splits the same way:
// g*_+- = w phi +- w (phi a + phi_norm c.n).
//
// Post-collision values are returned as offsets from w_l * anchor (rho for f,
// phi for g): the small part is formed directly, without the cancellation of
// subtracting w_l * anchor from a full population.
#pragma once
#include "lattice.hpp"

namespace mcb {
namespace fast {

// Pair k = 1..9: its "+" link is 2k-1.
template <int K> constexpr int PL = 2 * K - 1;

struct Consts {
    float w0, w1, w2;
    float coef1, coef2;            // 3 w1, 3 w2 (gradient / curvature weights)
    float r0, r1, r2;              // (1 - omega) w_c / (2 cs^4)
    float sharp_c, sigma, omega_g, one_m_omega_g;
    __device__ __forceinline__ explicit Consts(const Phys& p) {
        w0 = p.w[0];
        w1 = p.w[1];
        w2 = p.w[2];
        coef1 = 3.0f * w1;
        coef2 = 3.0f * w2;
        const float om = 1.0f - p.omega;
        r0 = om * (w0 / TWO_CS4);
        r1 = om * (w1 / TWO_CS4);
        r2 = om * (w2 / TWO_CS4);
        sharp_c = p.sharp_c;
        sigma = p.sigma;
        omega_g = p.omega_g;
        one_m_omega_g = 1.0f - p.omega_g;
    }
    template <int L> __device__ __forceinline__ float w() const {
        return WC[L] == 0 ? w0 : (WC[L] == 1 ? w1 : w2);
    }
};

// gradCalc: nb(l) returns phi(x + c_l).
template <class NB>
__device__ __forceinline__ void grad_normal(NB&& nb, const Consts& k, float& nx, float& ny, float& nz, float& ind) {
    auto d = [&](auto a, auto b) { return nb(a) - nb(b); };
    const float d12 = d(LinkC<1>{}, LinkC<2>{}), d34 = d(LinkC<3>{}, LinkC<4>{}), d56 = d(LinkC<5>{}, LinkC<6>{});
    const float d78 = d(LinkC<7>{}, LinkC<8>{}), d910 = d(LinkC<9>{}, LinkC<10>{});
    const float d1112 = d(LinkC<11>{}, LinkC<12>{}), d1314 = d(LinkC<13>{}, LinkC<14>{});
    const float d1516 = d(LinkC<15>{}, LinkC<16>{}), d1718 = d(LinkC<17>{}, LinkC<18>{});
    const float gx = k.coef1 * d12 + k.coef2 * ((d78 + d910) + (d1314 + d1516));
    const float gy = k.coef1 * d34 + k.coef2 * ((d78 + d1112) + (d1718 - d1314));
    const float gz = k.coef1 * d56 + k.coef2 * ((d910 + d1112) - (d1516 + d1718));
    const float g2 = gx * gx + gy * gy + gz * gz;
    const float r = __builtin_amdgcn_rsqf(g2 > 1e-9f ? g2 : 1e-9f);
    nx = gx * r;
    ny = gy * r;
    nz = gz * r;
    ind = g2 * r;
}

// curvatureCalc + the surface-tension force. nbn(l, ax, ay, az) loads the
// components of n(x + c_l) that link l needs.
template <class NBN>
__device__ __forceinline__ void curvature_force(NBN&& nbn, const Consts& k, float nx, float ny, float nz, float ind,
                                                float& fx, float& fy, float& fz) {
    float face = 0.0f, edge = 0.0f;
    for_links([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        if constexpr (l != 0) {
            float ax = 0.0f, ay = 0.0f, az = 0.0f;
            nbn(lc, ax, ay, az);
            if constexpr (WC[l] == 1) face = face + cdot<l>(ax, ay, az);
            else edge = edge + cdot<l>(ax, ay, az);
        }
    });
    const float curv = -(k.coef1 * face + k.coef2 * edge);
    const float m = k.sigma * curv * ind;
    fx = m * nx;
    fy = m * ny;
    fz = m * nz;
}

// Moments of the fluid populations and the per-cell terms of the collision.
struct Fluid {
    float rho, ux, uy, uz;
    float uu, beta, hb, wr0, wr1, wr2;  // 1.5|u|^2, beta, 1.5/rho, w_c rho
    float pxx, pyy, pzz, pxy, pxz, pyz;
    float trc;                           // cs^2 (pxx + pyy + pzz)
};

template <int L> __device__ __forceinline__ float wr_of(const Fluid& f) {
    return WC[L] == 0 ? f.wr0 : (WC[L] == 1 ? f.wr1 : f.wr2);
}

// a = c.u / cs^2, b = 1.5 c.F / rho and o = 1 + a^2/2 - 1.5|u|^2 of link L.
template <int L>
__device__ __forceinline__ void pair_terms(const Fluid& f, float fx, float fy, float fz, float& a, float& b,
                                           float& o) {
    a = 3.0f * cdot<L>(f.ux, f.uy, f.uz);
    b = f.hb * cdot<L>(fx, fy, fz);
    o = 1.0f + (0.5f * a * a - f.uu);
}

__device__ __forceinline__ Fluid fluid(const float (&fv)[Q], const Consts& k, float fx, float fy, float fz) {
    Fluid f;
    float s[10], dd[10];
    s[0] = fv[0];
    dd[0] = 0.0f;
    float rho = fv[0];
    for_links([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        if constexpr (l % 2 == 1) {
            s[(l + 1) / 2] = fv[l] + fv[l + 1];
            dd[(l + 1) / 2] = fv[l] - fv[l + 1];
        }
    });
    rho = (rho + (s[1] + s[2] + s[3])) + ((s[4] + s[5] + s[6]) + (s[7] + s[8] + s[9]));
    // pairs: 1 x, 2 y, 3 z, 4 (1,1,0), 5 (1,0,1), 6 (0,1,1), 7 (1,-1,0), 8 (1,0,-1), 9 (0,1,-1)
    const float jx = dd[1] + (dd[4] + dd[5]) + (dd[7] + dd[8]);
    const float jy = dd[2] + (dd[4] + dd[6]) + (dd[9] - dd[7]);
    const float jz = dd[3] + (dd[5] + dd[6]) - (dd[8] + dd[9]);
    const float ir = __builtin_amdgcn_rcpf(rho);
    f.rho = rho;
    f.ux = (jx + 0.5f * fx) * ir;
    f.uy = (jy + 0.5f * fy) * ir;
    f.uz = (jz + 0.5f * fz) * ir;
    f.uu = 1.5f * (f.ux * f.ux + f.uy * f.uy + f.uz * f.uz);
    f.hb = 1.5f * ir;
    f.beta = f.hb * (f.ux * fx + f.uy * fy + f.uz * fz);
    f.wr0 = k.w0 * rho;
    f.wr1 = k.w1 * rho;
    f.wr2 = k.w2 * rho;
    // fneq_+ + fneq_- of each pair: s - 2 w rho (o (1 + beta) - a b).
    float S[10];
    const float opb = 1.0f + f.beta;
    for_links([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        if constexpr (l % 2 == 1) {
            float a, b, o;
            pair_terms<l>(f, fx, fy, fz, a, b, o);
            const float E = wr_of<l>(f) * (o * opb - a * b);
            S[(l + 1) / 2] = s[(l + 1) / 2] - 2.0f * E;
        }
    });
    f.pxx = S[1] + (S[4] + S[5]) + (S[7] + S[8]);
    f.pyy = S[2] + (S[4] + S[6]) + (S[7] + S[9]);
    f.pzz = S[3] + (S[5] + S[6]) + (S[8] + S[9]);
    f.pxy = S[4] - S[7];
    f.pxz = S[5] - S[8];
    f.pyz = S[6] - S[9];
    f.trc = CSSQ * (f.pxx + f.pyy + f.pzz);
    return f;
}

// Q_l : Pi for the "+" link L of a pair (the same for its opposite link).
template <int L> __device__ __forceinline__ float q_pi(const Fluid& f) {
    constexpr int cx = CX[L], cy = CY[L], cz = CZ[L];
    float s = -f.trc;
    if constexpr (cx != 0) s = s + f.pxx;
    if constexpr (cy != 0) s = s + f.pyy;
    if constexpr (cz != 0) s = s + f.pzz;
    if constexpr (cx * cy != 0) s = cx * cy > 0 ? s + 2.0f * f.pxy : s - 2.0f * f.pxy;
    if constexpr (cx * cz != 0) s = cx * cz > 0 ? s + 2.0f * f.pxz : s - 2.0f * f.pxz;
    if constexpr (cy * cz != 0) s = cy * cz > 0 ? s + 2.0f * f.pyz : s - 2.0f * f.pyz;
    return s;
}

// Post-collision offsets f*_l - w_l rho of the pair whose "+" link is L.
template <int L>
__device__ __forceinline__ void fluid_pair(const Fluid& f, const Consts& k, float fx, float fy, float fz, float& op,
                                           float& om) {
    float a, b, o;
    pair_terms<L>(f, fx, fy, fz, a, b, o);
    const float wr = wr_of<L>(f);
    const float omb = 1.0f - f.beta;
    // w rho (o (1 - beta) + a b - 1) = w rho ((o - 1) - beta o + a b)
    const float even = wr * (((o - 1.0f) - f.beta * o) + a * b) + (WC[L] == 1 ? k.r1 : k.r2) * q_pi<L>(f);
    const float odd = wr * (a * omb + b * o);
    op = even + odd;
    om = even - odd;
}

// Rest link: f*_0 - w0 rho.
__device__ __forceinline__ float fluid_rest(const Fluid& f, const Consts& k) {
    // w0 rho ((1 - uu)(1 - beta) - 1) + r0 (-trc)
    return f.wr0 * (-f.uu - f.beta * (1.0f - f.uu)) - k.r0 * f.trc;
}

// Phase collision, odd part w (phi a + phi_norm c.n) of the pair whose "+"
// link is L (the even part is w phi, the anchor itself).
template <int L>
__device__ __forceinline__ float phase_odd(const Consts& k, float phi, float phi_norm, float ux, float uy, float uz,
                                           float nx, float ny, float nz) {
    const float a = 3.0f * cdot<L>(ux, uy, uz);
    return k.w<L>() * (phi * a + phi_norm * cdot<L>(nx, ny, nz));
}

}  // namespace fast
}  // namespace mcb
