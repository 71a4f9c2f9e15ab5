// Per-cell arithmetic of the 16-bit population forms: the same discrete
// model as lattice.hpp (gradCalc, curvatureCalc, momentiCalc,
// collisionFluid, collisionPhase of src/lbm.cu), evaluated with fewer
// operations. Results agree with lattice.hpp up to FP32 rounding, not bit
// for bit, so this is only used where populations are stored in 16 bits.
//
// D3Q19 links come in opposite pairs (2k-1, 2k), k = 1..9, with c_{2k} =
// -c_{2k-1}. With a = c.u / cs^2, b = 1.5 (c.F) / rho, beta = 1.5 (u.F) / rho
// and o = 1 + a^2/2 - 1.5 |u|^2 of a link, the reference's quantities are
//   feq - HeF = w rho (o + a)(1 + beta - b)      (momentiCalc, lbm.cu:200-212)
//   feq + HeF = w rho (o + a)(1 - beta + b)      (collisionFluid, lbm.cu:273-286)
// and a link's opposite flips the signs of a and b. The fluid collision is
// evaluated as a polynomial in c_l (fluid_poly); the phase collision
// (lbm.cu:319-326) splits into an even part w phi and an odd part:
// g*_+- = w phi +- w (phi a + phi_norm c.n).
//
// Post-collision values are returned as offsets (f*_l / w_l - 1 for f, the
// odd part for g): the small part is formed directly, without the
// cancellation of subtracting w_l * anchor from a full population.
#pragma once
#include "lattice.hpp"

namespace mcb {
namespace fast {

struct Consts {
    float w0, w1, w2;
    float coef1, coef2;            // 3 w1, 3 w2 (gradient / curvature weights)
    float sharp_c, sigma, omega_g, one_m_omega_g;
    // Coefficients of the polynomial form (fluid_poly).
    float m_rho, m_uf, m_x1, m_t, m_t2, tr_c;
    float wk1, wk2;                // w_c * K_c: 16-bit offset scale of f*_l / w_l - 1
    __device__ __forceinline__ explicit Consts(const Phys& p, float k1 = 1.0f, float k2 = 1.0f) {
        wk1 = p.w[1] * k1;
        wk2 = p.w[2] * k2;
        const float w = p.omega;
        m_rho = 4.5f * w;                 // 9/2 omega
        m_uf = 6.75f * (2.0f - w);        // 27/4 (2 - omega)
        m_x1 = 4.5f * (2.0f - w);
        m_t = 4.5f * (1.0f - w);
        m_t2 = 9.0f * (1.0f - w);
        tr_c = 1.5f * (1.0f - w);
        w0 = p.w[0];
        w1 = p.w[1];
        w2 = p.w[2];
        coef1 = 3.0f * w1;
        coef2 = 3.0f * w2;
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

// momentiCalc + collisionFluid (lbm.cu:133-288) for one cell as a polynomial
// in c_l. Input: the offsets t_l = f_l - w_l and the force F. With
// J = sum_l c_l f_l + F/2 = rho u, uF = u.F, rho_m = rho - 1.5 uF and
// rho_p = rho + 1.5 uF, the reference's post-collision population is
//   f*_l = w_l (rho (o + a)(1 - beta + b) + (9/2)(1 - omega)(c c : Pi - tr Pi / 3))
// (a, b, o, beta as above), which expands to
//   f*_l / w_l = P0 + c c : M + c.V + (27/4) (c.u)^2 (c.F)
// with
//   P0   = rho_m (1 - 1.5|u|^2) - 1.5 (1 - omega) tr Pi
//   M_ab = (9/2) [rho_m u_a u_b + X_ab + (1 - omega) Pi_ab],  X_ab = (u_a F_b + u_b F_a)/2
//   V    = 3 rho_m u + 1.5 (1 - 1.5|u|^2) F.
// Pi (momentiCalc's sum of c c (f - feq + HeF)) has a closed form through
// D3Q19's moment identities (sum w c_a c_b = delta/3, fourth moment
// (delta delta + delta delta + delta delta)/9): with T_ab = sum_l c_a c_b t_l,
//   Pi_ab = T_ab - rho_p u_a u_b + X_ab - delta_ab (rho - 1)/3,
// so M_ab = (9/2)[(omega rho - 1.5 (2 - omega) uF) u_a u_b + (2 - omega) X_ab
// + (1 - omega) T_ab] - delta_ab 1.5 (1 - omega)(rho - 1). On D3Q19 links
// (components 0, +-1) the cubic term is linear in the signs: a face link
// +-e_p gives +-u_p^2 F_p, an edge link s_p e_p + s_q e_q gives
// s_p (F_p (u_p^2 + u_q^2) + 2 u_p u_q F_q) + s_q (F_q (u_p^2 + u_q^2) + 2 u_p u_q F_p).
// So every link is one even part (shared by the pair) plus or minus one odd
// part, both sums of per-cell coefficients: a few adds per link instead of
// the per-link products of the reference's form. Same model, other rounding.
//
// emit(LinkC<l>, x) receives f*_l / w_l - 1 for l = 1..18; the function
// returns it for l = 0 and sets rho and u.
template <class Emit>
__device__ __forceinline__ float fluid_poly(const float (&t)[Q], const Consts& k, float fx, float fy, float fz,
                                           float& rho, float& ux, float& uy, float& uz, Emit&& emit) {
    // Pairs k = 1..9 (links 2k-1, 2k): x, y, z, (1,1,0), (1,0,1), (0,1,1),
    // (1,-1,0), (1,0,-1), (0,1,-1).
    float s[10], d[10];
    for_links([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        if constexpr (l % 2 == 1) {
            static_assert(CX[l] == -CX[l + 1] && CY[l] == -CY[l + 1] && CZ[l] == -CZ[l + 1], "opposite pairs");
            s[(l + 1) / 2] = t[l] + t[l + 1];
            d[(l + 1) / 2] = t[l] - t[l + 1];
        }
    });
    static_assert(CX[7] == 1 && CY[7] == 1 && CX[9] == 1 && CZ[9] == 1 && CY[11] == 1 && CZ[11] == 1 &&
                      CX[13] == 1 && CY[13] == -1 && CX[15] == 1 && CZ[15] == -1 && CY[17] == 1 && CZ[17] == -1,
                  "pair table");
    const float drho = (t[0] + (s[1] + s[2] + s[3])) + ((s[4] + s[5] + s[6]) + (s[7] + s[8] + s[9]));
    rho = 1.0f + drho;
    const float jx = (d[1] + (d[4] + d[5])) + (d[7] + d[8]);
    const float jy = (d[2] + (d[4] + d[6])) + (d[9] - d[7]);
    const float jz = (d[3] + (d[5] + d[6])) - (d[8] + d[9]);
    const float txx = (s[1] + (s[4] + s[5])) + (s[7] + s[8]);
    const float tyy = (s[2] + (s[4] + s[6])) + (s[7] + s[9]);
    const float tzz = (s[3] + (s[5] + s[6])) + (s[8] + s[9]);
    const float txy = s[4] - s[7], txz = s[5] - s[8], tyz = s[6] - s[9];

    const float ir = __builtin_amdgcn_rcpf(rho);
    ux = (jx + 0.5f * fx) * ir;
    uy = (jy + 0.5f * fy) * ir;
    uz = (jz + 0.5f * fz) * ir;
    const float uf = ux * fx + uy * fy + uz * fz;
    const float ux2 = ux * ux, uy2 = uy * uy, uz2 = uz * uz;
    const float u2 = ux2 + uy2 + uz2;
    const float pxy = ux * uy, pxz = ux * uz, pyz = uy * uz;
    const float rho_m = rho - 1.5f * uf, rho_p = rho + 1.5f * uf;

    // M (diagonal) and 2 M (off-diagonal).
    const float c1 = k.m_rho * rho - k.m_uf * uf;
    const float c2 = c1 + c1;
    const float dd = -k.tr_c * drho;
    const float mxx = (c1 * ux2 + (k.m_x1 * (ux * fx) + k.m_t * txx)) + dd;
    const float myy = (c1 * uy2 + (k.m_x1 * (uy * fy) + k.m_t * tyy)) + dd;
    const float mzz = (c1 * uz2 + (k.m_x1 * (uz * fz) + k.m_t * tzz)) + dd;
    const float m2xy = c2 * pxy + (k.m_x1 * (ux * fy + uy * fx) + k.m_t2 * txy);
    const float m2xz = c2 * pxz + (k.m_x1 * (ux * fz + uz * fx) + k.m_t2 * txz);
    const float m2yz = c2 * pyz + (k.m_x1 * (uy * fz + uz * fy) + k.m_t2 * tyz);
    const float tr_pi = ((txx + tyy + tzz) - rho_p * u2) + (uf - drho);
    const float p0m1 = ((drho - 1.5f * uf) - rho_m * (1.5f * u2)) - k.tr_c * tr_pi;

    // Odd coefficients.
    const float r3 = 3.0f * rho_m, hv = 1.5f - 2.25f * u2;
    const float vx = r3 * ux + hv * fx, vy = r3 * uy + hv * fy, vz = r3 * uz + hv * fz;
    const float afx = vx + 6.75f * (ux2 * fx), afy = vy + 6.75f * (uy2 * fy), afz = vz + 6.75f * (uz2 * fz);
    auto plane = [&](float vp, float vq, float up2, float uq2, float p2, float fp, float fq, float& ap, float& aq) {
        const float sq = up2 + uq2;
        ap = vp + 6.75f * (fp * sq + p2 * fq);
        aq = vq + 6.75f * (fq * sq + p2 * fp);
    };
    float axy_x, axy_y, axz_x, axz_z, ayz_y, ayz_z;
    plane(vx, vy, ux2, uy2, pxy + pxy, fx, fy, axy_x, axy_y);
    plane(vx, vz, ux2, uz2, pxz + pxz, fx, fz, axz_x, axz_z);
    plane(vy, vz, uy2, uz2, pyz + pyz, fy, fz, ayz_y, ayz_z);

    // Even parts.
    const float efx = p0m1 + mxx, efy = p0m1 + myy, efz = p0m1 + mzz;
    const float exy = efx + myy, exz = efx + mzz, eyz = efy + mzz;
    auto pair = [&](auto lc, float even, float odd) {
        constexpr int l = decltype(lc)::value;
        emit(LinkC<l>{}, even + odd);
        emit(LinkC<l + 1>{}, even - odd);
    };
    pair(LinkC<1>{}, efx, afx);
    pair(LinkC<3>{}, efy, afy);
    pair(LinkC<5>{}, efz, afz);
    pair(LinkC<7>{}, exy + m2xy, axy_x + axy_y);
    pair(LinkC<9>{}, exz + m2xz, axz_x + axz_z);
    pair(LinkC<11>{}, eyz + m2yz, ayz_y + ayz_z);
    pair(LinkC<13>{}, exy - m2xy, axy_x - axy_y);
    pair(LinkC<15>{}, exz - m2xz, axz_x - axz_z);
    pair(LinkC<17>{}, eyz - m2yz, ayz_y - ayz_z);
    return p0m1;
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
