// D3Q19 lattice and the per-cell arithmetic of the phase-field bubble LBM,
// shared by every HIP kernel of libmcb (staged operators and fused step).
//
// Link order and vectors follow src/var.cu:29-31 of the reference; weights
// are runtime scalars (w0, w1, w2) so both the reference's float literals
// (var.cu:40-45) and the unit-sum table can be used. Links are unrolled at
// compile time: every c_l component is a constant, so terms whose lattice
// coefficient is 0 are dropped and +-1 factors become adds/subtracts. For
// finite inputs this rounds exactly like the reference's generic loops
// (x*1 == x, x*(-1) == -x, and x + (+-0) == x for the non-zero partial sums
// involved; the only possible difference is the sign of an exact zero, which
// no later operation can observe). Every other operation is kept in the
// reference's order, so with -ffp-contract=off the kernels reproduce the
// transcription bit for bit.
#pragma once
#include <cstddef>
#include <cstdint>
#include <utility>

#if defined(__HIPCC__)
#define MCB_HD __host__ __device__ __forceinline__
#else
#define MCB_HD inline
#endif

namespace mcb {

constexpr int Q = 19;
constexpr int CX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0};
constexpr int CY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1};
constexpr int CZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1};
// Weight class of each link: 0 rest, 1 axis, 2 edge.
constexpr int WC[Q] = {0, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2};

// var.cu:22 evaluated in float; folded at compile time with IEEE rounding.
constexpr float CSSQ = 1.0f / 3.0f;
constexpr float INV_CSSQ = 1.0f / CSSQ;                 // momentiCalc/collisionPhase `1.0f/CSSQ`
constexpr float TWO_CS4 = 2.0f * CSSQ * CSSQ;           // lbm.cu:278 `2.0f*CSSQ*CSSQ`
constexpr float DIAG1 = 1.0f - CSSQ;                    // (c_a*c_a - CSSQ) for c_a = +-1
constexpr float DIAG0 = 0.0f - CSSQ;                    // (c_a*c_a - CSSQ) for c_a = 0

// Runtime physical constants, computed on the host as var.cu:21-25 does.
struct Phys {
    float w[3];       // w0, w1, w2
    float omega;      // 1/tau_f
    float sharp_c;    // SHARP_C
    float sigma;      // SIGMA
    float omega_g;    // 1/tau_g (only used when tau_g != 1)
};

// Loop-invariant derived constants, one set per thread.
struct Derived {
    float coef[3];     // 3.0f * W[l]                (lbm.cu:59, 115)
    float fcoef[3];    // W[l] / (2.0f*CSSQ*CSSQ)    (lbm.cu:278)
    float one_m_omega; // 1.0f - OMEGA               (lbm.cu:286)
    float one_m_omega_g;
    MCB_HD explicit Derived(const Phys& p) {
        for (int c = 0; c < 3; ++c) {
            coef[c] = 3.0f * p.w[c];
            fcoef[c] = p.w[c] / TWO_CS4;
        }
        one_m_omega = 1.0f - p.omega;
        one_m_omega_g = 1.0f - p.omega_g;
    }
};

template <int L> using LinkC = std::integral_constant<int, L>;

template <class Fn, int... Ls>
MCB_HD void for_links_impl(Fn&& fn, std::integer_sequence<int, Ls...>) {
    (fn(LinkC<Ls>{}), ...);
}
// Calls fn(LinkC<l>) for l = 0..18 in order.
template <class Fn> MCB_HD void for_links(Fn&& fn) {
    for_links_impl(fn, std::make_integer_sequence<int, Q>{});
}

// a + c*b for a constant lattice coefficient c in {-1, 0, 1}.
template <int C> MCB_HD float acc(float a, float b) {
    if constexpr (C > 0) return a + b;
    else if constexpr (C < 0) return a - b;
    else return a;
}

// c_x*vx + c_y*vy + c_z*vz of one link, in the reference's x,y,z order with
// the zero terms dropped (lbm.cu:116, 203, 273, 323, 325).
template <int L> MCB_HD float cdot(float vx, float vy, float vz) {
    constexpr int cx = CX[L], cy = CY[L], cz = CZ[L];
    if constexpr (cx != 0) {
        float s = cx > 0 ? vx : -vx;
        s = acc<cy>(s, vy);
        s = acc<cz>(s, vz);
        return s;
    } else if constexpr (cy != 0) {
        float s = cy > 0 ? vy : -vy;
        s = acc<cz>(s, vz);
        return s;
    } else if constexpr (cz != 0) {
        return cz > 0 ? vz : -vz;
    } else {
        return 0.0f;
    }
}

// ---------------------------------------------------------------- gradCalc
// lbm.cu:50-71. nb(l) returns phi(x + c_l).
template <class NB>
MCB_HD void grad_normal(NB&& nb, const Derived& d, float& nx, float& ny, float& nz, float& ind) {
    float gx = 0.0f, gy = 0.0f, gz = 0.0f;
    for_links([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        if constexpr (l != 0) {
            const float t = d.coef[WC[l]] * nb(lc);
            gx = acc<CX[l]>(gx, t);
            gy = acc<CY[l]>(gy, t);
            gz = acc<CZ[l]>(gz, t);
        }
    });
    const float gmag_sq = gx * gx + gy * gy + gz * gz;
    const float m = gmag_sq > 1e-9f ? gmag_sq : 1e-9f;   // fmaxf(gmag_sq, 1e-9)
    const float factor = 1.0f / sqrtf(m);                 // rsqrtf -> IEEE 1/sqrt
    nx = gx * factor;
    ny = gy * factor;
    nz = gz * factor;
    ind = gmag_sq * factor;
}

// ----------------------------------------------------------- curvatureCalc
// lbm.cu:103-124. nbn(l, nx, ny, nz) loads n(x + c_l).
template <class NBN>
MCB_HD void curvature_force(NBN&& nbn, const Phys& p, const Derived& d, float nx, float ny, float nz,
                            float ind, float& curv_out, float& fx, float& fy, float& fz) {
    float curv = 0.0f;
    for_links([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        if constexpr (l != 0) {
            float ax = 0.0f, ay = 0.0f, az = 0.0f;
            nbn(lc, ax, ay, az);
            curv = curv - d.coef[WC[l]] * cdot<l>(ax, ay, az);
        }
    });
    const float mult = p.sigma * curv;
    curv_out = curv;
    fx = mult * nx * ind;
    fy = mult * ny * ind;
    fz = mult * nz * ind;
}

// ------------------------------------------------------------- momentiCalc
struct Moments {
    float rho, ux, uy, uz;
    float pxx, pyy, pzz, pxy, pxz, pyz;
};

// lbm.cu:161-224. fv[l] are the 19 pre-collision populations of the cell.
MCB_HD Moments moments(const float (&fv)[Q], const Phys& p, float fx, float fy, float fz) {
    Moments m;
    float rho = 0.0f;
    for_links([&](auto lc) { rho += fv[decltype(lc)::value]; });
    const float invRho = 1.0f / rho;
    const float sumUx = invRho * (fv[1] - fv[2] + fv[7] - fv[8] + fv[9] - fv[10] + fv[13] - fv[14] + fv[15] - fv[16]);
    const float sumUy = invRho * (fv[3] - fv[4] + fv[7] - fv[8] + fv[11] - fv[12] + fv[14] - fv[13] + fv[17] - fv[18]);
    const float sumUz = invRho * (fv[5] - fv[6] + fv[9] - fv[10] + fv[11] - fv[12] + fv[16] - fv[15] + fv[18] - fv[17]);
    const float ux = sumUx + fx * 0.5f * invRho;
    const float uy = sumUy + fy * 0.5f * invRho;
    const float uz = sumUz + fz * 0.5f * invRho;
    const float uu = 0.5f * (ux * ux + uy * uy + uz * uz) * INV_CSSQ;
    const float invRhoCssq = 1.0f / (rho * CSSQ);
    float fneq[Q];
    for_links([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        const float udotc = cdot<l>(ux, uy, uz) * INV_CSSQ;
        const float eqBase = rho * (udotc + 0.5f * udotc * udotc - uu);
        const float common = p.w[WC[l]] * (rho + eqBase);
        const float HeF = 0.5f * common *
                          (((float)CX[l] - ux) * fx + ((float)CY[l] - uy) * fy + ((float)CZ[l] - uz) * fz) *
                          invRhoCssq;
        const float feq = common - HeF;
        fneq[l] = fv[l] - feq;
    });
    m.pxx = fneq[1] + fneq[2] + fneq[7] + fneq[8] + fneq[9] + fneq[10] + fneq[13] + fneq[14] + fneq[15] + fneq[16];
    m.pyy = fneq[3] + fneq[4] + fneq[7] + fneq[8] + fneq[11] + fneq[12] + fneq[13] + fneq[14] + fneq[17] + fneq[18];
    m.pzz = fneq[5] + fneq[6] + fneq[9] + fneq[10] + fneq[11] + fneq[12] + fneq[15] + fneq[16] + fneq[17] + fneq[18];
    m.pxy = fneq[7] - fneq[13] + fneq[8] - fneq[14];
    m.pxz = fneq[9] - fneq[15] + fneq[10] - fneq[16];
    m.pyz = fneq[11] - fneq[17] + fneq[12] - fneq[18];
    m.rho = rho;
    m.ux = ux;
    m.uy = uy;
    m.uz = uz;
    return m;
}

// ---------------------------------------------------------- collisionFluid
// Per-cell invariants of lbm.cu:263-265.
struct FluidCell {
    float rho, ux, uy, uz, fx, fy, fz;
    float pxx, pyy, pzz, pxy, pxz, pyz;
    float uu, invRhoCssq;
};

MCB_HD FluidCell fluid_cell(const Moments& m, float fx, float fy, float fz) {
    FluidCell c;
    c.rho = m.rho; c.ux = m.ux; c.uy = m.uy; c.uz = m.uz;
    c.fx = fx; c.fy = fy; c.fz = fz;
    c.pxx = m.pxx; c.pyy = m.pyy; c.pzz = m.pzz;
    c.pxy = m.pxy; c.pxz = m.pxz; c.pyz = m.pyz;
    c.uu = 0.5f * (m.ux * m.ux + m.uy * m.uy + m.uz * m.uz) / CSSQ;   // division, as lbm.cu:263
    c.invRhoCssq = 1.0f / (m.rho * CSSQ);
    return c;
}

// Post-collision population of link L (lbm.cu:273-286).
template <int L>
MCB_HD float fluid_post(const FluidCell& c, const Phys& p, const Derived& d) {
    constexpr int cx = CX[L], cy = CY[L], cz = CZ[L];
    const float w = p.w[WC[L]];
    const float udotc = cdot<L>(c.ux, c.uy, c.uz) * INV_CSSQ;
    const float feq = w * (c.rho + c.rho * (udotc + 0.5f * udotc * udotc - c.uu));
    const float HeF = 0.5f * feq *
                      (((float)cx - c.ux) * c.fx + ((float)cy - c.uy) * c.fy + ((float)cz - c.uz) * c.fz) *
                      c.invRhoCssq;
    constexpr float kx = cx != 0 ? DIAG1 : DIAG0;
    constexpr float ky = cy != 0 ? DIAG1 : DIAG0;
    constexpr float kz = cz != 0 ? DIAG1 : DIAG0;
    float s = kx * c.pxx + ky * c.pyy;
    s = s + kz * c.pzz;
    if constexpr (cx * cy != 0) s = acc<cx * cy>(s, 2.0f * c.pxy);
    if constexpr (cx * cz != 0) s = acc<cx * cz>(s, 2.0f * c.pxz);
    if constexpr (cy * cz != 0) s = acc<cy * cz>(s, 2.0f * c.pyz);
    const float fneq = d.fcoef[WC[L]] * s;
    return feq + d.one_m_omega * fneq + HeF;
}

// ---------------------------------------------------------- collisionPhase
struct PhaseCell {
    float phi, ux, uy, uz, nx, ny, nz, phi_norm;
};

MCB_HD PhaseCell phase_cell(const Phys& p, float phi, float ux, float uy, float uz, float nx, float ny, float nz) {
    PhaseCell c;
    c.phi = phi; c.ux = ux; c.uy = uy; c.uz = uz; c.nx = nx; c.ny = ny; c.nz = nz;
    c.phi_norm = p.sharp_c * phi * (1.0f - phi);   // lbm.cu:319
    return c;
}

// lbm.cu:323-326: the reference's full-relaxation phase collision.
template <int L>
MCB_HD float phase_post(const PhaseCell& c, const Phys& p) {
    const float w = p.w[WC[L]];
    const float udotc = cdot<L>(c.ux, c.uy, c.uz) * INV_CSSQ;
    const float geq = w * c.phi * (1.0f + udotc);
    const float Ai = c.phi_norm * cdot<L>(c.nx, c.ny, c.nz);
    return geq + w * Ai;
}

// Phase relaxation time extension (no reference counterpart): BGK towards
// the reference's augmented equilibrium; equals phase_post at omega_g = 1.
template <int L>
MCB_HD float phase_post_bgk(const PhaseCell& c, const Phys& p, const Derived& d, float gpre) {
    return d.one_m_omega_g * gpre + p.omega_g * phase_post<L>(c, p);
}

}  // namespace mcb
