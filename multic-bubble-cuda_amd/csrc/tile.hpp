// Tile bodies of the two passes of a step, shared by the single-case and the
// batched kernels and by every population storage form.
//
// Storage forms (template parameter S):
//   StoreF32  : populations stored as FP32, read and written as computed.
//               This form reproduces the FP32 transcription bit for bit.
//   StoreF16 / StoreBF16 : populations stored in 16 bits as offsets:
//                 f*_l(y) = w_l + t_l(y) / K_l              (from the lattice weight)
//                 g*_l(y) = w_l phi(y) + t_l(y) / K_l       (tau_g != 1)
//               where phi(y) is the phase the source cell had when it
//               collided (an FP32 field). At tau_g = 1 g is not stored as
//               populations at all but in a moment form: (phi, u, n) in FP32
//               for StoreF32 (phase_tile_gm; phase_tile_gm_rn recomputes n
//               from phi instead), (phi, B) with B in FP16 for the
//               16-bit forms (phase_tile_ab). The offsets are small (density, velocity
//               and sharpening terms), so 16 bits keep about 1e-5 of w_l
//               (FP16). K_l is a power of two near 2^12 / w_l: an
//               exact scale, so encoding and decoding carry no systematic
//               bias (a rounded 1/w_l would scale every population by the
//               same w_l * fl(1/w_l) != 1 and drift the mass each step), and
//               large enough that offsets down to ~1e-8 w_l stay in FP16's
//               normal range instead of losing bits as subnormals. FP16's
//               largest value then bounds an offset to 12 w_0, 18 w_1, 18 w_2.
//               All arithmetic stays FP32. Each cell's rounding residual of
//               links 1..18 is folded into its rest link (l = 0), so the
//               cell's total is stored with one 16-bit rounding only.
// FP32 with g in moment form can also keep f as the 13 values collisionFluid
// reads (collide_tile_fm) instead of 19 populations.
#pragma once
#include "fused.hpp"
#include "lattice_fast.hpp"

namespace mcb {

// Offset scale K of the weight classes (rest, face, edge) and its inverse.
__host__ __device__ constexpr float off_scale(int wc) {
    return wc == 0 ? 16384.0f : (wc == 1 ? 65536.0f : 131072.0f);
}
__host__ __device__ constexpr float off_unscale(int wc) {
    return wc == 0 ? 0x1p-14f : (wc == 1 ? 0x1p-16f : 0x1p-17f);
}

// Population of a 16-bit form from its anchor and its decoded offset t.
template <int L> __device__ __forceinline__ float pop_from_offset(const Phys& p, float anchor, float t) {
    return __builtin_fmaf(p.w[WC[L]], anchor, t * off_unscale(WC[L]));
}
// Offset of population v against its anchor (before rounding to 16 bits).
template <int L> __device__ __forceinline__ float offset_of_pop(const Phys& p, float anchor, float v) {
    return __builtin_fmaf(-p.w[WC[L]], anchor, v) * off_scale(WC[L]);
}

__device__ __forceinline__ int wrapmod(int v, int n) {
    v %= n;
    return v < 0 ? v + n : v;
}

// REFERENCE boundaries (CaseView::bc): whether (x, y, z) is a face cell (or
// lies beyond one), and the interior cell boundaryConditions copies a face
// cell's phi from (src/lbm.cu:403-425 run z faces, then y, then x: every
// coordinate clamped to [1, n - 2]).
__device__ __forceinline__ bool ref_face(const Geo& G, int x, int y, int z) {
    return x <= 0 || x >= G.nx - 1 || y <= 0 || y >= G.ny - 1 || z <= 0 || z >= G.nz - 1;
}
__device__ __forceinline__ unsigned ref_clamped(const Geo& G, int x, int y, int z) {
    x = min(max(x, 1), G.nx - 2);
    y = min(max(y, 1), G.ny - 2);
    z = min(max(z, 1), G.nz - 2);
    return (unsigned)z * (unsigned)G.nxy + (unsigned)(y * G.nx + x);
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
__host__ __device__ __forceinline__ unsigned xcd_order(unsigned lid, unsigned total) {
    const unsigned q = total / kXcds, r = total % kXcds;
    const unsigned xcd = lid % kXcds, slot = lid / kXcds;
    return (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + slot;
}
// A milder renumbering: the device keeps working on one contiguous stretch
// of tiles, as in dispatch order, but each XCD takes runs of `run`
// consecutive tiles of it (within every window of 8 * run workgroups, XCD j
// takes tiles [window + j * run, window + (j + 1) * run)), so x-neighbouring
// tiles of a run share their halo lines in one L2. A last partial window
// keeps dispatch order. Bijective on [0, total).
__host__ __device__ __forceinline__ unsigned xcd_runs(unsigned lid, unsigned total, unsigned run) {
    const unsigned w = kXcds * run;
    if (lid >= total / w * w) return lid;
    return lid / w * w + (lid % kXcds) * run + (lid / kXcds) % run;
}

// ------------------------------------------------------------ storage forms

struct StoreF32 {
    static constexpr bool kAnchored = false;
    static constexpr unsigned kBytes = 4;
};
struct StoreF16 {
    static constexpr bool kAnchored = true;
    static constexpr unsigned kBytes = 2;
    static __device__ __forceinline__ float dec(unsigned short b) { return (float)__builtin_bit_cast(_Float16, b); }
    static __device__ __forceinline__ unsigned short enc(float v) {
        return __builtin_bit_cast(unsigned short, (_Float16)v);
    }
};
struct StoreBF16 {
    static constexpr bool kAnchored = true;
    static constexpr unsigned kBytes = 2;
    static __device__ __forceinline__ float dec(unsigned short b) { return (float)__builtin_bit_cast(__bf16, b); }
    static __device__ __forceinline__ unsigned short enc(float v) { return __builtin_bit_cast(unsigned short, (__bf16)v); }
};

// The 19 populations of one buffer behind one buffer resource (see Planes
// in common.hpp); 16-bit forms move raw offsets and are always kept in
// planes.
template <class S> struct PopPlanes;
template <> struct PopPlanes<StoreF32> : Planes {
    using Raw = float;   // a population as loaded (see raw())
    __device__ __forceinline__ PopPlanes(const void* base, const Geo& G) : Planes(static_cast<const float*>(base), G) {}
    template <int L> __device__ __forceinline__ Raw raw(unsigned vb, unsigned sb) const { return ldb<L>(vb, sb); }
    static __device__ __forceinline__ float dec(Raw v) { return v; }
};
template <class S> struct PopPlanes {
    __amdgpu_buffer_rsrc_t r;
    unsigned plane_bytes;
    __device__ __forceinline__ PopPlanes(const void* base, const Geo& G)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)(19u * (unsigned)G.plane * 2u),
                                              kBufferConfigWord)),
          plane_bytes((unsigned)G.plane * 2u) {}
    __device__ __forceinline__ unsigned col(unsigned e) const { return e << 1; }
    __device__ __forceinline__ unsigned zb(unsigned zc) const { return zc << 1; }
    template <int L> __device__ __forceinline__ float ld(unsigned elem) const {
        return S::dec(__builtin_amdgcn_raw_buffer_load_b16(r, elem << 1, L * plane_bytes, 0));
    }
    // Stores the 16-bit form of v and returns the value it decodes to.
    template <int L> __device__ __forceinline__ float st(unsigned elem, float v) const {
        const unsigned short b = S::enc(v);
        __builtin_amdgcn_raw_buffer_store_b16(b, r, elem << 1, L * plane_bytes, MCB_STORE16_AUX);
        return S::dec(b);
    }
    // Split addressing, as Planes::ldb / stb.
    template <int L> __device__ __forceinline__ float ldb(unsigned vb, unsigned sb) const {
        return S::dec(__builtin_amdgcn_raw_buffer_load_b16(r, vb, L * plane_bytes + sb, 0));
    }
    // The stored 16 bits, decoded later with dec(): lets a caller issue the
    // loads of the next plane early and keep them in flight as raw bits.
    using Raw = unsigned short;
    template <int L> __device__ __forceinline__ Raw raw(unsigned vb, unsigned sb) const {
        return __builtin_amdgcn_raw_buffer_load_b16(r, vb, L * plane_bytes + sb, 0);
    }
    static __device__ __forceinline__ float dec(Raw b) { return S::dec(b); }
    template <int L> __device__ __forceinline__ float stb(unsigned vb, unsigned sb, float v) const {
        const unsigned short b = S::enc(v);
        __builtin_amdgcn_raw_buffer_store_b16(b, r, vb, L * plane_bytes + sb, MCB_STORE16_AUX);
        return S::dec(b);
    }
};

// B = 3 phi u + phi_n n, the vector of the 16-bit forms' (phi, B) moment
// form of g, is kept in FP16 scaled by 2^10: an exact scale, |B| (a few
// tenths at most) stays far below FP16's largest value / 2^10, and values
// down to ~6e-8 stay normal. B only moves phi between cells (its terms cancel
// in the sum of phi), so its rounding does not touch mass conservation.
constexpr float kBScale = 1024.0f, kBUnscale = 0x1p-10f;
__device__ __forceinline__ unsigned short enc_b(float v) {
    return __builtin_bit_cast(unsigned short, (_Float16)(v * kBScale));
}
__device__ __forceinline__ float ld_b(const float* field, unsigned elem) {
    typedef const unsigned short __attribute__((address_space(1)))* gp16;
    return (float)__builtin_bit_cast(_Float16, ((gp16)field)[elem]) * kBUnscale;
}
// 16-bit stores into one field behind a buffer resource (as Field).
struct Field16 {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ Field16(const void* base, unsigned bytes)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, kBufferConfigWord)) {}
    __device__ __forceinline__ void st(unsigned vb, unsigned sb, unsigned short v) const {
        __builtin_amdgcn_raw_buffer_store_b16(v, r, vb, sb, 0);
    }
};

// In-plane byte offsets of the 9 columns (x - c_x, y - c_y) that a cell's
// pull-form gathers read, in the layout of the population buffers (planes'
// col). They do not depend on z, so a tile computes them once; per plane only
// the wave-uniform plane offset (zb) changes, and it goes into the scalar
// offset of the access.
struct Columns {
    unsigned b[9];
    template <class P>
    __device__ __forceinline__ Columns(const unsigned (&xs)[3], const unsigned (&ys)[3], const P& planes) {
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int i = 0; i < 3; ++i) b[j * 3 + i] = planes.col(ys[j] + xs[i]);
    }
    template <int L> __device__ __forceinline__ unsigned src() const { return b[(1 - CY[L]) * 3 + (1 - CX[L])]; }
    __device__ __forceinline__ unsigned me() const { return b[4]; }
};

// ------------------------------------------------------------------ phase

// Moment form (FP32, tau_g = 1): g*_l(y) is recomputed from the source cell's
// phi, u and normal n with collisionPhase's own expressions (lbm.cu:319-326),
// so the phase sum carries the same bits as with stored populations. n is
// either stored by the collide pass (CaseView::gm_n set) or recomputed from
// phi around the source with gradCalc's expressions (lbm.cu:50-71), which
// gives the bits the collide pass computed. The initial state is initDist's
// g_l(x) = w_l phi0(x) (init.cu:69-79).

// n at cell (x, y, z) of the state phi (c.phi_prev), read from global memory,
// as the collide pass computed it; coordinates up to one cell outside the
// lattice name its periodic image. A ghost-layer face keeps n = 0 (gradCalc
// updates interior cells only).
__device__ __forceinline__ void normal_at(const CaseView& c, int x, int y, int z, float& nx, float& ny, float& nz) {
    const Geo& G = c.geo;
    if (c.bc == MCB_BC_REFERENCE && ref_face(G, x, y, z)) {
        nx = ny = nz = 0.0f;
        return;
    }
    const Derived d(c.phys);
    auto nb = [&](auto lc) {
        constexpr int l = decltype(lc)::value;
        return ld32(c.phi_prev, (unsigned)(wrapmod(z + CZ[l], G.nz) * G.nxy + wrapmod(y + CY[l], G.ny) * G.nx +
                                           wrapmod(x + CX[l], G.nx)));
    };
    float ind;
    grad_normal(nb, d, nx, ny, nz, ind);
}

// g*_l of the source cell (x, y, z) (periodic images allowed), from global
// memory (the moment form's readout as populations).
template <int L> __device__ __forceinline__ float g_from_moments(const CaseView& c, int x, int y, int z) {
    const Geo& G = c.geo;
    const unsigned o = (unsigned)(wrapmod(z, G.nz) * G.nxy + wrapmod(y, G.ny) * G.nx + wrapmod(x, G.nx));
    float nx, ny, nz;
    if (c.gm_n[0]) {
        nx = ld32(c.gm_n[0], o);
        ny = ld32(c.gm_n[1], o);
        nz = ld32(c.gm_n[2], o);
    } else {
        normal_at(c, x, y, z, nx, ny, nz);
    }
    const PhaseCell pc = phase_cell(c.phys, ld32(c.phi_prev, o), ld32(c.gm_u[0], o), ld32(c.gm_u[1], o),
                                    ld32(c.gm_u[2], o), nx, ny, nz);
    return phase_post<L>(pc, c.phys);
}

// phiCalc (src/lbm.cu:18-25) for one cell of a pull-form state: the
// post-stream g_l(x) is g*_l(x - c_l). xs/ys/zs: offsets of x-1, x, x+1 etc.
// From stored populations, or before the first step from initDist's
// g_l = w_l phi0 (GM_INITIAL); the moment form runs phase_tile_gm (FP32) or
// phase_tile_ab (16-bit forms) instead.
// col: the columns of xs/ys in gp's layout.
template <class S>
__device__ __forceinline__ float phase_sum(const CaseView& c, const PopPlanes<S>& gp, const Columns& col,
                                           const unsigned (&xs)[3], const unsigned (&ys)[3],
                                           const unsigned (&zs)[3]) {
    float sum = 0.0f;
    if (c.gm == GM_INITIAL) {
        const float p0 = ld32(c.phi_prev, zs[1] + ys[1] + xs[1]);
        for_links([&](auto lc) { sum += c.phys.w[WC[decltype(lc)::value]] * p0; });
        return sum;
    }
    for_links([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        const float v = gp.template ldb<l>(col.template src<l>(), gp.zb(zs[1 - CZ[l]]));
        if constexpr (S::kAnchored)
            sum += pop_from_offset<l>(c.phys, ld32(c.phi_prev, zs[1 - CZ[l]] + ys[1 - CY[l]] + xs[1 - CX[l]]), v);
        else sum += v;
    });
    return sum;
}

// Links that a halo-ring cell sends into the tile, by side of the ring:
// 0 = row y0-1 (c_y = +1), 1 = row y0+TY (c_y = -1), 2 = column x0-1
// (c_x = +1, corners included), 3 = column x0+TX (c_x = -1). With p = s e_a
// (a the side's axis, s its sign) and b the other in-plane axis, the five
// links are p, p + e_b, p - e_b, p + e_z, p - e_z, in that order.
constexpr int RING_LINK[4][5] = {{3, 7, 14, 11, 17}, {4, 13, 8, 18, 12}, {1, 7, 13, 9, 15}, {2, 14, 8, 16, 10}};
constexpr bool ring_links_ok() {
    for (int side = 0; side < 4; ++side) {
        const int s = side % 2 == 0 ? 1 : -1;
        const int pa[3] = {side < 2 ? 0 : s, side < 2 ? s : 0, 0};
        const int eb[3] = {side < 2 ? 1 : 0, side < 2 ? 0 : 1, 0};
        const int want[5][3] = {{pa[0], pa[1], 0},
                                {pa[0] + eb[0], pa[1] + eb[1], 0},
                                {pa[0] - eb[0], pa[1] - eb[1], 0},
                                {pa[0], pa[1], 1},
                                {pa[0], pa[1], -1}};
        for (int k = 0; k < 5; ++k) {
            const int l = RING_LINK[side][k];
            if (CX[l] != want[k][0] || CY[l] != want[k][1] || CZ[l] != want[k][2]) return false;
        }
    }
    return true;
}
static_assert(ring_links_ok(), "ring link table");

// One link value of collisionPhase at tau_g = 1 (lbm.cu:323-326, phase_post)
// from the link's dot products cu = c.u and cn = c.n, with wphi = w phi formed
// once per weight class. Same operations in the same order as phase_post.
__device__ __forceinline__ float phase_link(float w, float wphi, float phi_norm, float cu, float cn) {
    return wphi * (1.0f + cu * INV_CSSQ) + w * (phi_norm * cn);
}

// Cells of the 1-cell ring around a TX x TY tile that send links into it
// (two rows of TX cells and two columns of TY + 2, corners included).
template <int TX, int TY> constexpr int gm_ring_cells() { return 2 * TX + 2 * (TY + 2); }

// LDS slots of phase_tile_gm_rn's link values, link-major. Link l (1..18) has a
// row-major region of the cells that lie in the tile or send l into it:
// x in [-max(c_x, 0), TX - 1 + max(-c_x, 0)] and likewise y, i.e.
// (TX + |c_x|) x (TY + |c_y|) cells (the rest link stays in a register).
// A tile cell writes every link at its own position and reads link l at its
// position - c_l, both at a compile-time offset from one of two row bases;
// a ring corner's links that enter neither the region nor the tile go to one
// spare slot that nothing reads.
template <int TX, int TY> struct GmSlots {
    static constexpr int w(int l) { return TX + (CX[l] != 0 ? 1 : 0); }
    static constexpr int h(int l) { return TY + (CY[l] != 0 ? 1 : 0); }
    static constexpr int lo_x(int l) { return CX[l] > 0 ? -1 : 0; }
    static constexpr int lo_y(int l) { return CY[l] > 0 ? -1 : 0; }
    static constexpr int base(int l) {
        int b = 0;
        for (int k = 1; k < l; ++k) b += w(k) * h(k);
        return b;
    }
    static constexpr int kSpare = base(Q);   // after the 18 regions
    static constexpr int kSlots = kSpare + 1;
    static constexpr bool inside(int l, int sx, int sy) {
        return sx >= lo_x(l) && sx < lo_x(l) + w(l) && sy >= lo_y(l) && sy < lo_y(l) + h(l);
    }
    static constexpr int at(int l, int sx, int sy) { return base(l) + (sy - lo_y(l)) * w(l) + (sx - lo_x(l)); }
    // Offsets from the row base sy * w(l) + sx of a tile cell: its own slot,
    // and the slot of its source x - c_l.
    static constexpr int own(int l) { return base(l) - lo_y(l) * w(l) - lo_x(l); }
    static constexpr int src(int l) { return own(l) - CY[l] * w(l) - CX[l]; }
};

template <int TX, int TY> constexpr bool gm_slots_ok() {
    using L = GmSlots<TX, TY>;
    for (int l = 1; l < Q; ++l) {
        if (L::own(l) < 0 || L::src(l) < 0) return false;
        // Every source of the tile and every tile cell lies in the region.
        for (int sy = -1; sy <= TY; ++sy)
            for (int sx = -1; sx <= TX; ++sx) {
                const bool tile = sx >= 0 && sx < TX && sy >= 0 && sy < TY;
                const bool sends = sx + CX[l] >= 0 && sx + CX[l] < TX && sy + CY[l] >= 0 && sy + CY[l] < TY;
                if ((tile || sends) && !L::inside(l, sx, sy)) return false;
            }
    }
    return true;
}

// Moment-form phase pass with stored normals (CaseView::gm_n set: the
// collide pass stores n beside u) for the TX x TY column (x0, y0) over
// planes [zb, ze). For destination plane z, every cell y of the tile and of the
// 1-cell ring around it writes to LDS the values g*_l(y) that links l carry
// from it into plane z (from its plane z - c_z(l)), computed with
// collisionPhase's own operations (phase_post, lbm.cu:319-326); each
// destination then reads its 19 values and sums them in link order, so the
// result carries the stored populations' bits.
//   Each value is computed once. In-plane opposite links l, l' (c_l' = -c_l)
// of a cell share their work: their dot products are exact negatives (IEEE
// rounding is sign-symmetric), so one evaluation gives both.
//   Thread (tx, ty) owns tile cell (tx, ty), whose source fields (phi, u, n)
// are kept in a 3-plane register window (z, z+1, z+2 in flight). On plane z
// it writes its in-plane links and its c_z = -1 links from plane z+1, and
// computes its c_z = +1 links from plane z, held (5 registers) for plane
// z+1. Threads tid < gm_ring_cells also own one ring cell, of which they
// compute only the five links that enter the tile: its plane z+1 is read at
// the start of plane z and used after the tile cell's work; what later
// planes need of it is held (5 values). This keeps a lane within 80 VGPRs,
// i.e. three 8-wave workgroups per CU. Slots are compile-time (the z loop
// is unrolled by 6, the period of the window and of the ring's held
// values). Must be called by all TX * TY threads (it contains barriers).
template <int TX, int TY>
__device__ __forceinline__ void phase_tile_gm(const CaseView& c, int x0, int y0, int zb, int ze) {
    constexpr int NT = TX * TY;
    constexpr int HX = TX + 2, HY = TY + 2, HN = HX * HY;
    constexpr int NR = gm_ring_cells<TX, TY>();
    static_assert(NR <= NT, "a lane per ring cell");
    __shared__ float s_g[HN * Q];   // [halo cell][link]: g*_l of the source cell

    const Geo& G = c.geo;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
    const float w1 = c.phys.w[1], w2 = c.phys.w[2];
    auto fetch = [&](float (&r)[7], unsigned o) {
        r[0] = ld32(c.phi_prev, o);
        r[1] = ld32(c.gm_u[0], o);
        r[2] = ld32(c.gm_u[1], o);
        r[3] = ld32(c.gm_u[2], o);
        r[4] = ld32(c.gm_n[0], o);
        r[5] = ld32(c.gm_n[1], o);
        r[6] = ld32(c.gm_n[2], o);
    };
    auto plane_of = [&](int z) { return (unsigned)(wrapmod(z, G.nz) * G.nxy); };
    auto cell = [&](const float (&r)[7]) { return phase_cell(c.phys, r[0], r[1], r[2], r[3], r[4], r[5], r[6]); };

    // ---- tile cell
    const int x = x0 + tx, y = y0 + ty;
    const bool valid = x < G.nx && y < G.ny;
    const int e0 = (ty + 1) * HX + (tx + 1);
    // Cells past the lattice edge (partial tiles) take their periodic image.
    const unsigned own = (unsigned)(wrapmod(y, G.ny) * G.nx + wrapmod(x, G.nx));
    float f[3][7];     // window: phi, ux, uy, uz, nx, ny, nz on planes by slot
    float hold[5];     // c_z = +1 links for plane z+1, from plane z
    // ---- ring cell (side 0: row y0-1, 1: row y0+TY, 2: column x0-1, 3: column x0+TX)
    const bool ring = tid < NR;
    int hx, hy, side;
    if (tid < TX) { hx = tid + 1; hy = 0; side = 0; }
    else if (tid < 2 * TX) { hx = tid - TX + 1; hy = TY + 1; side = 1; }
    else if (tid < 2 * TX + TY + 2) { hx = 0; hy = tid - 2 * TX; side = 2; }
    else { hx = TX + 1; hy = min(tid - 2 * TX - TY - 2, TY + 1); side = 3; }
    const unsigned roff = (unsigned)(wrapmod(y0 - 1 + hy, G.ny) * G.nx + wrapmod(x0 - 1 + hx, G.nx));
    const int rbase = (hy * HX + hx) * Q;
    // The side's five link numbers, 5 bits each.
    unsigned rlinks = 0;
#pragma unroll
    for (int k = 0; k < 5; ++k)
        rlinks |= (unsigned)(side == 0 ? RING_LINK[0][k]
                                       : side == 1 ? RING_LINK[1][k] : side == 2 ? RING_LINK[2][k] : RING_LINK[3][k])
                  << (5 * k);
    auto rslot = [&](int k) { return s_g + rbase + (int)((rlinks >> (5 * k)) & 31u); };
    const bool ry = side < 2;                     // the side's axis is y
    const float rs = (side & 1) ? -1.0f : 1.0f;   // and its sign
    float rin[3];      // in-plane links of the ring cell for the next plane
    float rhold[2];    // its c_z = +1 link for planes by parity
    // From the ring cell's plane p: in-plane links (for plane p) into rin,
    // c_z = -1 link (for plane p-1) returned, c_z = +1 link (plane p+1) to up.
    auto ring_plane = [&](const float (&r)[7], float& up) {
        const PhaseCell q = cell(r);
        const float pu = rs * (ry ? q.uy : q.ux), qu = ry ? q.ux : q.uy;
        const float pn = rs * (ry ? q.ny : q.nx), qn = ry ? q.nx : q.ny;
        const float w1phi = w1 * q.phi, w2phi = w2 * q.phi;
        rin[0] = phase_link(w1, w1phi, q.phi_norm, pu, pn);
        rin[1] = phase_link(w2, w2phi, q.phi_norm, pu + qu, pn + qn);
        rin[2] = phase_link(w2, w2phi, q.phi_norm, pu - qu, pn - qn);
        up = phase_link(w2, w2phi, q.phi_norm, pu + q.uz, pn + q.nz);
        return phase_link(w2, w2phi, q.phi_norm, pu - q.uz, pn - q.nz);
    };

    // Window slot of plane p: (p - zb) mod 3 (plane zb-1 passes through slot
    // 2); the ring's held c_z = +1 link for plane p: rhold[(p - zb) mod 2].
    fetch(f[2], plane_of(zb - 1) + own);
    fetch(f[0], plane_of(zb) + own);
    fetch(f[1], plane_of(zb + 1) + own);
    {
        const PhaseCell pm = cell(f[2]);
        int k = 0;
        for_links([&](auto lc) {
            constexpr int l = decltype(lc)::value;
            if constexpr (CZ[l] > 0) { hold[k] = phase_post<l>(pm, c.phys); ++k; }
        });
        if (ring) {
            float r[7];
            fetch(r, plane_of(zb - 1) + roff);
            ring_plane(r, rhold[0]);
            fetch(r, plane_of(zb) + roff);
            ring_plane(r, rhold[1]);   // leaves plane zb's in-plane links in rin
        }
    }
    auto iter = [&](int z, auto p0, auto pp, auto pn, auto ph) {
        constexpr int P0 = decltype(p0)::value, PP = decltype(pp)::value, PN = decltype(pn)::value;
        constexpr int PH = decltype(ph)::value;
        float r[7];
        if (ring) fetch(r, plane_of(z + 1) + roff);            // used after the tile cell's work
        if (z + 1 < ze) fetch(f[PN], plane_of(z + 2) + own);   // read by the next plane
        float* dst = s_g + e0 * Q;
        {   // c_z = +1 links from plane z-1, computed on the previous plane
            int k = 0;
            for_links([&](auto lc) {
                constexpr int l = decltype(lc)::value;
                if constexpr (CZ[l] > 0) { dst[l] = hold[k]; ++k; }
            });
        }
        {   // in-plane links from plane z
            const PhaseCell c0 = cell(f[P0]);
            dst[0] = phase_post<0>(c0, c.phys);
            for_links([&](auto lc) {
                constexpr int l = decltype(lc)::value;
                if constexpr (l != 0 && CZ[l] == 0 && l % 2 == 1) {
                    static_assert(CX[l + 1] == -CX[l] && CY[l + 1] == -CY[l] && CZ[l + 1] == 0, "pair");
                    const float wl = c.phys.w[WC[l]];
                    const float ud = cdot<l>(c0.ux, c0.uy, c0.uz) * INV_CSSQ;
                    const float wa = wl * (c0.phi_norm * cdot<l>(c0.nx, c0.ny, c0.nz));
                    const float wphi = wl * c0.phi;
                    dst[l] = wphi * (1.0f + ud) + wa;
                    dst[l + 1] = wphi * (1.0f + -ud) + -wa;
                }
            });
            int k = 0;
            for_links([&](auto lc) {
                constexpr int l = decltype(lc)::value;
                if constexpr (CZ[l] > 0) { hold[k] = phase_post<l>(c0, c.phys); ++k; }
            });
        }
        {   // c_z = -1 links from plane z+1
            const PhaseCell cp = cell(f[PP]);
            for_links([&](auto lc) {
                constexpr int l = decltype(lc)::value;
                if constexpr (CZ[l] < 0) dst[l] = phase_post<l>(cp, c.phys);
            });
        }
        if (ring) {
            *rslot(3) = rhold[PH];
            *rslot(0) = rin[0];
            *rslot(1) = rin[1];
            *rslot(2) = rin[2];
            *rslot(4) = ring_plane(r, rhold[PH]);
        }
        __syncthreads();
        if (valid) {
            float sum = 0.0f;
            for_links([&](auto lc) {
                constexpr int l = decltype(lc)::value;
                sum += s_g[(e0 - CY[l] * HX - CX[l]) * Q + l];
            });
            // Ghost-layer faces take boundaryConditions' copy instead.
            if (c.bc == MCB_BC_REFERENCE && ref_face(G, x, y, z)) sum = ld32(c.phi_prev, ref_clamped(G, x, y, z));
            st32(c.phi, (unsigned)z * (unsigned)G.nxy + own, sum);
        }
        __syncthreads();   // s_g is rewritten by the next plane
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    for (int z = zb; z < ze; z += 6) {
        iter(z, I0{}, I1{}, I2{}, I0{});
        if (z + 1 < ze) iter(z + 1, I1{}, I2{}, I0{}, I1{});
        if (z + 2 < ze) iter(z + 2, I2{}, I0{}, I1{}, I0{});
        if (z + 3 < ze) iter(z + 3, I0{}, I1{}, I2{}, I1{});
        if (z + 4 < ze) iter(z + 4, I1{}, I2{}, I0{}, I0{});
        if (z + 5 < ze) iter(z + 5, I2{}, I0{}, I1{}, I1{});
    }
}

// Moment-form phase pass with recomputed normals (CaseView::gm_n null) for
// the TX x TY column (x0, y0) over planes [zb, ze). For destination plane z, every cell y of the tile and of the
// 1-cell ring around it writes to LDS the values g*_l(y) that links l carry
// from it into plane z (from its plane z - c_z(l)), computed with
// collisionPhase's own operations (phase_post, lbm.cu:319-326) from its phi,
// its u and its normal n; each destination then sums its 19 values in link
// order, so the result carries the stored populations' bits.
//   n is not state: it is recomputed here with gradCalc's operations from
// the same phi the collide pass computed it from (this pass's input), so a
// step moves phi and u (16 B per cell) instead of phi, u and n (28 B), and
// the collide pass stores u only. phi planes with a 2-cell halo are staged
// in a 3-slot LDS ring (a plane is loaded once per chunk plus the chunk's
// 4 boundary planes).
//   Each link value is computed once. In-plane opposite links l, l'
// (c_l' = -c_l) of a cell share their work: their dot products are exact
// negatives (IEEE rounding is sign-symmetric), so one evaluation gives both.
//   Thread (tx, ty) owns tile cell (tx, ty): a 3-plane register window of
// its u (z, z+1, z+2 in flight) and its n on planes z, z+1. On plane z it
// computes n on plane z+1, writes its in-plane links and its c_z = -1 links
// from plane z+1, and computes its c_z = +1 links from plane z, held (5
// registers) for plane z+1. Threads tid < gm_ring_cells also own one ring
// cell, of which they compute the normal and only the five links that enter
// the tile; what later planes need of it is held (5 values). Slots are
// compile-time (the z loop is unrolled by 6, the period of the window and of
// the ring's held values). Must be called by all TX * TY threads (it contains
// barriers).
template <int TX, int TY>
__device__ __forceinline__ void phase_tile_gm_rn(const CaseView& c, int x0, int y0, int zb, int ze) {
    constexpr int NT = TX * TY;
    constexpr int PX = TX + 4, PY = TY + 4, PN = PX * PY;   // phi plane + 2-halo
    constexpr int KP = (PN + NT - 1) / NT;
    constexpr int NR = gm_ring_cells<TX, TY>();
    using SL = GmSlots<TX, TY>;
    static_assert(NR <= NT, "a lane per ring cell");
    static_assert(gm_slots_ok<TX, TY>(), "link slots");
    static_assert(SL::kSlots < 16384, "LDS offsets");
    __shared__ float s_g[SL::kSlots];
    __shared__ float s_phi[3][PN];

    const Geo& G = c.geo;
    const Derived d(c.phys);
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
    const float w1 = c.phys.w[1], w2 = c.phys.w[2];
    const unsigned fbytes = (unsigned)(G.n * 4);
    const Field phi_f(c.phi_prev, fbytes), ux_f(c.gm_u[0], fbytes), uy_f(c.gm_u[1], fbytes),
        uz_f(c.gm_u[2], fbytes);
    auto plane4 = [&](int z) { return (unsigned)(wrapmod(z, G.nz) * G.nxy) * 4u; };
    auto fetch_u = [&](float (&u)[3], unsigned cell4, unsigned z4) {
        u[0] = ux_f.ld(cell4, z4);
        u[1] = uy_f.ld(cell4, z4);
        u[2] = uz_f.ld(cell4, z4);
    };

    // In-plane byte offsets of the staged phi cells this thread loads.
    unsigned poff[KP];
#pragma unroll
    for (int q = 0; q < KP; ++q) {
        const int e = tid + q * NT;
        unsigned off = 0;
        if (e < PN) {
            const int lx = e % PX, ly = e / PX;
            off = (unsigned)(wrapmod(y0 - 2 + ly, G.ny) * G.nx + wrapmod(x0 - 2 + lx, G.nx)) * 4u;
        }
        poff[q] = off;
    }
    auto load_phi = [&](int z, float* dst) {
        const unsigned z4 = plane4(z);
#pragma unroll
        for (int q = 0; q < KP; ++q) {
            const int e = tid + q * NT;
            if (q < KP - 1 || e < PN) dst[e] = phi_f.ld(poff[q], z4);
        }
    };
    // Software-pipelined staging inside the z loop: fetch_phi() loads plane
    // z+3 into registers during plane z, put_phi() writes it to its slot once
    // every read of plane z (the slot's previous plane) is done.
    float preg[KP];
    auto fetch_phi = [&](int z) {
        const unsigned z4 = plane4(z);
#pragma unroll
        for (int q = 0; q < KP; ++q) {
            const int e = tid + q * NT;
            if (q < KP - 1 || e < PN) preg[q] = phi_f.ld(poff[q], z4);
        }
    };
    auto put_phi = [&](float* dst) {
#pragma unroll
        for (int q = 0; q < KP; ++q) {
            const int e = tid + q * NT;
            if (q < KP - 1 || e < PN) dst[e] = preg[q];
        }
    };
    // n of the cell staged at pc on the plane whose phi is p0 (pm below, pp
    // above), as the collide pass computes it; (gx, gy, gz) its lattice
    // coordinates for the ghost-layer test.
    auto normal = [&](int pc, int gx, int gy, int gz, const float* pm, const float* p0, const float* pp,
                      float (&n)[3]) {
        if (c.bc == MCB_BC_REFERENCE && ref_face(G, gx, gy, gz)) {
            n[0] = n[1] = n[2] = 0.0f;
            return;
        }
        auto nb = [&](auto lc) {
            constexpr int l = decltype(lc)::value;
            return pick<CZ[l]>(pm, p0, pp)[pc + CY[l] * PX + CX[l]];
        };
        float ind;
        grad_normal(nb, d, n[0], n[1], n[2], ind);
    };

    // ---- tile cell
    const int x = x0 + tx, y = y0 + ty;
    const bool valid = x < G.nx && y < G.ny;
    const int pc_own = (ty + 2) * PX + tx + 2;
    // Cells past the lattice edge (partial tiles) take their periodic image.
    const unsigned own4 = (unsigned)(wrapmod(y, G.ny) * G.nx + wrapmod(x, G.nx)) * 4u;
    const int row0 = ty * TX + tx, row1 = ty * (TX + 1) + tx;   // row bases of widths TX, TX + 1
    auto own_slot = [&](auto lc) -> float& {
        constexpr int l = decltype(lc)::value;
        return s_g[(SL::w(l) == TX ? row0 : row1) + SL::own(l)];
    };
    float u[3][3];     // window: u on planes by slot
    float nn[2][3];    // n on planes by parity
    float hold[5];     // c_z = +1 links for plane z+1, from plane z
    // ---- ring cell (side 0: row y0-1, 1: row y0+TY, 2: column x0-1, 3: column x0+TX)
    const bool ring = tid < NR;
    int hx, hy, side;
    if (tid < TX) { hx = tid + 1; hy = 0; side = 0; }
    else if (tid < 2 * TX) { hx = tid - TX + 1; hy = TY + 1; side = 1; }
    else if (tid < 2 * TX + TY + 2) { hx = 0; hy = tid - 2 * TX; side = 2; }
    else { hx = TX + 1; hy = min(tid - 2 * TX - TY - 2, TY + 1); side = 3; }
    const unsigned roff4 = (unsigned)(wrapmod(y0 - 1 + hy, G.ny) * G.nx + wrapmod(x0 - 1 + hx, G.nx)) * 4u;
    const int rgx = x0 - 1 + hx, rgy = y0 - 1 + hy;
    const int pc_ring = (hy + 1) * PX + hx + 1;
    // LDS slots of the side's five links at this cell.
    int rslot[5];
    {
        const int sx = hx - 1, sy = hy - 1;
        auto fill = [&](auto sc) {
            constexpr int s = decltype(sc)::value;
            auto one = [&](auto kc) {
                constexpr int k = decltype(kc)::value;
                constexpr int l = RING_LINK[s][k];
                rslot[k] = SL::inside(l, sx, sy) ? SL::at(l, sx, sy) : SL::kSpare;
            };
            one(std::integral_constant<int, 0>{});
            one(std::integral_constant<int, 1>{});
            one(std::integral_constant<int, 2>{});
            one(std::integral_constant<int, 3>{});
            one(std::integral_constant<int, 4>{});
        };
        if (side == 0) fill(std::integral_constant<int, 0>{});
        else if (side == 1) fill(std::integral_constant<int, 1>{});
        else if (side == 2) fill(std::integral_constant<int, 2>{});
        else fill(std::integral_constant<int, 3>{});
    }
    const bool ry = side < 2;                     // the side's axis is y
    const float rs = (side & 1) ? -1.0f : 1.0f;   // and its sign
    float rin[3];      // in-plane links of the ring cell for the next plane
    float rhold[2];    // its c_z = +1 link for planes by parity
    // From the ring cell's plane p: in-plane links (for plane p) into rin,
    // c_z = -1 link (for plane p-1) returned, c_z = +1 link (plane p+1) to up.
    auto ring_plane = [&](const PhaseCell& q, float& up) {
        const float pu = rs * (ry ? q.uy : q.ux), qu = ry ? q.ux : q.uy;
        const float pn = rs * (ry ? q.ny : q.nx), qn = ry ? q.nx : q.ny;
        const float w1phi = w1 * q.phi, w2phi = w2 * q.phi;
        rin[0] = phase_link(w1, w1phi, q.phi_norm, pu, pn);
        rin[1] = phase_link(w2, w2phi, q.phi_norm, pu + qu, pn + qn);
        rin[2] = phase_link(w2, w2phi, q.phi_norm, pu - qu, pn - qn);
        up = phase_link(w2, w2phi, q.phi_norm, pu + q.uz, pn + q.nz);
        return phase_link(w2, w2phi, q.phi_norm, pu - q.uz, pn - q.nz);
    };
    auto cell = [&](float phi, const float (&uu)[3], const float (&n)[3]) {
        return phase_cell(c.phys, phi, uu[0], uu[1], uu[2], n[0], n[1], n[2]);
    };

    // Prologue. Slot of plane p: (p - zb) mod 3 for the phi ring and the u
    // window (planes zb-2 and zb-1 pass through slots 1 and 2); parity
    // (p - zb) mod 2 for n and the ring's held c_z = +1 link.
    load_phi(zb - 2, s_phi[1]);
    load_phi(zb - 1, s_phi[2]);
    load_phi(zb, s_phi[0]);
    fetch_u(u[2], own4, plane4(zb - 1));
    fetch_u(u[0], own4, plane4(zb));
    fetch_u(u[1], own4, plane4(zb + 1));
    float rp[3];   // the ring cell's u on a prologue plane
    if (ring) fetch_u(rp, roff4, plane4(zb - 1));
    __syncthreads();
    {   // plane zb-1: the tile cell's c_z = +1 links, the ring cell's
        float nm[3];
        normal(pc_own, x, y, zb - 1, s_phi[1], s_phi[2], s_phi[0], nm);
        const PhaseCell pm = cell(s_phi[2][pc_own], u[2], nm);
        int k = 0;
        for_links([&](auto lc) {
            constexpr int l = decltype(lc)::value;
            if constexpr (CZ[l] > 0) { hold[k] = phase_post<l>(pm, c.phys); ++k; }
        });
        if (ring) {
            float rn[3];
            normal(pc_ring, rgx, rgy, zb - 1, s_phi[1], s_phi[2], s_phi[0], rn);
            ring_plane(cell(s_phi[2][pc_ring], rp, rn), rhold[0]);
        }
    }
    __syncthreads();
    load_phi(zb + 1, s_phi[1]);
    if (ring) fetch_u(rp, roff4, plane4(zb));
    __syncthreads();
    {   // plane zb: the tile cell's n, the ring cell's in-plane and c_z = +1 links
        normal(pc_own, x, y, zb, s_phi[2], s_phi[0], s_phi[1], nn[0]);
        if (ring) {
            float rn[3];
            normal(pc_ring, rgx, rgy, zb, s_phi[2], s_phi[0], s_phi[1], rn);
            ring_plane(cell(s_phi[0][pc_ring], rp, rn), rhold[1]);   // leaves plane zb's in-plane links in rin
        }
    }
    __syncthreads();
    load_phi(zb + 2, s_phi[2]);
    __syncthreads();

    auto iter = [&](int z, auto p0, auto pp, auto pn, auto ph) {
        constexpr int P0 = decltype(p0)::value, PP = decltype(pp)::value, PN = decltype(pn)::value;
        constexpr int PH = decltype(ph)::value;
        float ru[3];
        if (ring) fetch_u(ru, roff4, plane4(z + 1));   // used after the tile cell's work
        if (z + 1 < ze) {                              // read by the next plane
            fetch_u(u[PN], own4, plane4(z + 2));
            fetch_phi(z + 3);
        }
        // n on plane z+1 (phi planes z, z+1, z+2 in slots P0, PP, PN)
        normal(pc_own, x, y, z + 1, s_phi[P0], s_phi[PP], s_phi[PN], nn[PH ^ 1]);
        {   // c_z = +1 links from plane z-1, computed on the previous plane
            int k = 0;
            for_links([&](auto lc) {
                constexpr int l = decltype(lc)::value;
                if constexpr (CZ[l] > 0) { own_slot(lc) = hold[k]; ++k; }
            });
        }
        float g0;
        {   // in-plane links from plane z
            const PhaseCell c0 = cell(s_phi[P0][pc_own], u[P0], nn[PH]);
            g0 = phase_post<0>(c0, c.phys);
            for_links([&](auto lc) {
                constexpr int l = decltype(lc)::value;
                if constexpr (l != 0 && CZ[l] == 0 && l % 2 == 1) {
                    static_assert(CX[l + 1] == -CX[l] && CY[l + 1] == -CY[l] && CZ[l + 1] == 0, "pair");
                    const float wl = c.phys.w[WC[l]];
                    const float ud = cdot<l>(c0.ux, c0.uy, c0.uz) * INV_CSSQ;
                    const float wa = wl * (c0.phi_norm * cdot<l>(c0.nx, c0.ny, c0.nz));
                    const float wphi = wl * c0.phi;
                    own_slot(lc) = wphi * (1.0f + ud) + wa;
                    own_slot(std::integral_constant<int, l + 1>{}) = wphi * (1.0f + -ud) + -wa;
                }
            });
            int k = 0;
            for_links([&](auto lc) {
                constexpr int l = decltype(lc)::value;
                if constexpr (CZ[l] > 0) { hold[k] = phase_post<l>(c0, c.phys); ++k; }
            });
        }
        {   // c_z = -1 links from plane z+1
            const PhaseCell cp = cell(s_phi[PP][pc_own], u[PP], nn[PH ^ 1]);
            for_links([&](auto lc) {
                constexpr int l = decltype(lc)::value;
                if constexpr (CZ[l] < 0) own_slot(lc) = phase_post<l>(cp, c.phys);
            });
        }
        if (ring) {
            float rn[3];
            normal(pc_ring, rgx, rgy, z + 1, s_phi[P0], s_phi[PP], s_phi[PN], rn);
            s_g[rslot[3]] = rhold[PH];
            s_g[rslot[0]] = rin[0];
            s_g[rslot[1]] = rin[1];
            s_g[rslot[2]] = rin[2];
            s_g[rslot[4]] = ring_plane(cell(s_phi[PP][pc_ring], ru, rn), rhold[PH]);
        }
        __syncthreads();
        if (z + 1 < ze) put_phi(s_phi[P0]);   // plane z+3 replaces plane z
        if (valid) {
            float sum = 0.0f;
            sum += g0;
            for_links([&](auto lc) {
                constexpr int l = decltype(lc)::value;
                if constexpr (l != 0) sum += s_g[(SL::w(l) == TX ? row0 : row1) + SL::src(l)];
            });
            // Ghost-layer faces take boundaryConditions' copy instead.
            if (c.bc == MCB_BC_REFERENCE && ref_face(G, x, y, z)) sum = ld32(c.phi_prev, ref_clamped(G, x, y, z));
            st32(c.phi, (unsigned)z * (unsigned)G.nxy + own4 / 4u, sum);
        }
        __syncthreads();   // s_g is rewritten by the next plane, phi slot P0 read
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    for (int z = zb; z < ze; z += 6) {
        iter(z, I0{}, I1{}, I2{}, I0{});
        if (z + 1 < ze) iter(z + 1, I1{}, I2{}, I0{}, I1{});
        if (z + 2 < ze) iter(z + 2, I2{}, I0{}, I1{}, I0{});
        if (z + 3 < ze) iter(z + 3, I0{}, I1{}, I2{}, I1{});
        if (z + 4 < ze) iter(z + 4, I1{}, I2{}, I0{}, I0{});
        if (z + 5 < ze) iter(z + 5, I2{}, I0{}, I1{}, I1{});
    }
}

// Sum of a 64-bit integer over the workgroup (exact, hence the same in any
// order); valid in thread 0. Contains a barrier.
template <int NT> __device__ __forceinline__ long long block_sum_q(long long v) {
    static_assert(NT % 64 == 0, "whole waves");
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __shared__ long long s_wq[NT / 64];
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    if (tid % 64 == 0) s_wq[tid / 64] = v;
    __syncthreads();
    long long s = 0;
    if (tid == 0)
        for (int i = 0; i < NT / 64; ++i) s += s_wq[i];
    return s;
}

// A moment-form phase work item's fixed-point partial sum of phi (see
// MassQ), stored for launch_mass_finish.
template <int TX, int TY>
__device__ __forceinline__ void massq_partial(const CaseView& c, int x0, int y0, int z0, long long part) {
    const Geo& G = c.geo;
    const unsigned gx = (unsigned)(G.nx + TX - 1) / TX, gy = (unsigned)(G.ny + TY - 1) / TY;
    const unsigned tile = ((unsigned)(z0 / c.zc) * gy + (unsigned)(y0 / TY)) * gx + (unsigned)(x0 / TX);
    const long long s = block_sum_q<TX * TY>(part);
    if (threadIdx.y == 0 && threadIdx.x == 0) reinterpret_cast<long long*>(c.partial)[tile] = s;
}

// Moment form of the 16-bit storage forms (not bit-exact, so the algebra is
// free): collisionPhase's g*_l(y) = w_l phi (1 + 3 c_l.u) + w_l phi_n c_l.n is
// w_l (phi + c_l.B) with B = 3 phi u + phi_n n, so the state keeps phi and
// the vector B (in gm_u) and the phase pass sums
//   phi(x) = w0 phi(x) + w1 sum_faces (phi + c.B)(x - c) + w2 sum_edges (...).
// 4 fields per source cell instead of 7, and a handful of adds per link.
// The tile and its 1-cell halo are staged plane by plane into a 3-plane LDS
// ring (the next plane's loads in flight while a plane is summed). PAIRED:
// a lane stages two x-adjacent cells of an even-aligned pair (one 8-byte phi
// load and one 4-byte load per B component, the halo in one pass); it needs
// an even row length, so that every row, hence every pair, starts at an
// even cell. Otherwise a lane stages single cells. With a MassQ (16-bit
// storage) every phi gets the pending mass-drift correction and enters the
// fixed-point sum of the pass.
template <int TX, int TY, bool PAIRED>
__device__ __forceinline__ void phase_tile_ab_impl(const CaseView& c, int x0, int y0, int zb, int ze,
                                                   float (*s_a)[4][(TX + 2) * (TY + 2)]) {
    constexpr int NT = TX * TY;
    constexpr int HX = TX + 2, HY = TY + 2, HN = HX * HY;
    constexpr int KH = (HN + NT - 1) / NT;
    constexpr int PX = TX / 2 + 2, NPR = PX * HY;   // pairs per halo row, in the halo
    static_assert(TX % 2 == 0 && NPR <= NT, "a lane per pair");
    constexpr int KS = PAIRED ? 1 : KH;             // staging items per lane

    const Geo& G = c.geo;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
    const float w0 = c.phys.w[0], w1 = c.phys.w[1], w2 = c.phys.w[2];
    // Staging items: single halo cells e = tid + q NT, or the pair starting
    // at halo column 2 px - 1 of halo row hy (tid = hy PX + px).
    unsigned hoff[KS];
    int pe = 0;
#pragma unroll
    for (int q = 0; q < KS; ++q) {
        unsigned off = 0;
        if constexpr (PAIRED) {
            if (tid < NPR) {
                const int px = tid % PX, hy = tid / PX;
                off = (unsigned)(wrapmod(y0 - 1 + hy, G.ny) * G.nx + wrapmod(x0 - 2 + 2 * px, G.nx));
                pe = hy * HX + 2 * px - 1;
            }
        } else {
            const int e = tid + q * NT;
            if (e < HN) {
                const int hx = e % HX, hy = e / HX;
                off = (unsigned)(wrapmod(y0 - 1 + hy, G.ny) * G.nx + wrapmod(x0 - 1 + hx, G.nx));
            }
        }
        hoff[q] = off;
    }
    typedef float f32x2 __attribute__((ext_vector_type(2)));
    typedef const f32x2 __attribute__((address_space(1)))* gf2;
    typedef const unsigned __attribute__((address_space(1)))* gu32;
    typedef const unsigned short __attribute__((address_space(1)))* gu16;
    auto dec = [](unsigned short h) { return (float)__builtin_bit_cast(_Float16, h) * kBUnscale; };
    float reg[KS][4];
    unsigned rb[3];   // PAIRED: B components of the pair, two FP16 each
    f32x2 rp;         // PAIRED: phi of the pair
    auto fetch = [&](int z) {
        const unsigned base = (unsigned)(wrapmod(z, G.nz) * G.nxy);
        if constexpr (PAIRED) {
            if (tid < NPR) {
                const unsigned o = base + hoff[0];
                rp = *((gf2)c.phi_prev + (o >> 1));
#pragma unroll
                for (int k = 0; k < 3; ++k) rb[k] = *(gu32)((gu16)c.gm_u[k] + o);
            }
        } else {
#pragma unroll
            for (int q = 0; q < KS; ++q) {
                const int e = tid + q * NT;
                if (q < KS - 1 || e < HN) {
                    const unsigned o = base + hoff[q];
                    reg[q][0] = ld32(c.phi_prev, o);
                    reg[q][1] = ld_b(c.gm_u[0], o);
                    reg[q][2] = ld_b(c.gm_u[1], o);
                    reg[q][3] = ld_b(c.gm_u[2], o);
                }
            }
        }
    };
    auto put = [&](int slot) {
        float (*dst)[HN] = s_a[slot];
        if constexpr (PAIRED) {
            if (tid < NPR) {
                // Halo columns 2 px - 1 (from px = 1) and 2 px (up to px = TX / 2).
                const int px = tid % PX;
                if (px >= 1) {
                    dst[0][pe] = rp.x;
#pragma unroll
                    for (int k = 0; k < 3; ++k) dst[1 + k][pe] = dec((unsigned short)(rb[k] & 0xffffu));
                }
                if (px <= TX / 2) {
                    dst[0][pe + 1] = rp.y;
#pragma unroll
                    for (int k = 0; k < 3; ++k) dst[1 + k][pe + 1] = dec((unsigned short)(rb[k] >> 16));
                }
            }
        } else {
#pragma unroll
            for (int q = 0; q < KS; ++q) {
                const int e = tid + q * NT;
                if (q < KS - 1 || e < HN)
                    for (int k = 0; k < 4; ++k) dst[k][e] = reg[q][k];
            }
        }
    };
    const int x = x0 + tx, y = y0 + ty;
    const bool valid = x < G.nx && y < G.ny;
    const int e0 = (ty + 1) * HX + (tx + 1);
    const bool mq = c.massq != nullptr;   // uniform per case
    const float delta = mq ? (float)c.massq->delta[c.mslot] : 0.0f;
    long long qpart = 0;
    // Ring slots of planes z-1, z, z+1, rotated once per plane.
    int sm = 0, s0 = 1, sp = 2;
    fetch(zb - 1);
    put(sm);
    fetch(zb);
    put(s0);
    fetch(zb + 1);
    for (int z = zb; z < ze; ++z) {
        put(sp);   // plane z+1, fetched during the previous plane
        __syncthreads();
        if (z + 1 < ze) fetch(z + 2);
        if (valid) {
            float faces = 0.0f, edges = 0.0f;
            for_links([&](auto lc) {
                constexpr int l = decltype(lc)::value;
                if constexpr (l != 0) {
                    const float(*p)[HN] = s_a[CZ[l] > 0 ? sm : (CZ[l] < 0 ? sp : s0)];
                    const int e = e0 - CY[l] * HX - CX[l];
                    float v = p[0][e];
                    if constexpr (CX[l] != 0) v = CX[l] > 0 ? v + p[1][e] : v - p[1][e];
                    if constexpr (CY[l] != 0) v = CY[l] > 0 ? v + p[2][e] : v - p[2][e];
                    if constexpr (CZ[l] != 0) v = CZ[l] > 0 ? v + p[3][e] : v - p[3][e];
                    if constexpr (WC[l] == 1) faces += v;
                    else edges += v;
                }
            });
            const float own = s_a[s0][0][e0];
            float v = w0 * own + w1 * faces + w2 * edges;
            if (mq) {
                v += delta;
                qpart += (long long)__builtin_llrint((double)v * kMassQScale);
            }
            st32(c.phi, (unsigned)(z * G.nxy + y * G.nx + x), v);
        }
        __syncthreads();   // the slot of plane z-1 is rewritten next
        const int t = sm;
        sm = s0;
        s0 = sp;
        sp = t;
    }
    if (mq) massq_partial<TX, TY>(c, x0, y0, zb, qpart);
}

template <int TX, int TY>
__device__ __forceinline__ void phase_tile_ab(const CaseView& c, int x0, int y0, int zb, int ze) {
    __shared__ float s_a[3][4][(TX + 2) * (TY + 2)];   // ring of 3 planes: phi, Bx, By, Bz
    if ((c.geo.nx & 1) == 0) phase_tile_ab_impl<TX, TY, true>(c, x0, y0, zb, ze, s_a);
    else phase_tile_ab_impl<TX, TY, false>(c, x0, y0, zb, ze, s_a);
}

// ---------------------------------------------------------------- collide

// One cell of the collide pass for a 16-bit form (see lattice_fast.hpp), from
// its gathered f offsets fr (raw 16-bit): surface-tension force, moments, and
// the post-collision offsets of f,
// written against w_l, and of g, written against w_l phi of this cell (or, in
// moment form, this cell's u and n). Links 1..18 go first; their rounding
// residuals (in population units) then go into the rest link.
template <int MODE, class S, class NBN>
__device__ __forceinline__ void cell16(const CaseView& c, const fast::Consts& k, const PopPlanes<S>& fsrc,
                                       const PopPlanes<S>& gsrc, const PopPlanes<S>& fdst,
                                       const PopPlanes<S>& gdst, const Columns& col, unsigned me4,
                                       const unsigned (&zs)[3], const unsigned short (&fr)[Q], NBN&& nbn,
                                       float ncx, float ncy, float ncz, float ind, float phic) {
    float Fx, Fy, Fz;
    fast::curvature_force(nbn, k, ncx, ncy, ncz, ind, Fx, Fy, Fz);
    float tv[Q];   // f_l - w_l
    for_links([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        tv[l] = S::dec(fr[l]) * off_unscale(WC[l]);
    });
    const unsigned z1 = zs[1];
    const unsigned fb = (unsigned)(c.geo.n * 4);   // bytes of one FP32 field
    // Stores the 16-bit form of t = offset * K and adds its rounding residual
    // (population units) to r.
    auto put = [&](auto lc, const PopPlanes<S>& dst, float t, float& r) {
        constexpr int l = decltype(lc)::value;
        r += (t - dst.template stb<l>(col.me(), z1 * 2u, t)) * off_unscale(WC[l]);
    };
    float rf = 0.0f;
    auto emit = [&](auto lc, float x) {   // x = f*_l / w_l - 1
        constexpr int l = decltype(lc)::value;
        if constexpr (MODE != COLLIDE_MACROS) put(lc, fdst, x * (WC[l] == 1 ? k.wk1 : k.wk2), rf);
    };
    float rho, ux, uy, uz;
    const float x0 = fast::fluid_poly(tv, k, Fx, Fy, Fz, rho, ux, uy, uz, emit);
    if constexpr (MODE == COLLIDE_MACROS) {
        Field(c.rho, fb).st(me4, z1 * 4u, rho);
        Field(c.ux, fb).st(me4, z1 * 4u, ux);
        Field(c.uy, fb).st(me4, z1 * 4u, uy);
        Field(c.uz, fb).st(me4, z1 * 4u, uz);
        Field(c.fx, fb).st(me4, z1 * 4u, Fx);
        Field(c.fy, fb).st(me4, z1 * 4u, Fy);
        Field(c.fz, fb).st(me4, z1 * 4u, Fz);
        Field(c.nx, fb).st(me4, z1 * 4u, ncx);
        Field(c.ny, fb).st(me4, z1 * 4u, ncy);
        Field(c.nz, fb).st(me4, z1 * 4u, ncz);
    } else {
        fdst.template stb<0>(col.me(), z1 * 2u, (k.w0 * x0 + rf) * off_scale(0));

        const float phin = k.sharp_c * phic * (1.0f - phic);
        if (MODE == COLLIDE_STEP && c.gm != GM_POPULATIONS) {
            // Moment form of g (tau_g = 1): B = 3 phi u + phi_n n (phase_tile_ab).
            const float p3 = 3.0f * phic;
            const unsigned fb2 = (unsigned)(c.geo.n * 2);
            Field16(c.gm_u[0], fb2).st(col.me(), z1 * 2u, enc_b(p3 * ux + phin * ncx));
            Field16(c.gm_u[1], fb2).st(col.me(), z1 * 2u, enc_b(p3 * uy + phin * ncy));
            Field16(c.gm_u[2], fb2).st(col.me(), z1 * 2u, enc_b(p3 * uz + phin * ncz));
            return;
        }
        // g offset of link l after the phase collision, from its odd part.
        auto g_off = [&](auto lc, float odd) {
            constexpr int l = decltype(lc)::value;
            if constexpr (MODE == COLLIDE_STEP) {
                return odd;
            } else {
                // (1 - omega_g)(g_pre - w phi) + omega_g (+-odd), with
                // g_pre - w phi = w (anchor - phi) + t / K.
                const unsigned vb = col.template src<l>(), zb = zs[1 - CZ[l]];
                const float pre = __builtin_fmaf(k.template w<l>(), Field(c.phi_prev, fb).ld(vb * 2u, zb * 4u) - phic,
                                                 gsrc.template ldb<l>(vb, zb * 2u) * off_unscale(WC[l]));
                return k.one_m_omega_g * pre + k.omega_g * odd;
            }
        };
        float rg = 0.0f;
        for_links([&](auto lc) {
            constexpr int l = decltype(lc)::value;
            if constexpr (l % 2 == 1) {
                const float odd = fast::phase_odd<l>(k, phic, phin, ux, uy, uz, ncx, ncy, ncz);
                put(LinkC<l>{}, gdst, g_off(LinkC<l>{}, odd) * off_scale(WC[l]), rg);
                put(LinkC<l + 1>{}, gdst, g_off(LinkC<l + 1>{}, -odd) * off_scale(WC[l]), rg);
            }
        });
        gdst.template stb<0>(col.me(), z1 * 2u, (g_off(LinkC<0>{}, 0.0f) + rg) * off_scale(0));
    }
}

// gradCalc + curvatureCalc + momentiCalc + collisionFluid + collisionPhase
// for the TX x TY column (x0, y0) of case c over planes [zb, ze). phi planes
// (2-cell halo) live in a 4-slot LDS ring and the normals (1-cell halo) in a
// 3-slot ring, so the gradient is computed once per tile-plus-halo cell and
// each phi value is fetched about (1 + 4/zc) times. Must be called by all
// TX*TY threads of the workgroup (it contains barriers).
template <int TX, int TY, int MODE, class S>
__device__ __forceinline__ void collide_tile(const CaseView& c, int x0, int y0, int zb, int ze) {
    constexpr int NT = TX * TY;
    constexpr int PX = TX + 4, PY = TY + 4, PN = PX * PY;  // phi plane + 2-halo
    constexpr int HX = TX + 2, HY = TY + 2, HN = HX * HY;  // normal plane + 1-halo
    constexpr int KP = (PN + NT - 1) / NT;
    constexpr int KH = (HN + NT - 1) / NT;
    __shared__ float s_phi[4][PN];
    __shared__ float s_n[3][3][HN];
    __shared__ float s_ind[2][NT];

    const Geo& G = c.geo;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
    const Derived d(c.phys);
    const fast::Consts k16(c.phys, off_scale(1), off_scale(2));
    const PopPlanes<S> fsrc(c.fsrc, G), fdst(c.fdst, G);
    const PopPlanes<S> gsrc(c.gsrc, G), gdst(c.gdst, G);

    // In-plane byte offsets of the halo phi elements this thread stages.
    unsigned poff[KP];
#pragma unroll
    for (int q = 0; q < KP; ++q) {
        const int e = tid + q * NT;
        unsigned off = 0;
        if (e < PN) {
            const int lx = e % PX, ly = e / PX;
            off = (unsigned)(wrapmod(y0 - 2 + ly, G.ny) * G.nx + wrapmod(x0 - 2 + lx, G.nx)) * 4u;
        }
        poff[q] = off;
    }

    const Field phi_f(c.phi, (unsigned)(G.n * 4));
    auto load_phi = [&](int z) {
        const unsigned zb4 = (unsigned)(wrapmod(z, G.nz) * G.nxy) * 4u;
        float* dst = s_phi[z & 3];
#pragma unroll
        for (int q = 0; q < KP; ++q) {
            const int e = tid + q * NT;
            if (q < KP - 1 || e < PN) dst[e] = phi_f.ld(poff[q], zb4);
        }
    };
    // Software-pipelined phi staging inside the z loop: fetch_phi() loads a
    // plane into registers while the previous plane's cells are computed,
    // put_phi() writes it to its LDS slot at the top of the next iteration.
    float preg[KP];
    auto fetch_phi = [&](int z) {
        const unsigned zb4 = (unsigned)(wrapmod(z, G.nz) * G.nxy) * 4u;
#pragma unroll
        for (int q = 0; q < KP; ++q) {
            const int e = tid + q * NT;
            if (q < KP - 1 || e < PN) preg[q] = phi_f.ld(poff[q], zb4);
        }
    };
    auto put_phi = [&](int z) {
        float* dst = s_phi[z & 3];
#pragma unroll
        for (int q = 0; q < KP; ++q) {
            const int e = tid + q * NT;
            if (q < KP - 1 || e < PN) dst[e] = preg[q];
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
                const int cc = (hy + 1) * PX + (hx + 1);
                float nx, ny, nz, ind;
                auto nb = [&](auto lc) {
                    constexpr int l = decltype(lc)::value;
                    return pick<CZ[l]>(pm, p0, pp)[cc + CY[l] * PX + CX[l]];
                };
                if constexpr (S::kAnchored) {
                    fast::grad_normal(nb, k16, nx, ny, nz, ind);
                } else if (c.bc == MCB_BC_REFERENCE && ref_face(G, x0 - 1 + hx, y0 - 1 + hy, z)) {
                    // gradCalc updates interior cells only: a face keeps n = 0.
                    nx = ny = nz = ind = 0.0f;
                } else {
                    grad_normal(nb, d, nx, ny, nz, ind);
                }
                s_n[sl][0][e] = nx;
                s_n[sl][1][e] = ny;
                s_n[sl][2][e] = nz;
                if (hx >= 1 && hx <= TX && hy >= 1 && hy <= TY) ind_dst[(hy - 1) * TX + (hx - 1)] = ind;
            }
        }
    };

    const int x = x0 + tx, y = y0 + ty;
    const bool valid = x < G.nx && y < G.ny;
    // Unsigned 32-bit offsets: the gathers use a scalar base + 32-bit vector
    // offset instead of 64-bit per-lane addresses.
    const unsigned xs[3] = {valid ? (unsigned)shift<-1>(x, G.nx) : 0u, (unsigned)x,
                            valid ? (unsigned)shift<1>(x, G.nx) : 0u};
    const unsigned ys[3] = {valid ? (unsigned)(shift<-1>(y, G.ny) * G.nx) : 0u, (unsigned)(y * G.nx),
                            valid ? (unsigned)(shift<1>(y, G.ny) * G.nx) : 0u};
    const Columns col(xs, ys, fsrc);   // f and g share the layout
    const unsigned me4 = (ys[1] + xs[1]) * 4u;   // FP32 fields

    // Prologue: phi planes zb-2..zb+1, normals of planes zb-1 and zb.
    load_phi(zb - 2);
    load_phi(zb - 1);
    load_phi(zb);
    load_phi(zb + 1);
    fetch_phi(zb + 2);
    __syncthreads();
    comp_n(zb - 1);
    comp_n(zb);
    __syncthreads();

    // The 19 population gathers of plane z are issued before the plane's
    // barriers and normals, so they are in flight while those run (raw bits
    // in registers; 16-bit forms decode at use).
    using Raw = typename PopPlanes<S>::Raw;
    auto fetch_f = [&](Raw (&r)[Q], int z) {
        const unsigned zs[3] = {(unsigned)(shift<-1>(z, G.nz) * G.nxy), (unsigned)(z * G.nxy),
                                (unsigned)(shift<1>(z, G.nz) * G.nxy)};
        for_links([&](auto lc) {
            constexpr int l = decltype(lc)::value;
            r[l] = fsrc.template raw<l>(col.template src<l>(), fsrc.zb(zs[1 - CZ[l]]));
        });
    };
    // Plane z of the tile, given its gathered populations fr.
    auto plane = [&](int z, const Raw (&fr)[Q]) {
        // Slot (z+2)&3 held plane z-2, last read by comp_n(z-1) before the
        // previous barrier.
        put_phi(z + 2);
        __syncthreads();
        // n slot of z+1 held plane z-2 (read last in iteration z-1).
        comp_n(z + 1);
        __syncthreads();
        if (z + 1 < ze) fetch_phi(z + 3);
        if (!valid) return;
        if (c.bc == MCB_BC_REFERENCE && ref_face(G, x, y, z)) {
            // Faces never collide (interior guards of src/lbm.cu); their
            // macroscopic fields keep the values the reference never
            // updates: rho = 1 (initTensor), u, F and n = 0.
            if constexpr (MODE == COLLIDE_MACROS) {
                const unsigned fb = (unsigned)(G.n * 4), z4 = (unsigned)(z * G.nxy) * 4u;
                Field(c.rho, fb).st(me4, z4, 1.0f);
                float* const zero[] = {c.ux, c.uy, c.uz, c.fx, c.fy, c.fz, c.nx, c.ny, c.nz};
                for (float* q : zero) Field(q, fb).st(me4, z4, 0.0f);
            }
            return;
        }

        // Wave-uniform plane offsets (elements) of z-1, z, z+1.
        const unsigned zs[3] = {(unsigned)(shift<-1>(z, G.nz) * G.nxy), (unsigned)(z * G.nxy),
                                (unsigned)(shift<1>(z, G.nz) * G.nxy)};
        const unsigned z1 = zs[1];
        const unsigned fb = (unsigned)(G.n * 4);
        const int sm = nslot(z - 1), s0 = nslot(z), sp = nslot(z + 1);
        const int e0 = (ty + 1) * HX + (tx + 1);
        const float ncx = s_n[s0][0][e0], ncy = s_n[s0][1][e0], ncz = s_n[s0][2][e0];
        const float ind = s_ind[z & 1][tid];

        const float phic = s_phi[z & 3][(ty + 2) * PX + (tx + 2)];
        auto nbn = [&](auto lc, float& ax, float& ay, float& az) {
            constexpr int l = decltype(lc)::value;
            const int slot = pick<CZ[l]>(sm, s0, sp);
            const int e = e0 + CY[l] * HX + CX[l];
            if constexpr (CX[l] != 0) ax = s_n[slot][0][e];
            if constexpr (CY[l] != 0) ay = s_n[slot][1][e];
            if constexpr (CZ[l] != 0) az = s_n[slot][2][e];
        };
        if constexpr (S::kAnchored) {
            cell16<MODE>(c, k16, fsrc, gsrc, fdst, gdst, col, me4, zs, fr, nbn, ncx, ncy, ncz, ind, phic);
        } else {
            float curv, Fx, Fy, Fz;
            curvature_force(nbn, c.phys, d, ncx, ncy, ncz, ind, curv, Fx, Fy, Fz);
            const Moments m = moments(fr, c.phys, Fx, Fy, Fz);

            if constexpr (MODE == COLLIDE_MACROS) {
                Field(c.rho, fb).st(me4, z1 * 4u, m.rho);
                Field(c.ux, fb).st(me4, z1 * 4u, m.ux);
                Field(c.uy, fb).st(me4, z1 * 4u, m.uy);
                Field(c.uz, fb).st(me4, z1 * 4u, m.uz);
                Field(c.fx, fb).st(me4, z1 * 4u, Fx);
                Field(c.fy, fb).st(me4, z1 * 4u, Fy);
                Field(c.fz, fb).st(me4, z1 * 4u, Fz);
                Field(c.nx, fb).st(me4, z1 * 4u, ncx);
                Field(c.ny, fb).st(me4, z1 * 4u, ncy);
                Field(c.nz, fb).st(me4, z1 * 4u, ncz);
            } else {
                const FluidCell fc = fluid_cell(m, Fx, Fy, Fz);
                const PhaseCell pc = phase_cell(c.phys, phic, m.ux, m.uy, m.uz, ncx, ncy, ncz);
                // BGK phase collision (tau_g != 1): the 19 pre-collision g
                // (post-stream, gathered) are loaded together here, so their
                // latency overlaps the fluid collision instead of stalling each
                // g* store in turn.
                float gpre[Q];
                if constexpr (MODE == COLLIDE_STEP_BGK) {
                    for_links([&](auto lc) {
                        constexpr int l = decltype(lc)::value;
                        gpre[l] = gsrc.template ldb<l>(col.template src<l>(), gsrc.zb(zs[1 - CZ[l]]));
                    });
                }
                for_links([&](auto lc) {
                    constexpr int l = decltype(lc)::value;
                    fdst.template stb<l>(col.me(), fdst.zb(z1), fluid_post<l>(fc, c.phys, d));
                });
                if (MODE == COLLIDE_STEP && c.gm != GM_POPULATIONS) {
                    // Moment form: keep what collisionPhase needs (phic is the
                    // phi buffer of this step, the next step's phi_prev): u,
                    // and n unless the next phase pass recomputes it from phi.
                    Field(c.gm_u[0], fb).st(me4, z1 * 4u, m.ux);
                    Field(c.gm_u[1], fb).st(me4, z1 * 4u, m.uy);
                    Field(c.gm_u[2], fb).st(me4, z1 * 4u, m.uz);
                    if (c.gm_n[0]) {
                        Field(c.gm_n[0], fb).st(me4, z1 * 4u, ncx);
                        Field(c.gm_n[1], fb).st(me4, z1 * 4u, ncy);
                        Field(c.gm_n[2], fb).st(me4, z1 * 4u, ncz);
                    }
                    return;
                }
                for_links([&](auto lc) {
                    constexpr int l = decltype(lc)::value;
                    if constexpr (MODE == COLLIDE_STEP) {
                        gdst.template stb<l>(col.me(), gdst.zb(z1), phase_post<l>(pc, c.phys));
                    } else {
                        gdst.template stb<l>(col.me(), gdst.zb(z1), phase_post_bgk<l>(pc, c.phys, d, gpre[l]));
                    }
                });
            }
        }
    };

    for (int z = zb; z < ze; ++z) {
        Raw r[Q];
        fetch_f(r, z);
        plane(z, r);
    }
}

// ------------------------------------------------------- f in moment form

// f in moment form (the moment representation of f; no counterpart in the
// reference): the state keeps, per cell, the 13 values collisionFluid reads
// (lbm.cu:258-287) from its last collision instead of the 19 post-collision
// populations, one plane each: rho, ux, uy, uz, pxx, pyy, pzz, pxy, pxz, pyz
// (Pi^neq), fx, fy, fz. f*_l(y) is collisionFluid of those values, so the
// collide pass rebuilds f*_l(x - c_l) from the source cell's moments with
// collisionFluid's own operations: the same bits as stored populations.

// The 13 moment planes of one buffer behind one buffer resource.
struct FmPlanes {
    __amdgpu_buffer_rsrc_t r;
    unsigned plane_bytes;
    __device__ __forceinline__ FmPlanes(const void* base, unsigned plane_elems)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)(kFm * plane_elems * 4u),
                                              kBufferConfigWord)),
          plane_bytes(plane_elems * 4u) {}
    __device__ __forceinline__ float ld(int k, unsigned vb, unsigned sb) const {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vb, (unsigned)k * plane_bytes + sb, 0));
    }
    __device__ __forceinline__ void st(int k, unsigned vb, unsigned sb, float v) const {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, vb, (unsigned)k * plane_bytes + sb,
                                              MCB_STORE_AUX);
    }
};

// collisionFluid's per-cell values (fluid_cell) from the 13 stored moments.
__device__ __forceinline__ FluidCell fluid_cell_of(const float (&v)[kFm]) {
    Moments m;
    m.rho = v[0];
    m.ux = v[1];
    m.uy = v[2];
    m.uz = v[3];
    m.pxx = v[4];
    m.pyy = v[5];
    m.pzz = v[6];
    m.pxy = v[7];
    m.pxz = v[8];
    m.pyz = v[9];
    return fluid_cell(m, v[10], v[11], v[12]);
}

// Collide pass with f in moment form (FP32, g in moment form, periodic) for
// the TX x TY column (x0, y0) over planes [zb, ze). For destination plane z,
// every source cell of the tile and of the 1-cell ring around it writes to
// LDS (the link-major GmSlots layout of phase_tile_gm_rn) the values f*_l it
// sends into plane z, each computed once from its moments: in iteration z a
// tile cell builds its plane z+1 (collisionFluid's per-cell values, then
// fluid_post per link), writes its c_z = -1 links at once and holds its
// in-plane links (9, rest link first) for plane z+1 and its c_z = +1 links
// (5) for plane z+2; a ring cell does the same for the five links it sends
// into the tile. Cells past the lattice edge (partial tiles) act as their
// periodic image. Each destination then reads its 19 values and continues as
// collide_tile: normals and curvature from the LDS phi ring, momentiCalc, and
// the 13 moments of this collision (or, with COLLIDE_MACROS, the
// observables) stored aligned; with stored normals, n as well. Must be called
// by all TX * TY threads of the workgroup (it contains barriers).
template <int TX, int TY, int MODE>
__device__ __forceinline__ void collide_tile_fm(const CaseView& c, int x0, int y0, int zb, int ze) {
    constexpr int NT = TX * TY;
    constexpr int PX = TX + 4, PY = TY + 4, PN = PX * PY;  // phi plane + 2-halo
    constexpr int HX = TX + 2, HY = TY + 2, HN = HX * HY;  // normal plane + 1-halo
    constexpr int KP = (PN + NT - 1) / NT;
    constexpr int KH = (HN + NT - 1) / NT;
    constexpr int NR = gm_ring_cells<TX, TY>();
    using SL = GmSlots<TX, TY>;
    static_assert(NR <= NT, "a lane per ring cell");
    static_assert(gm_slots_ok<TX, TY>(), "link slots");
    static_assert(MODE == COLLIDE_STEP || MODE == COLLIDE_MACROS, "f in moment form: tau_g = 1");
    // phi in 3 slots (collide_tile keeps 4; this pass does not read plane z's
    // phi after the normals), so that with s_f two workgroups fit per CU.
    __shared__ float s_phi[3][PN];
    __shared__ float s_n[3][3][HN];
    __shared__ float s_ind[2][NT];
    __shared__ float s_f[SL::kSlots];

    const Geo& G = c.geo;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
    const Derived d(c.phys);
    const FmPlanes msrc(c.fsrc, (unsigned)G.plane), mdst(c.fdst, (unsigned)G.plane);
    auto plane4 = [&](int z) { return (unsigned)(wrapmod(z, G.nz) * G.nxy) * 4u; };

    // ---- phi ring and normals, as collide_tile (phi slot of plane z: (z + 3) mod 3, z >= -3)
    auto pslot = [](int z) { return (z + 3) % 3; };
    unsigned poff[KP];
#pragma unroll
    for (int q = 0; q < KP; ++q) {
        const int e = tid + q * NT;
        unsigned off = 0;
        if (e < PN) {
            const int lx = e % PX, ly = e / PX;
            off = (unsigned)(wrapmod(y0 - 2 + ly, G.ny) * G.nx + wrapmod(x0 - 2 + lx, G.nx)) * 4u;
        }
        poff[q] = off;
    }
    const Field phi_f(c.phi, (unsigned)(G.n * 4));
    auto load_phi = [&](int z) {
        const unsigned z4 = plane4(z);
        float* dst = s_phi[pslot(z)];
#pragma unroll
        for (int q = 0; q < KP; ++q) {
            const int e = tid + q * NT;
            if (q < KP - 1 || e < PN) dst[e] = phi_f.ld(poff[q], z4);
        }
    };
    float preg[KP];
    auto fetch_phi = [&](int z) {
        const unsigned z4 = plane4(z);
#pragma unroll
        for (int q = 0; q < KP; ++q) {
            const int e = tid + q * NT;
            if (q < KP - 1 || e < PN) preg[q] = phi_f.ld(poff[q], z4);
        }
    };
    auto put_phi = [&](int z) {
        float* dst = s_phi[pslot(z)];
#pragma unroll
        for (int q = 0; q < KP; ++q) {
            const int e = tid + q * NT;
            if (q < KP - 1 || e < PN) dst[e] = preg[q];
        }
    };
    auto nslot = [](int z) { return (z + 3) % 3; };  // z >= -3
    auto comp_n = [&](int z) {
        const float* pm = s_phi[pslot(z - 1)];
        const float* p0 = s_phi[pslot(z)];
        const float* pp = s_phi[pslot(z + 1)];
        const int sl = nslot(z);
        float* ind_dst = s_ind[z & 1];
#pragma unroll
        for (int q = 0; q < KH; ++q) {
            const int e = tid + q * NT;
            if (q < KH - 1 || e < HN) {
                const int hx = e % HX, hy = e / HX;
                const int cc = (hy + 1) * PX + (hx + 1);
                float nx, ny, nz, ind;
                auto nb = [&](auto lc) {
                    constexpr int l = decltype(lc)::value;
                    return pick<CZ[l]>(pm, p0, pp)[cc + CY[l] * PX + CX[l]];
                };
                grad_normal(nb, d, nx, ny, nz, ind);
                s_n[sl][0][e] = nx;
                s_n[sl][1][e] = ny;
                s_n[sl][2][e] = nz;
                if (hx >= 1 && hx <= TX && hy >= 1 && hy <= TY) ind_dst[(hy - 1) * TX + (hx - 1)] = ind;
            }
        }
    };

    // ---- tile cell
    const int x = x0 + tx, y = y0 + ty;
    const bool valid = x < G.nx && y < G.ny;
    const unsigned own4 = (unsigned)(wrapmod(y, G.ny) * G.nx + wrapmod(x, G.nx)) * 4u;
    const int row0 = ty * TX + tx, row1 = ty * (TX + 1) + tx;   // row bases of widths TX, TX + 1
    auto own_slot = [&](auto lc) -> float& {
        constexpr int l = decltype(lc)::value;
        return s_f[(SL::w(l) == TX ? row0 : row1) + SL::own(l)];
    };
    auto load_m = [&](float (&v)[kFm], unsigned cell4, unsigned z4) {
#pragma unroll
        for (int k = 0; k < kFm; ++k) v[k] = msrc.ld(k, cell4, z4);
    };
    float hin[9];      // in-plane links (rest first, then c_z = 0 links in link order) for the next plane
    float hup[2][5];   // c_z = +1 links by parity of their destination plane
    // From a source plane's per-cell values: its c_z = -1 links written for
    // the destination plane below, its in-plane links into hin and its c_z =
    // +1 links into up.
    auto tile_source = [&](const FluidCell& fc, bool write_down, float (&up)[5]) {
        int ki = 0, ku = 0;
        for_links([&](auto lc) {
            constexpr int l = decltype(lc)::value;
            const float v = fluid_post<l>(fc, c.phys, d);
            if constexpr (CZ[l] < 0) {
                if (write_down) own_slot(lc) = v;
            } else if constexpr (CZ[l] == 0) {
                hin[ki] = v;
                ++ki;
            } else {
                up[ku] = v;
                ++ku;
            }
        });
    };
    auto write_held = [&](const float (&up)[5]) {
        int ki = 1, ku = 0;
        for_links([&](auto lc) {
            constexpr int l = decltype(lc)::value;
            if constexpr (l != 0 && CZ[l] == 0) {
                own_slot(lc) = hin[ki];
                ++ki;
            } else if constexpr (CZ[l] > 0) {
                own_slot(lc) = up[ku];
                ++ku;
            }
        });
    };

    // ---- ring cell (side 0: row y0-1, 1: row y0+TY, 2: column x0-1, 3: column x0+TX)
    const bool ring = tid < NR;
    int hx, hy, side;
    if (tid < TX) { hx = tid + 1; hy = 0; side = 0; }
    else if (tid < 2 * TX) { hx = tid - TX + 1; hy = TY + 1; side = 1; }
    else if (tid < 2 * TX + TY + 2) { hx = 0; hy = tid - 2 * TX; side = 2; }
    else { hx = TX + 1; hy = min(tid - 2 * TX - TY - 2, TY + 1); side = 3; }
    const unsigned roff4 = (unsigned)(wrapmod(y0 - 1 + hy, G.ny) * G.nx + wrapmod(x0 - 1 + hx, G.nx)) * 4u;
    int rslot[5];
    {
        const int sx = hx - 1, sy = hy - 1;
        auto fill = [&](auto sc) {
            constexpr int s = decltype(sc)::value;
            auto one = [&](auto kc) {
                constexpr int k = decltype(kc)::value;
                constexpr int l = RING_LINK[s][k];
                rslot[k] = SL::inside(l, sx, sy) ? SL::at(l, sx, sy) : SL::kSpare;
            };
            one(std::integral_constant<int, 0>{});
            one(std::integral_constant<int, 1>{});
            one(std::integral_constant<int, 2>{});
            one(std::integral_constant<int, 3>{});
            one(std::integral_constant<int, 4>{});
        };
        if (side == 0) fill(std::integral_constant<int, 0>{});
        else if (side == 1) fill(std::integral_constant<int, 1>{});
        else if (side == 2) fill(std::integral_constant<int, 2>{});
        else fill(std::integral_constant<int, 3>{});
    }
    float rin[3];    // the ring cell's in-plane links for the next plane
    float rup[2];    // its c_z = +1 link by parity of the destination plane
    // The side's five links of a ring cell's plane: in-plane into rin, c_z =
    // +1 into up, c_z = -1 returned.
    auto ring_source = [&](const FluidCell& fc, float& up) {
        float down = 0.0f;
        auto by_side = [&](auto sc) {
            constexpr int s = decltype(sc)::value;
            rin[0] = fluid_post<RING_LINK[s][0]>(fc, c.phys, d);
            rin[1] = fluid_post<RING_LINK[s][1]>(fc, c.phys, d);
            rin[2] = fluid_post<RING_LINK[s][2]>(fc, c.phys, d);
            up = fluid_post<RING_LINK[s][3]>(fc, c.phys, d);
            down = fluid_post<RING_LINK[s][4]>(fc, c.phys, d);
        };
        if (side == 0) by_side(std::integral_constant<int, 0>{});
        else if (side == 1) by_side(std::integral_constant<int, 1>{});
        else if (side == 2) by_side(std::integral_constant<int, 2>{});
        else by_side(std::integral_constant<int, 3>{});
        return down;
    };

    // Prologue: normals of planes zb-1 and zb (phi planes zb-2..zb+1 through
    // the three slots); source planes zb-1 (c_z = +1 links for plane zb) and
    // zb (in-plane links for plane zb, c_z = +1 links for plane zb+1).
    load_phi(zb - 2);
    load_phi(zb - 1);
    load_phi(zb);
    {
        float v[kFm];
        load_m(v, own4, plane4(zb - 1));
        tile_source(fluid_cell_of(v), false, hup[0]);
        load_m(v, own4, plane4(zb));
        tile_source(fluid_cell_of(v), false, hup[1]);   // leaves plane zb's in-plane links in hin
        if (ring) {
            load_m(v, roff4, plane4(zb - 1));
            ring_source(fluid_cell_of(v), rup[0]);
            load_m(v, roff4, plane4(zb));
            ring_source(fluid_cell_of(v), rup[1]);      // leaves plane zb's in-plane links in rin
        }
    }
    __syncthreads();
    comp_n(zb - 1);
    __syncthreads();   // plane zb+1 replaces plane zb-2
    load_phi(zb + 1);
    fetch_phi(zb + 2);
    __syncthreads();
    comp_n(zb);
    __syncthreads();   // plane zb+2 replaces plane zb-1 in the first iteration

    auto iter = [&](int z, auto par) {
        constexpr int PAR = decltype(par)::value;
        float v[kFm], rv[kFm];
        load_m(v, own4, plane4(z + 1));                 // used after the plane's barrier and normals
        if (ring) load_m(rv, roff4, plane4(z + 1));
        // The phi slot of z+2 held plane z-1, last read by comp_n(z) before
        // the previous barrier; the n slot of z+1 held plane z-2 (read last
        // in iteration z-1).
        put_phi(z + 2);
        __syncthreads();
        comp_n(z + 1);
        if (z + 1 < ze) fetch_phi(z + 3);
        // Sources: plane z (held in-plane links), z-1 (held c_z = +1 links),
        // z+1 (built now).
        const float f0 = hin[0];
        write_held(hup[PAR]);
        tile_source(fluid_cell_of(v), true, hup[PAR]);
        if (ring) {
            s_f[rslot[0]] = rin[0];
            s_f[rslot[1]] = rin[1];
            s_f[rslot[2]] = rin[2];
            s_f[rslot[3]] = rup[PAR];
            s_f[rslot[4]] = ring_source(fluid_cell_of(rv), rup[PAR]);
        }
        __syncthreads();
        if (!valid) return;
        float fv[Q];
        fv[0] = f0;
        for_links([&](auto lc) {
            constexpr int l = decltype(lc)::value;
            if constexpr (l != 0) fv[l] = s_f[(SL::w(l) == TX ? row0 : row1) + SL::src(l)];
        });
        const unsigned fb = (unsigned)(G.n * 4), z4 = (unsigned)(z * G.nxy) * 4u;
        const int sm = nslot(z - 1), s0 = nslot(z), sp = nslot(z + 1);
        const int e0 = (ty + 1) * HX + (tx + 1);
        const float ncx = s_n[s0][0][e0], ncy = s_n[s0][1][e0], ncz = s_n[s0][2][e0];
        const float ind = s_ind[z & 1][tid];
        auto nbn = [&](auto lc, float& ax, float& ay, float& az) {
            constexpr int l = decltype(lc)::value;
            const int slot = pick<CZ[l]>(sm, s0, sp);
            const int e = e0 + CY[l] * HX + CX[l];
            if constexpr (CX[l] != 0) ax = s_n[slot][0][e];
            if constexpr (CY[l] != 0) ay = s_n[slot][1][e];
            if constexpr (CZ[l] != 0) az = s_n[slot][2][e];
        };
        float curv, Fx, Fy, Fz;
        curvature_force(nbn, c.phys, d, ncx, ncy, ncz, ind, curv, Fx, Fy, Fz);
        const Moments m = moments(fv, c.phys, Fx, Fy, Fz);
        if constexpr (MODE == COLLIDE_MACROS) {
            Field(c.rho, fb).st(own4, z4, m.rho);
            Field(c.ux, fb).st(own4, z4, m.ux);
            Field(c.uy, fb).st(own4, z4, m.uy);
            Field(c.uz, fb).st(own4, z4, m.uz);
            Field(c.fx, fb).st(own4, z4, Fx);
            Field(c.fy, fb).st(own4, z4, Fy);
            Field(c.fz, fb).st(own4, z4, Fz);
            Field(c.nx, fb).st(own4, z4, ncx);
            Field(c.ny, fb).st(own4, z4, ncy);
            Field(c.nz, fb).st(own4, z4, ncz);
        } else {
            const float out[kFm] = {m.rho, m.ux, m.uy, m.uz, m.pxx, m.pyy, m.pzz, m.pxy, m.pxz, m.pyz, Fx, Fy, Fz};
#pragma unroll
            for (int k = 0; k < kFm; ++k) mdst.st(k, own4, z4, out[k]);
            if (c.gm_n[0]) {
                Field(c.gm_n[0], fb).st(own4, z4, ncx);
                Field(c.gm_n[1], fb).st(own4, z4, ncy);
                Field(c.gm_n[2], fb).st(own4, z4, ncz);
            }
        }
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    for (int z = zb; z < ze; z += 2) {
        iter(z, I0{});
        if (z + 1 < ze) iter(z + 1, I1{});
    }
}

// ---------------------------------------------------------------- batched

// One work item per workgroup: a TX x TY x zc block of one case. The host
// builds the work table case by case, x fastest, then renumbers it with
// xcd_order so each XCD walks contiguous tiles.
// Sum of v over the NT threads of the workgroup in a fixed order (butterfly
// within each wave64, then the waves in order); the result is valid in
// thread 0. Contains a barrier.
template <int NT> __device__ __forceinline__ double block_sum(double v) {
    static_assert(NT % 64 == 0, "whole waves");
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __shared__ double s_w[NT / 64];
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    if (tid % 64 == 0) s_w[tid / 64] = v;
    __syncthreads();
    double s = 0.0;
    if (tid == 0)
        for (int i = 0; i < NT / 64; ++i) s += s_w[i];
    return s;
}

// A stored-g phase work item's partial sum of phi (see MassCell), stored
// for launch_mass_finish, which turns the partials into the next step's
// delta.
template <int TX, int TY>
__device__ __forceinline__ void mass_partial(const CaseView& c, const WorkItem& w, double part) {
    const Geo& G = c.geo;
    const unsigned gx = (unsigned)(G.nx + TX - 1) / TX, gy = (unsigned)(G.ny + TY - 1) / TY;
    const unsigned tile = ((unsigned)(w.z0 / c.zc) * gy + (unsigned)(w.y0 / TY)) * gx + (unsigned)(w.x0 / TX);
    const double s = block_sum<TX * TY>(part);
    if (threadIdx.y == 0 && threadIdx.x == 0) c.partial[tile] = s;
}

template <int TX, int TY, class S>
__global__ __launch_bounds__(TX * TY) void k_phase_batch(const CaseView* __restrict__ cases,
                                                         const WorkItem* __restrict__ work) {
    const WorkItem w = work[blockIdx.x];
    const CaseView c = cases[w.c];
    const Geo& G = c.geo;
    if (c.gm == GM_MOMENTS) {   // uniform per workgroup
        if constexpr (S::kAnchored) phase_tile_ab<TX, TY>(c, w.x0, w.y0, w.z0, min(w.z0 + c.zc, G.nz));
        else phase_tile_gm<TX, TY>(c, w.x0, w.y0, w.z0, min(w.z0 + c.zc, G.nz));
        return;
    }
    // 16-bit phase populations carry the mass bookkeeping (a barrier inside).
    const bool mass = S::kAnchored && c.gm == GM_POPULATIONS;
    const int x = w.x0 + (int)threadIdx.x, y = w.y0 + (int)threadIdx.y;
    const bool valid = x < G.nx && y < G.ny;
    if (!mass && !valid) return;
    const unsigned xs[3] = {valid ? (unsigned)shift<-1>(x, G.nx) : 0u, (unsigned)x,
                            valid ? (unsigned)shift<1>(x, G.nx) : 0u};
    const unsigned ys[3] = {valid ? (unsigned)(shift<-1>(y, G.ny) * G.nx) : 0u, (unsigned)(y * G.nx),
                            valid ? (unsigned)(shift<1>(y, G.ny) * G.nx) : 0u};
    const PopPlanes<S> gp(c.gsrc, G);
    const Columns col(xs, ys, gp);
    const int ze = min(w.z0 + c.zc, G.nz);
    double part = 0.0;
    float delta = 0.0f;
    if (mass) delta = (float)c.mass->delta[c.mslot];
    if (valid) {
        for (int z = w.z0; z < ze; ++z) {
            const unsigned zs[3] = {(unsigned)(shift<-1>(z, G.nz) * G.nxy), (unsigned)(z * G.nxy),
                                    (unsigned)(shift<1>(z, G.nz) * G.nxy)};
            float v = phase_sum<S>(c, gp, col, xs, ys, zs);
            if (mass) {
                v += delta;
                part += (double)v;
            }
            st32(c.phi, zs[1] + ys[1] + xs[1], v);
        }
    }
    if (mass) mass_partial<TX, TY>(c, w, part);
}

// 16-bit collide with more than 10 waves per workgroup (the 176 x 4 tile):
// at its natural 88 VGPRs one workgroup fits per CU, at 80 two (no scratch
// spills for that shape at tau_g = 1): 33.1k -> 37.6k MLUPS on 176^3 cases.
// Other shapes, and the tau_g != 1 form, keep their natural count (at 80
// VGPRs they spill; the other shapes ran 2-5% slower).
template <int TX, int TY, int MODE, class S> constexpr int collide_waves_per_eu() {
    return S::kAnchored && MODE != COLLIDE_STEP_BGK && TX * TY > 640 ? 6 : 1;
}

template <int TX, int TY, int MODE, class S>
__global__ __launch_bounds__(TX * TY)
__attribute__((amdgpu_waves_per_eu(collide_waves_per_eu<TX, TY, MODE, S>())))
void k_collide_batch(const CaseView* __restrict__ cases, const WorkItem* __restrict__ work) {
    const WorkItem w = work[blockIdx.x];
    const CaseView c = cases[w.c];
    collide_tile<TX, TY, MODE, S>(c, w.x0, w.y0, w.z0, min(w.z0 + c.zc, c.geo.nz));
}

}  // namespace mcb
