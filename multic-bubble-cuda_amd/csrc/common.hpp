// Internal helpers shared by the libmcb translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdio>
#include <string>
#include "../../include/mcb.h"
#include "../../include/mcb_ops.h"
#include "lattice.hpp"

namespace mcb {

// Lattice geometry of one case: sides, cells, the population plane stride
// and the layout of its FP32 population buffers (16-bit ones are always in
// planes).
//
// Population layouts (cell index i = z*nxy + y*nx + x, link l):
//   planes (blocked = 0): element l*plane + i, one plane of cells per link;
//   blocks (blocked = 1): element (i >> 6)*19*64 + l*64 + (i & 63), groups of
//     64 consecutive cells holding their 19 links side by side.
// Blocks need nxy to be a multiple of 64, so a z plane is a whole number of
// groups and an offset splits into a wave-uniform z part (z*nxy*19 elements)
// and a per-lane in-plane part. A cell's 19 links then lie within 19*64
// elements instead of 19 planes apart, and a wave's gathers and stores open
// fewer DRAM pages (256^3: 0.577 -> 0.544 ms per FP32 collide).
struct Geo {
    int nx, ny, nz;
    int nxy;          // nx*ny
    long long n;      // nx*ny*nz
    long long plane;  // elements between population planes (>= n)
    int blocked = 0;  // FP32 populations in 64-cell blocks (see above)
};

#ifndef MCB_POP_BLOCKS
#define MCB_POP_BLOCKS 1   // 0: FP32 populations always in planes
#endif

// Element offset of link l of cell i in a population buffer of layout blk.
__host__ __device__ __forceinline__ long long pop_index(const Geo& G, int blk, int l, long long i) {
    return blk ? (i >> 6) * (19 * 64) + l * 64 + (i & 63) : l * G.plane + i;
}

__device__ __forceinline__ int wrap_lo(int a, int n) { return a < 0 ? a + n : a; }
__device__ __forceinline__ int wrap_hi(int a, int n) { return a >= n ? a - n : a; }
// a in [-n, 2n)
__device__ __forceinline__ int wrap1(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

// Coordinate of x + c for a constant c in {-1, 0, 1} with periodic wrap.
template <int C> __device__ __forceinline__ int shift(int a, int n) {
    if constexpr (C > 0) return wrap_hi(a + 1, n);
    else if constexpr (C < 0) return wrap_lo(a - 1, n);
    else return a;
}

// Element access at a 32-bit element offset from a wave-uniform base. The
// byte offset is formed in 32 bits, so the access compiles to a global
// load/store with the base in scalar registers and one 32-bit vector offset
// (no per-lane 64-bit address arithmetic). Needs elem * 4 < 2^32, i.e. planes
// of fewer than 2^30 cells (enforced by parameter validation). The pointer is
// taken as a global-memory (address space 1) pointer: fields reached through
// a CaseView read from device memory would otherwise be generic pointers and
// compile to flat accesses, which share the LDS counter and serialise the
// LDS staging against the gathers.
typedef const float __attribute__((address_space(1)))* gconst_fp;
typedef float __attribute__((address_space(1)))* g_fp;
__device__ __forceinline__ float ld32(const float* __restrict__ base, unsigned elem) {
    const gconst_fp p = (gconst_fp)base;
    return *(gconst_fp)((const char __attribute__((address_space(1)))*)p + (elem << 2));
}
__device__ __forceinline__ void st32(float* __restrict__ base, unsigned elem, float v) {
    const g_fp p = (g_fp)base;
    *(g_fp)((char __attribute__((address_space(1)))*)p + (elem << 2)) = v;
}
// The same with a byte offset that the caller keeps precomputed (no shift
// per access). Callers fold wave-uniform parts (the plane) into base.
__device__ __forceinline__ float ld32b(const float* __restrict__ base, unsigned byte_off) {
    return *(gconst_fp)((const char __attribute__((address_space(1)))*)(gconst_fp)base + byte_off);
}
__device__ __forceinline__ void st32b(float* __restrict__ base, unsigned byte_off, float v) {
    *(g_fp)((char __attribute__((address_space(1)))*)(g_fp)base + byte_off) = v;
}

// The 19 populations of one FP32 buffer (either layout of Geo) behind a
// single buffer resource: base address in four scalar registers, link l
// selected by the scalar offset l * lb, the cell by a 32-bit vector byte
// offset. This keeps one descriptor live instead of 19 64-bit plane
// pointers. Accesses beyond
// the resource's extent read 0 and drop stores instead of faulting. Needs
// 19 * plane * 4 < 2^32 (checked on the host before the fused path is used).
constexpr int kBufferConfigWord = 0x00020000;   // raw 32-bit buffer, gfx9 family
// Cache-policy bits of the population stores (gfx950: bit 1 = non-temporal).
// The post-collision planes are read back one full pass later, long after a
// 5 GB working set has cycled through L2 and the 256 MB Infinity Cache, so
// they are written non-temporally: measured 0.31 -> 0.28 ms for the following
// phase pass at 256^3, because fewer dirty lines are evicted during it.
#ifndef MCB_STORE_AUX
#define MCB_STORE_AUX 2
#endif
// The same bits for the 2-byte stores of the 16-bit forms.
#ifndef MCB_STORE16_AUX
#define MCB_STORE16_AUX 2
#endif
struct Planes {
    __amdgpu_buffer_rsrc_t r;
    unsigned lb;   // bytes between the links of a cell
    unsigned bb;   // bytes per group of 64 cells
    unsigned zm;   // bytes per unit of a z plane's cell offset z*nxy
    __device__ __forceinline__ Planes(const float* base, const Geo& G)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(base), (short)0, (int)(19u * (unsigned)G.plane * 4u),
                                              kBufferConfigWord)),
          lb(G.blocked ? 256u : (unsigned)G.plane * 4u),
          bb(G.blocked ? 19u * 256u : 256u),
          zm(G.blocked ? 76u : 4u) {}
    // Byte offset of cell index e (whole, or within its z plane) for link 0.
    __device__ __forceinline__ unsigned col(unsigned e) const { return (e >> 6) * bb + (e & 63u) * 4u; }
    // Byte offset of the z plane whose first cell index is zc (= z*nxy).
    __device__ __forceinline__ unsigned zb(unsigned zc) const { return zc * zm; }
    template <int L> __device__ __forceinline__ float ld(unsigned elem) const {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, col(elem), L * lb, 0));
    }
    template <int L> __device__ __forceinline__ void st(unsigned elem, float v) const {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, col(elem), L * lb, MCB_STORE_AUX);
    }
    // Split addressing: vb = per-lane in-plane byte offset (col of the
    // in-plane cell index, loop invariant), sb = wave-uniform byte offset of
    // the z plane (zb), which goes into the scalar offset next to the link.
    template <int L> __device__ __forceinline__ float ldb(unsigned vb, unsigned sb) const {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vb, L * lb + sb, 0));
    }
    template <int L> __device__ __forceinline__ void stb(unsigned vb, unsigned sb, float v) const {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, vb, L * lb + sb, MCB_STORE_AUX);
    }
};

// One FP32 field behind a buffer resource, addressed like Planes::ldb / stb:
// a per-lane byte offset that the caller keeps (loop-invariant) and a
// wave-uniform byte offset (the z plane), so no 64-bit per-lane address is
// formed per access. bytes: the field's extent (accesses beyond it read 0 and
// drop stores).
struct Field {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ Field(const float* base, unsigned bytes)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(base), (short)0, (int)bytes, kBufferConfigWord)) {}
    __device__ __forceinline__ float ld(unsigned vb, unsigned sb) const {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vb, sb, 0));
    }
    __device__ __forceinline__ void st(unsigned vb, unsigned sb, float v) const {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, vb, sb, 0);
    }
};

// Population plane stride (cells) of an n-cell lattice: n rounded up to 256
// cells, so every plane starts 1 KiB-aligned, plus one more 256-cell block.
// Without it the planes of a power-of-two lattice sit a power of two apart
// (2^25 bytes for 16-bit planes at 256^3), and the 19 concurrent plane
// streams of a collide pass contend for the same HBM channels: the skew
// measured 0.422 -> 0.407 ms per 16-bit step at 256^3 (FP32 unchanged).
#ifndef MCB_PLANE_SKEW
#define MCB_PLANE_SKEW 256
#endif
inline long long plane_stride_cells(long long n) { return (n + 255) / 256 * 256 + MCB_PLANE_SKEW; }

// Largest plane stride (cells) the fused kernels address through Planes.
constexpr long long kMaxFusedPlane = 0xFFFFFFFFLL / (19 * 4);

inline mcb_status hip_status(hipError_t e) {
    if (e == hipSuccess) return MCB_OK;
    if (e == hipErrorOutOfMemory) return MCB_ERR_OUT_OF_MEMORY;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return MCB_ERR_NO_DEVICE;
    return MCB_ERR_HIP;
}

inline Phys phys_from_consts(const mcb_op_consts& c) {
    Phys p;
    p.w[0] = c.w[0]; p.w[1] = c.w[1]; p.w[2] = c.w[2];
    p.omega = c.omega;
    p.sharp_c = c.sharp_c;
    p.sigma = c.sigma;
    p.omega_g = c.omega_g;
    return p;
}

}  // namespace mcb
