// Internal helpers shared by the libmcb translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdio>
#include <string>
#include "../../include/mcb.h"
#include "../../include/mcb_ops.h"
#include "lattice.hpp"

namespace mcb {

// Lattice geometry of one case: sides, cells, and the population plane stride.
struct Geo {
    int nx, ny, nz;
    int nxy;          // nx*ny
    long long n;      // nx*ny*nz
    long long plane;  // elements between population planes (>= n)
};

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
