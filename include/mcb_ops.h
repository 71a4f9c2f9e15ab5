/*
 * libmcb operator face: one C entry point per reference kernel of
 * src/kernels.cuh (kreeyto/MULTIC-BUBBLE-CUDA), with the same names,
 * the same device-pointer arguments in the same order and meaning, and the
 * same SoA layout (src/header.cuh:20-21: x fastest; populations as 19
 * consecutive planes of NX*NY*NZ floats).
 *
 * Differences from the CUDA launches in src/main.cu:
 *  - the constants the reference reads from __constant__ memory (CSSQ, OMEGA,
 *    SHARP_C, SIGMA, W, CIX/CIY/CIZ; src/var.cu:6-11, 103-112) arrive in a
 *    mcb_op_consts passed by pointer, so there is no global state;
 *  - each call takes the HIP stream (hipStream_t, or NULL for the default
 *    stream) and returns an mcb_status instead of failing silently;
 *  - c->bc_mode = MCB_BC_REFERENCE gives the reference's interior guards;
 *    MCB_BC_PERIODIC updates every cell with periodic neighbours.
 * All pointers are device pointers. Calls are asynchronous on the stream.
 */
#ifndef MCB_OPS_H
#define MCB_OPS_H

#include "mcb.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MCB_OPS_MAX_BUBBLES 8

typedef struct mcb_op_consts {
    float w[3];          /* W[0], W[1..6], W[7..18] (src/var.cu:40-45) */
    float omega;         /* OMEGA = 1/tau (src/var.cu:23) */
    float sharp_c;       /* SHARP_C (src/var.cu:24) */
    float sigma;         /* SIGMA (src/var.cu:25) */
    float omega_g;       /* 1/tau_g (extension; unused when tau_g_is_one) */
    int tau_g_is_one;    /* 1: collisionPhase exactly as src/lbm.cu:324-326 */
    int bc_mode;         /* mcb_bc_mode */
    int n_bubbles;       /* initPhase bubbles (reference: one, R=20 at N/2) */
    float bubbles[4 * MCB_OPS_MAX_BUBBLES];
    float width;         /* initPhase tanh width (reference 3) */
} mcb_op_consts;

/* Constants derived from a parameter set exactly as src/var.cu:21-25. */
mcb_status mcb_op_consts_from_params(const mcb_params* p, mcb_op_consts* c);

/* src/kernels.cuh:6-15, src/init.cu:3-25 */
mcb_status mcb_initTensor(float* pxx, float* pyy, float* pzz, float* pxy, float* pxz, float* pyz,
                          float* rho, int NX, int NY, int NZ, const mcb_op_consts* c, void* stream);
/* src/kernels.cuh:17-20, src/init.cu:27-49 (device tanhf; the context API
 * initialises on the host instead so initial bits match the oracle) */
mcb_status mcb_initPhase(float* phi, int NX, int NY, int NZ, const mcb_op_consts* c, void* stream);
/* src/kernels.cuh:22-28, src/init.cu:51-80 */
mcb_status mcb_initDist(const float* rho, const float* phi, float* f, float* g,
                        int NX, int NY, int NZ, const mcb_op_consts* c, void* stream);
/* src/kernels.cuh:32-36, src/lbm.cu:5-26 */
mcb_status mcb_phiCalc(float* phi, const float* g, int NX, int NY, int NZ,
                       const mcb_op_consts* c, void* stream);
/* src/kernels.cuh:38-45, src/lbm.cu:34-72 (rsqrtf computed as 1/sqrtf) */
mcb_status mcb_gradCalc(const float* phi, float* normx, float* normy, float* normz, float* indicator,
                        int NX, int NY, int NZ, const mcb_op_consts* c, void* stream);
/* src/kernels.cuh:47-57, src/lbm.cu:80-125 */
mcb_status mcb_curvatureCalc(float* curvature, const float* indicator, const float* normx,
                             const float* normy, const float* normz, float* ffx, float* ffy, float* ffz,
                             int NX, int NY, int NZ, const mcb_op_consts* c, void* stream);
/* src/kernels.cuh:59-75, src/lbm.cu:133-225 */
mcb_status mcb_momentiCalc(float* ux, float* uy, float* uz, float* rho, const float* ffx, const float* ffy,
                           const float* ffz, const float* f, float* pxx, float* pyy, float* pzz,
                           float* pxy, float* pxz, float* pyz, int NX, int NY, int NZ,
                           const mcb_op_consts* c, void* stream);
/* src/kernels.cuh:77-93, src/lbm.cu:233-288 (push to x + c_l, in place) */
mcb_status mcb_collisionFluid(float* f, const float* ux, const float* uy, const float* uz,
                              const float* ffx, const float* ffy, const float* ffz, const float* rho,
                              const float* pxx, const float* pyy, const float* pzz,
                              const float* pxy, const float* pxz, const float* pyz,
                              int NX, int NY, int NZ, const mcb_op_consts* c, void* stream);
/* src/kernels.cuh:95-105, src/lbm.cu:290-329 (+ tau_g extension) */
mcb_status mcb_collisionPhase(float* g, const float* ux, const float* uy, const float* uz,
                              const float* phi, const float* normx, const float* normy, const float* normz,
                              int NX, int NY, int NZ, const mcb_op_consts* c, void* stream);
/* src/kernels.cuh:107-111, src/lbm.cu:337-357 (true modulo wrap) */
mcb_status mcb_streamingCalc(const float* g, float* g_out, int NX, int NY, int NZ,
                             const mcb_op_consts* c, void* stream);
/* src/kernels.cuh:113-119, src/lbm.cu:365-401 */
mcb_status mcb_fgBoundary(float* f, const float* rho, float* g, const float* phi,
                          int NX, int NY, int NZ, const mcb_op_consts* c, void* stream);
/* src/kernels.cuh:121-124, src/lbm.cu:403-425 (edge race resolved in the
 * serial order z-faces, y-faces, x-faces) */
mcb_status mcb_boundaryConditions(float* phi, int NX, int NY, int NZ, const mcb_op_consts* c, void* stream);
/* src/kernels.cuh:126-130, src/lbm.cu:430-461 (no observable effect) */
mcb_status mcb_movingWallBC(float* f, const float* rho, int NX, int NY, int NZ,
                            const mcb_op_consts* c, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* MCB_OPS_H */
