/*
 * cube_pm.h -- C-ABI of the differentiable PM-only force (particle-mesh
 * gravity on one periodic mesh) and its vector-Jacobian products.
 *
 * The reference has no counterpart: its closest pieces are PM1's mathematics
 * (particle_mesh.f90:36-205: TSC deposit, FFT convolution with Green_3D,
 * 4-point FD gradient, TSC interpolation) with apm_range = 0 (Green.f90:22-71)
 * and the commented single-PM case(0) (particle_mesh.f90:115-125).  These
 * entry points evaluate that operator on float particle positions, so that a
 * multi-step PM forward can be differentiated: each has a hand-written
 * adjoint (transpose of interpolation, of the FD stencil and of the
 * convolution, and the weight-derivative gathers).
 *
 * Conventions.  A batch of B independent periodic meshes of n^3 points,
 * [b][z][y][x] with x fastest.  Positions are float, in mesh units, any real
 * value (taps wrap periodically), x[3*p + d] for particle p of batch b =
 * p / np.  TSC: nearest mesh point i = floor(x + 1/2), d = x - i, weights
 * ((1/2 - d)^2/2, 3/4 - d^2, (1/2 + d)^2/2) on points i-1, i, i+1.
 *
 *   rho(m)   = sum_p W(x_p - m)                       (deposit)
 *   phi      = c2r(G * r2c(rho)) / n^3                (G = Green_3D(n, apm_soft, 0, 1))
 *   F_d(m)   = sum_s c_s phi(m + s e_d), c = (1/12, -2/3, 2/3, -1/12) at s = -2, -1, 1, 2
 *   acc_p,d  = sum_m W(x_p - m) F_d(m)                (interpolation)
 *
 * All pointers are DEVICE pointers on the context's device; work is enqueued
 * on the caller's HIP stream (`stream`, a hipStream_t; NULL = default
 * stream), and calls on one context must be ordered on one stream.  Sums
 * into meshes use integer (fixed-point) atomics, so every result is
 * independent of thread order.  Return codes are the CUBE_ERR_* of
 * cube_p3m.h.
 */
#ifndef CUBE_PM_H
#define CUBE_PM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cube_pm_config {
  int32_t n;        /* mesh points per dimension (a size the FFT supports)   */
  int32_t batch;    /* B independent meshes                                  */
  float apm_soft;   /* Green_3D softening in mesh cells (> 0)                */
  int32_t lattice;  /* 0: x holds absolute positions.  L > 0: x holds
                       displacements from a lattice of L^3 sites per mesh:
                       particle p of a batch (np = L^3) sits at site
                       p mod L^3, x fastest, at ((i + 1/2) n / L) per axis,
                       plus x.  Floor indices and fractions then carry the
                       rounding of the displacement, not of the absolute
                       coordinate (up to n cells), and gradients with respect
                       to x are those with respect to the positions.         */
} cube_pm_config;

typedef struct cube_pm_ctx cube_pm_ctx;

int cube_pm_create(const cube_pm_config* cfg, cube_pm_ctx** out);
int cube_pm_destroy(cube_pm_ctx* ctx);
const char* cube_pm_last_error(const cube_pm_ctx* ctx);

/* The kernel G on the half-complex grid, n*n*(n/2+1) floats, k_x fastest
 * (host pointer; synchronous). */
int cube_pm_green(cube_pm_ctx* ctx, float* G_host);

/* Force meshes F (B*3*n^3 floats, [b][d][z][y][x]) and accelerations
 * acc (B*np*3) at positions x (B*np*3).  F may be NULL (context scratch);
 * the VJP needs the F of the same positions. */
int cube_pm_force(cube_pm_ctx* ctx, const float* x, int64_t np, float* F, float* acc, void* stream);

/* g_x = (d acc / d x)^T g_acc, through both the interpolation weights and
 * the meshes (deposit -> convolution -> FD).  F is the force mesh of the
 * same x, or NULL: the force meshes of x are then recomputed into context
 * scratch first (callers that checkpoint positions only).  g_x is
 * overwritten. */
int cube_pm_force_vjp(cube_pm_ctx* ctx, const float* x, int64_t np, const float* F, const float* g_acc, float* g_x,
                      void* stream);

/* CIC density on B meshes of nw^3 points of the positions x*scale
 * (cell i covers [i, i+1), weights 1-d, d): rho (B*nw^3). */
int cube_pm_cic(cube_pm_ctx* ctx, const float* x, int64_t np, int32_t nw, float scale, float* rho, void* stream);

/* cube_pm_cic without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * fixed-point accumulator lives in a fixed device array of the library,
 * 2^23 64-bit cells (64 MiB) per GPU, held until the process ends; calls with
 * batch*nw^3 above 2^23 return CUBE_ERR_ARG.  Every call on one GPU shares
 * that array (each call clears it first), so calls that overlap on one GPU,
 * from any context or stream, race: order them.  Same result as cube_pm_cic,
 * bit for bit. */
int cube_pm_cic_noalloc(cube_pm_ctx* ctx, const float* x, int64_t np, int32_t nw, float scale, float* rho,
                        void* stream);

/* g_x = (d rho / d x)^T g_rho for cube_pm_cic.  g_x is overwritten. */
int cube_pm_cic_vjp(cube_pm_ctx* ctx, const float* x, int64_t np, int32_t nw, float scale, const float* g_rho,
                    float* g_x, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CUBE_PM_H */
