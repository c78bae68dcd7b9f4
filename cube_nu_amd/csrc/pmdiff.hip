// pmdiff.hip -- differentiable PM-only force on periodic meshes and its
// vector-Jacobian products (the C-ABI of include/cube_pm.h).
//
// Forward:  rho = TSC deposit, phi = c2r(G r2c(rho))/n^3, F_d = D_d phi (the
// 4-point stencil of particle_mesh.f90:237-242), acc = TSC interpolation of F.
// Adjoint of acc with respect to the positions, given g = dL/dacc:
//   Fbar_d   = TSC deposit of g_d                     (transpose of interpolation)
//   phibar   = sum_d D_d^T Fbar_d = -sum_d D_d Fbar_d (D_d is antisymmetric)
//   rhobar   = c2r(G r2c(phibar))/n^3                 (G real and even: self-adjoint)
//   dL/dx_p  = sum_m grad W(x_p - m) [F(m).g_p + rhobar(m)]
// Mesh sums use integer atomics on fixed-point values: 2^-32 units for the
// TSC density, 2^-31 for the CIC density, and for the signed adjoint
// deposits a power-of-two scale chosen from max|g| so that every
// contribution is a 32-bit integer and no sum can overflow.  Integer addition is
// associative, so every output is independent of thread order.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "../../include/cube_p3m.h"
#include "../../include/cube_pm.h"
#include "common.h"
#include "fft.h"
#include "pm.h"

// The allocation-free CIC entry point's device scratch (sized and explained
// with kScratchCells below).
static __device__ __align__(256) unsigned long long scratch[(int64_t)1 << 23];

namespace cube {
namespace {

template <typename T>
struct PMBuf {
  T* p = nullptr;
  size_t n = 0;
  T* ensure(size_t m) {
    if (m > n) {
      release();
      const hipError_t e = hipMalloc(&p, m * sizeof(T));
      if (e != hipSuccess) {
        p = nullptr;
        throw Error(CUBE_ERR_HIP, "hipMalloc of " + std::to_string(m * sizeof(T)) + " bytes failed");
      }
      n = m;
    }
    return p;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~PMBuf() { release(); }
};

constexpr double kFix = 4294967296.0;  // 2^32

// Fixed device scratch of cube_pm_cic_noalloc: the fixed-point CIC
// accumulator (batch * nw^3 64-bit cells, at most kScratchCells = 64 MiB).
// Each GPU the library runs on holds its own copy of the array, at its own
// address, until the process ends.  In host code `scratch` names a host-side
// stand-in, never the device memory: host code uses only the per-device
// address from hipGetSymbolAddress (scratch_address below).
constexpr int64_t kScratchCells = (int64_t)1 << 23;
constexpr int kMaxDevices = 64;

std::mutex g_scratch_mu;
unsigned long long* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device` (which must be current),
// looked up once per device.
unsigned long long* scratch_address(int device) {
  if (device < 0 || device >= kMaxDevices) throw Error(CUBE_ERR_ARG, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    static_assert(sizeof(::scratch) == sizeof(unsigned long long) * kScratchCells, "scratch size");
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(::scratch));
    if (e != hipSuccess || !ptr)
      throw Error(CUBE_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_ptr[device] = (unsigned long long*)ptr;
  }
  return g_scratch_ptr[device];
}

__device__ inline int wrapi(int i, int n) {
  i %= n;
  return i < 0 ? i + n : i;
}

// TSC weights and their derivatives on points i-1, i, i+1.
__device__ inline void tsc(float x, int& i, float w[3], float dw[3]) {
  const float fi = floorf(x + 0.5f);
  i = (int)fi;
  const float d = x - fi;
  const float a = 0.5f - d, b = 0.5f + d;
  w[0] = 0.5f * a * a;
  w[1] = 0.75f - d * d;
  w[2] = 0.5f * b * b;
  dw[0] = -a;
  dw[1] = -2.0f * d;
  dw[2] = b;
}

// CIC weights on points i, i+1 (cell i covers [i, i+1)).
__device__ inline void cic(float y, int& i, float w[2]) {
  const float fi = floorf(y);
  i = (int)fi;
  const float d = y - fi;
  w[0] = 1.0f - d;
  w[1] = d;
}

// Particle positions: absolute mesh coordinates (L = 0), or displacements
// from a Lagrangian lattice of L^3 sites per mesh (L > 0): particle p of a
// batch sits at site q = p mod L^3 (x fastest) whose position along each
// axis is (i + 1/2) n / L, plus its displacement.  With displacements the
// floor index and the fraction come from the integer and fractional parts of
// the site position (exact in FP64) plus the FP32 displacement, so their
// rounding is that of the displacement (|s| of a few cells), not of the
// absolute coordinate (up to n cells).
struct Lattice {
  int L = 0;          // sites per dimension; 0 = absolute positions
  double step = 0.0;  // n / L mesh cells between sites
};

// Site position along one axis times `scale`, split into integer and fraction.
__device__ inline void site_parts(const Lattice& lt, int64_t p, int d, double scale, int& ip, float& fp) {
  const int64_t q = p % ((int64_t)lt.L * lt.L * lt.L);
  const int64_t c = d == 0 ? q % lt.L : (d == 1 ? (q / lt.L) % lt.L : q / ((int64_t)lt.L * lt.L));
  const double v = ((double)c + 0.5) * lt.step * scale;
  const double f = floor(v);
  ip = (int)f;
  fp = (float)(v - f);
}

// TSC floor index and weights of coordinate d of particle p (p counts over
// the whole batch; the site is p mod L^3).
__device__ inline void tsc_pos(const float* __restrict__ x, int64_t p, int d, const Lattice& lt, int& i, float w[3],
                               float dw[3]) {
  if (lt.L == 0) {
    tsc(x[3 * p + d], i, w, dw);
    return;
  }
  int ip;
  float fp;
  site_parts(lt, p, d, 1.0, ip, fp);
  tsc(fp + x[3 * p + d], i, w, dw);  // the part below one cell, then the integer offset
  i += ip;
}

__device__ inline void cic_pos(const float* __restrict__ x, int64_t p, int d, const Lattice& lt, float scale, int& i,
                               float w[2]) {
  if (lt.L == 0) {
    cic(x[3 * p + d] * scale, i, w);
    return;
  }
  int ip;
  float fp;
  site_parts(lt, p, d, (double)scale, ip, fp);
  cic(fp + x[3 * p + d] * scale, i, w);
  i += ip;
}

__global__ void k_zero_u32(unsigned int* a, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = 0;
}

__global__ void k_zero_u64(unsigned long long* a, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = 0;
}

// rho += W (fixed point 2^-32); one thread per particle.
__global__ void k_tsc_deposit(const float* __restrict__ x, int64_t np, int64_t ntot, int n, Lattice lt,
                              unsigned long long* __restrict__ acc) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ntot) return;
  const int64_t b = p / np, n3 = (int64_t)n * n * n;
  int i[3];
  float w[3][3], dw[3][3];
  for (int d = 0; d < 3; ++d) tsc_pos(x, p, d, lt, i[d], w[d], dw[d]);
  unsigned long long* a = acc + b * n3;
  for (int k3 = 0; k3 < 3; ++k3) {
    const int z = wrapi(i[2] - 1 + k3, n);
    for (int k2 = 0; k2 < 3; ++k2) {
      const int y = wrapi(i[1] - 1 + k2, n);
      const float wyz32 = (w[1][k2] * w[2][k3]) * 4294967296.0f;  // weights <= 0.42: fits a u32
      for (int k1 = 0; k1 < 3; ++k1) {
        const int xx = wrapi(i[0] - 1 + k1, n);
        atomicAdd(&a[((int64_t)z * n + y) * n + xx], (unsigned long long)__float2uint_rn(w[0][k1] * wyz32));
      }
    }
  }
}

// out = acc * unit (as float), acc cleared for the next use.
__global__ void k_fix_to_float(unsigned long long* __restrict__ acc, int64_t n, double unit, float* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  out[t] = (float)((double)acc[t] * unit);
  acc[t] = 0;
}

// F[b][d] = D_d phi[b] (periodic).
__global__ void k_fd_force(const float* __restrict__ phi, int B, int n, float* __restrict__ F) {
  const int64_t n3 = (int64_t)n * n * n;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * n3) return;
  const int64_t b = t / n3, m = t % n3;
  const int x = (int)(m % n), y = (int)((m / n) % n), z = (int)(m / ((int64_t)n * n));
  const float* ph = phi + b * n3;
  const FDWeights<float> W;
  const int s[4] = {-2, -1, 1, 2};
  const float c[4] = {W.w0, W.w1, W.w2, W.w3};
  float f[3] = {0, 0, 0};
  for (int q = 0; q < 4; ++q) {
    f[0] = f[0] + ph[((int64_t)z * n + y) * n + wrapi(x + s[q], n)] * c[q];
    f[1] = f[1] + ph[((int64_t)z * n + wrapi(y + s[q], n)) * n + x] * c[q];
    f[2] = f[2] + ph[((int64_t)wrapi(z + s[q], n) * n + y) * n + x] * c[q];
  }
  for (int d = 0; d < 3; ++d) F[(b * 3 + d) * n3 + m] = f[d];
}

// phibar[b] = -sum_d D_d Fbar[b][d] (transpose of k_fd_force).
__global__ void k_fd_adjoint(const float* __restrict__ Fb, int B, int n, float* __restrict__ phib) {
  const int64_t n3 = (int64_t)n * n * n;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * n3) return;
  const int64_t b = t / n3, m = t % n3;
  const int x = (int)(m % n), y = (int)((m / n) % n), z = (int)(m / ((int64_t)n * n));
  const FDWeights<float> W;
  const int s[4] = {-2, -1, 1, 2};
  const float c[4] = {W.w0, W.w1, W.w2, W.w3};
  const float* f0 = Fb + (b * 3 + 0) * n3;
  const float* f1 = Fb + (b * 3 + 1) * n3;
  const float* f2 = Fb + (b * 3 + 2) * n3;
  float v = 0;
  for (int q = 0; q < 4; ++q) {
    v = v + f0[((int64_t)z * n + y) * n + wrapi(x + s[q], n)] * c[q];
    v = v + f1[((int64_t)z * n + wrapi(y + s[q], n)) * n + x] * c[q];
    v = v + f2[((int64_t)wrapi(z + s[q], n) * n + y) * n + x] * c[q];
  }
  phib[t] = -v;
}

// acc_p = sum W F.
__global__ void k_tsc_interp(const float* __restrict__ x, int64_t np, int64_t ntot, int n, Lattice lt,
                             const float* __restrict__ F, float* __restrict__ acc) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ntot) return;
  const int64_t b = p / np, n3 = (int64_t)n * n * n;
  int i[3];
  float w[3][3], dw[3][3];
  for (int d = 0; d < 3; ++d) tsc_pos(x, p, d, lt, i[d], w[d], dw[d]);
  const float* Fb = F + b * 3 * n3;
  float a[3] = {0, 0, 0};
  for (int k3 = 0; k3 < 3; ++k3) {
    const int z = wrapi(i[2] - 1 + k3, n);
    for (int k2 = 0; k2 < 3; ++k2) {
      const int y = wrapi(i[1] - 1 + k2, n);
      const float wyz = w[1][k2] * w[2][k3];
      for (int k1 = 0; k1 < 3; ++k1) {
        const int64_t m = ((int64_t)z * n + y) * n + wrapi(i[0] - 1 + k1, n);
        const float wt = w[0][k1] * wyz;
        for (int d = 0; d < 3; ++d) a[d] = fmaf(wt, Fb[d * n3 + m], a[d]);
      }
    }
  }
  for (int d = 0; d < 3; ++d) acc[3 * p + d] = a[d];
}

// gmax = max |g| (float bits; non-negative floats order as integers).
// Grid-stride over a fixed grid, then one atomic per block.
constexpr int kMaxBlocks = 1024;
__global__ void __launch_bounds__(256) k_absmax(const float* __restrict__ g, int64_t n, unsigned int* __restrict__ gmax) {
  __shared__ float red[4];
  float v = 0.0f;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    v = fmaxf(v, fabsf(g[t]));
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_down(v, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float m = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    if (m > 0) atomicMax(gmax, __float_as_uint(m));
  }
}

// Power-of-two scale S with max|g| * S in [2^30, 2^31): every contribution
// W g S (W <= 0.42) rounds to a 32-bit integer, and a mesh sum of at most
// 27 * ntot of them stays far inside 64 bits.  The rounding is 2^-31 of
// max|g| per contribution.
__device__ inline double fix_scale(const unsigned int* gmax) {
  const float m = __uint_as_float(*gmax);
  if (!(m > 0)) return 1.0;
  return ldexp(1.0, 30 - ilogbf(m));
}

// Fbar[b][d] += W g_d in fixed point of scale S.
__global__ void k_tsc_deposit_vec(const float* __restrict__ x, const float* __restrict__ g, int64_t np, int64_t ntot,
                                  int n, Lattice lt, const unsigned int* __restrict__ gmax,
                                  unsigned long long* __restrict__ acc) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ntot) return;
  const double S = fix_scale(gmax);
  const int64_t b = p / np, n3 = (int64_t)n * n * n;
  int i[3];
  float w[3][3], dw[3][3];
  for (int d = 0; d < 3; ++d) tsc_pos(x, p, d, lt, i[d], w[d], dw[d]);
  const float gs[3] = {(float)(g[3 * p] * S), (float)(g[3 * p + 1] * S), (float)(g[3 * p + 2] * S)};
  unsigned long long* a = acc + b * 3 * n3;
  for (int k3 = 0; k3 < 3; ++k3) {
    const int z = wrapi(i[2] - 1 + k3, n);
    for (int k2 = 0; k2 < 3; ++k2) {
      const int y = wrapi(i[1] - 1 + k2, n);
      const float wyz = w[1][k2] * w[2][k3];
      for (int k1 = 0; k1 < 3; ++k1) {
        const int64_t m = ((int64_t)z * n + y) * n + wrapi(i[0] - 1 + k1, n);
        const float wt = w[0][k1] * wyz;
        for (int d = 0; d < 3; ++d) {
          const int c = __float2int_rn(wt * gs[d]);
          if (c != 0) atomicAdd(&a[d * n3 + m], (unsigned long long)(long long)c);
        }
      }
    }
  }
}

// Tiled deposits.  The mesh is cut into kTile^3-cell tiles; a particle is
// listed under every tile its 3^3 TSC taps reach (one to eight), and one
// workgroup per tile sums the taps that fall inside it in LDS and writes the
// tile's cells once: one global atomic per listing instead of 27 (or 81) per
// particle.  The contributions are the integers the one-thread-per-particle
// form adds, so every sum, and every output, is the same bit for bit.
constexpr int kTile = 8;

__device__ inline void tile_span(int i, int n, int& t0, int& t1) {
  // tiles of the taps i-1 .. i+1 (periodic); t1 == t0 when they share one
  const int a = wrapi(i - 1, n) / kTile, b = wrapi(i + 1, n) / kTile;
  t0 = a;
  t1 = b;
}

template <int FILL>
__global__ void k_tile_bin(const float* __restrict__ x, int64_t np, int64_t ntot, int n, Lattice lt,
                           int* __restrict__ count, const int64_t* __restrict__ start, int* __restrict__ cursor,
                           int* __restrict__ ents) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ntot) return;
  const int64_t b = p / np;
  const int nt = n / kTile;
  int t0[3], t1[3];
  for (int d = 0; d < 3; ++d) {
    int i;
    float w[3], dw[3];
    tsc_pos(x, p, d, lt, i, w, dw);
    tile_span(i, n, t0[d], t1[d]);
  }
  for (int cz = 0; cz < 2; ++cz) {
    if (cz && t1[2] == t0[2]) break;
    const int tz = cz ? t1[2] : t0[2];
    for (int cy = 0; cy < 2; ++cy) {
      if (cy && t1[1] == t0[1]) break;
      const int ty = cy ? t1[1] : t0[1];
      for (int cx = 0; cx < 2; ++cx) {
        if (cx && t1[0] == t0[0]) break;
        const int tx = cx ? t1[0] : t0[0];
        const int64_t t = ((b * nt + tz) * nt + ty) * nt + tx;
        if (FILL) ents[start[t] + atomicAdd(&cursor[t], 1)] = (int)p;
        else atomicAdd(&count[t], 1);
      }
    }
  }
}

// One workgroup per tile: the listed particles' taps inside the tile, summed
// in LDS (u64; signed values as two's complement), then the tile's cells
// written as floats: NC = 1, the density (contribution u32(W 2^32), out =
// sum / 2^32); NC = 3, the adjoint's vector (contribution int(W g S), out =
// sum / S).
template <int NC>
__global__ void __launch_bounds__(256) k_tile_deposit(const float* __restrict__ x, const float* __restrict__ g,
                                                      int64_t np, int n, Lattice lt,
                                                      const int64_t* __restrict__ start,
                                                      const int* __restrict__ ents,
                                                      const unsigned int* __restrict__ gmax, float* __restrict__ out) {
  __shared__ unsigned long long acc[NC][kTile * kTile * kTile];
  const int nt = n / kTile;
  const int64_t t = blockIdx.x;
  const int tx = (int)(t % nt), ty = (int)((t / nt) % nt), tz = (int)((t / ((int64_t)nt * nt)) % nt);
  const int64_t b = t / ((int64_t)nt * nt * nt);
  for (int q = threadIdx.x; q < NC * kTile * kTile * kTile; q += blockDim.x) (&acc[0][0])[q] = 0;
  __syncthreads();
  const double S = NC == 3 ? fix_scale(gmax) : 1.0;
  const int org[3] = {tx * kTile, ty * kTile, tz * kTile};
  for (int64_t e = start[t] + threadIdx.x; e < start[t + 1]; e += blockDim.x) {
    const int64_t p = ents[e];
    int i[3];
    float w[3][3], dw[3][3];
    for (int d = 0; d < 3; ++d) tsc_pos(x, p, d, lt, i[d], w[d], dw[d]);
    float gs[3] = {0, 0, 0};
    if (NC == 3)
      for (int d = 0; d < 3; ++d) gs[d] = (float)(g[3 * p + d] * S);
    for (int k3 = 0; k3 < 3; ++k3) {
      const int lz = wrapi(i[2] - 1 + k3, n) - org[2];
      if (lz < 0 || lz >= kTile) continue;
      for (int k2 = 0; k2 < 3; ++k2) {
        const int ly = wrapi(i[1] - 1 + k2, n) - org[1];
        if (ly < 0 || ly >= kTile) continue;
        for (int k1 = 0; k1 < 3; ++k1) {
          const int lx = wrapi(i[0] - 1 + k1, n) - org[0];
          if (lx < 0 || lx >= kTile) continue;
          const int m = (lz * kTile + ly) * kTile + lx;
          if (NC == 1) {
            const float wyz32 = (w[1][k2] * w[2][k3]) * 4294967296.0f;
            atomicAdd(&acc[0][m], (unsigned long long)__float2uint_rn(w[0][k1] * wyz32));
          } else {
            const float wt = w[0][k1] * (w[1][k2] * w[2][k3]);
            for (int d = 0; d < NC; ++d) {
              const int c = __float2int_rn(wt * gs[d]);
              if (c != 0) atomicAdd(&acc[d][m], (unsigned long long)(long long)c);
            }
          }
        }
      }
    }
  }
  __syncthreads();
  const int64_t n3 = (int64_t)n * n * n;
  for (int q = threadIdx.x; q < kTile * kTile * kTile; q += blockDim.x) {
    const int lx = q % kTile, ly = (q / kTile) % kTile, lz = q / (kTile * kTile);
    const int64_t m = ((int64_t)(org[2] + lz) * n + (org[1] + ly)) * n + (org[0] + lx);
    for (int d = 0; d < NC; ++d) {
      const unsigned long long v = acc[d][q];
      out[(b * NC + d) * n3 + m] = NC == 1 ? (float)((double)v * (1.0 / kFix)) : (float)((double)(long long)v / S);
    }
  }
}

// out = (signed) acc / S, acc cleared.
__global__ void k_sfix_to_float(unsigned long long* __restrict__ acc, int64_t n, const unsigned int* __restrict__ gmax,
                                float* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double S = fix_scale(gmax);
  out[t] = (float)((double)(long long)acc[t] / S);
  acc[t] = 0;
}

// g_x = sum grad W (F.g + rhobar).
__global__ void k_tsc_vjp_gather(const float* __restrict__ x, int64_t np, int64_t ntot, int n, Lattice lt,
                                 const float* __restrict__ F, const float* __restrict__ g,
                                 const float* __restrict__ rhob, float* __restrict__ gx) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ntot) return;
  const int64_t b = p / np, n3 = (int64_t)n * n * n;
  int i[3];
  float w[3][3], dw[3][3];
  for (int d = 0; d < 3; ++d) tsc_pos(x, p, d, lt, i[d], w[d], dw[d]);
  const float gp[3] = {g[3 * p], g[3 * p + 1], g[3 * p + 2]};
  const float* Fb = F + b * 3 * n3;
  const float* rb = rhob + b * n3;
  float o[3] = {0, 0, 0};
  for (int k3 = 0; k3 < 3; ++k3) {
    const int z = wrapi(i[2] - 1 + k3, n);
    for (int k2 = 0; k2 < 3; ++k2) {
      const int y = wrapi(i[1] - 1 + k2, n);
      for (int k1 = 0; k1 < 3; ++k1) {
        const int64_t m = ((int64_t)z * n + y) * n + wrapi(i[0] - 1 + k1, n);
        const float s = fmaf(Fb[m], gp[0], fmaf(Fb[n3 + m], gp[1], fmaf(Fb[2 * n3 + m], gp[2], rb[m])));
        o[0] = fmaf(dw[0][k1] * w[1][k2] * w[2][k3], s, o[0]);
        o[1] = fmaf(w[0][k1] * dw[1][k2] * w[2][k3], s, o[1]);
        o[2] = fmaf(w[0][k1] * w[1][k2] * dw[2][k3], s, o[2]);
      }
    }
  }
  for (int d = 0; d < 3; ++d) gx[3 * p + d] = o[d];
}

// CIC density of x*scale on nw^3 (fixed point 2^-32).
__global__ void k_cic_deposit(const float* __restrict__ x, int64_t np, int64_t ntot, int nw, float scale, Lattice lt,
                              unsigned long long* __restrict__ acc) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ntot) return;
  const int64_t b = p / np, n3 = (int64_t)nw * nw * nw;
  int i[3];
  float w[3][2];
  for (int d = 0; d < 3; ++d) cic_pos(x, p, d, lt, scale, i[d], w[d]);
  unsigned long long* a = acc + b * n3;
  for (int k3 = 0; k3 < 2; ++k3)
    for (int k2 = 0; k2 < 2; ++k2)
      for (int k1 = 0; k1 < 2; ++k1) {
        const int64_t m = ((int64_t)wrapi(i[2] + k3, nw) * nw + wrapi(i[1] + k2, nw)) * nw + wrapi(i[0] + k1, nw);
        const float wt = w[0][k1] * w[1][k2] * w[2][k3];  // <= 1: fits a u32 at 2^31
        atomicAdd(&a[m], (unsigned long long)__float2uint_rn(wt * 2147483648.0f));
      }
}

__global__ void k_cic_vjp(const float* __restrict__ x, int64_t np, int64_t ntot, int nw, float scale, Lattice lt,
                          const float* __restrict__ grho, float* __restrict__ gx) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ntot) return;
  const int64_t b = p / np, n3 = (int64_t)nw * nw * nw;
  int i[3];
  float w[3][2];
  for (int d = 0; d < 3; ++d) cic_pos(x, p, d, lt, scale, i[d], w[d]);
  const float dw[2] = {-1.0f, 1.0f};
  const float* gr = grho + b * n3;
  float o[3] = {0, 0, 0};
  for (int k3 = 0; k3 < 2; ++k3)
    for (int k2 = 0; k2 < 2; ++k2)
      for (int k1 = 0; k1 < 2; ++k1) {
        const float v =
            gr[((int64_t)wrapi(i[2] + k3, nw) * nw + wrapi(i[1] + k2, nw)) * nw + wrapi(i[0] + k1, nw)];
        o[0] = fmaf(dw[k1] * w[1][k2] * w[2][k3], v, o[0]);
        o[1] = fmaf(w[0][k1] * dw[k2] * w[2][k3], v, o[1]);
        o[2] = fmaf(w[0][k1] * w[1][k2] * dw[k3], v, o[2]);
      }
  for (int d = 0; d < 3; ++d) gx[3 * p + d] = o[d] * scale;
}

inline unsigned nblk(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

}  // namespace
}  // namespace cube

using namespace cube;

struct cube_pm_ctx {
  int device = 0;
  cube_pm_config cfg{};
  int64_t n3 = 0;
  FFTCube<float> fc;
  PMBuf<float> G, mesh, Fs, Fbar;
  PMBuf<float2> work;
  PMBuf<unsigned long long> acc;   // B*3*n^3, kept zero between calls
  PMBuf<unsigned long long> accw;  // B*nw^3 for the CIC density, kept zero
  PMBuf<unsigned int> gmax;
  // tiled deposits: per-tile counts, list starts, fill cursors, particle lists
  PMBuf<int> tcount, tcursor, tents;
  PMBuf<int64_t> tstart;
  PMBuf<char> tscan;
  int nw_alloc = 0;
  Lattice lat;
  std::string err;

  bool tiled() const { return cfg.n % kTile == 0; }
  int64_t ntiles() const {
    const int64_t t = cfg.n / kTile;
    return (int64_t)cfg.batch * t * t * t;
  }
  // list every particle of x under the tiles its TSC taps reach
  void bin(const float* x, int64_t np, hipStream_t s) {
    const int n = cfg.n;
    const int64_t ntot = np * cfg.batch, nt = ntiles();
    tcount.ensure((size_t)nt);
    tcursor.ensure((size_t)nt);
    tstart.ensure((size_t)nt + 1);
    tents.ensure((size_t)ntot * 8);  // at most eight tiles per particle
    tscan.ensure(scan_tmp_bytes(nt));
    hipLaunchKernelGGL(k_zero_u32, dim3(nblk(nt)), dim3(256), 0, s, (unsigned int*)tcount.p, nt);
    hipLaunchKernelGGL(k_zero_u32, dim3(nblk(nt)), dim3(256), 0, s, (unsigned int*)tcursor.p, nt);
    hipLaunchKernelGGL((k_tile_bin<0>), dim3(nblk(ntot)), dim3(256), 0, s, x, np, ntot, n, lat, tcount.p, nullptr,
                       nullptr, nullptr);
    scan_counts(tcount.p, tstart.p, nt, tscan.p, tscan.n, s);
    hipLaunchKernelGGL((k_tile_bin<1>), dim3(nblk(ntot)), dim3(256), 0, s, x, np, ntot, n, lat, nullptr, tstart.p,
                       tcursor.p, tents.p);
    CUBE_HIP(hipGetLastError());
  }

  void check_np(int64_t np) const {
    if (np <= 0) throw Error(CUBE_ERR_ARG, "np must be positive");
    if (lat.L > 0 && np != (int64_t)lat.L * lat.L * lat.L)
      throw Error(CUBE_ERR_ARG, "with lattice displacements np must be lattice^3");
  }
  // rho -> phi -> F (B*3*n^3) at positions x
  void force_meshes(const float* x, int64_t np, float* F, hipStream_t s) {
    const int n = cfg.n, B = cfg.batch;
    const int64_t ntot = np * B;
    if (tiled()) {
      bin(x, np, s);
      hipLaunchKernelGGL((k_tile_deposit<1>), dim3((unsigned)ntiles()), dim3(256), 0, s, x, nullptr, np, n, lat,
                         tstart.p, tents.p, nullptr, mesh.p);
    } else {
      hipLaunchKernelGGL(k_tsc_deposit, dim3(nblk(ntot)), dim3(256), 0, s, x, np, ntot, n, lat, acc.p);
      hipLaunchKernelGGL(k_fix_to_float, dim3(nblk(B * n3)), dim3(256), 0, s, acc.p, B * n3, 1.0 / kFix, mesh.p);
    }
    fft_convolve<float>(fc, mesh.p, work.p, G.p, B, (float)n3, 0.0f, s);
    hipLaunchKernelGGL(k_fd_force, dim3(nblk(B * n3)), dim3(256), 0, s, mesh.p, B, n, F);
    CUBE_HIP(hipGetLastError());
  }
};

namespace {
template <class Fn>
int pm_guarded(cube_pm_ctx* ctx, Fn&& f) {
  if (!ctx) return CUBE_ERR_ARG;
  try {
    CUBE_HIP(hipSetDevice(ctx->device));
    f();
    ctx->err.clear();
    return CUBE_OK;
  } catch (const Error& e) {
    ctx->err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return CUBE_ERR_HIP;
  }
}
}  // namespace

extern "C" {

int cube_pm_create(const cube_pm_config* cfg, cube_pm_ctx** out) {
  if (!cfg || !out) return CUBE_ERR_ARG;
  *out = nullptr;
  if (cfg->n < 8 || cfg->batch < 1 || !(cfg->apm_soft > 0) || cfg->lattice < 0 || cfg->lattice > cfg->n)
    return CUBE_ERR_ARG;
  auto* c = new cube_pm_ctx();
  c->cfg = *cfg;
  c->lat.L = cfg->lattice;
  c->lat.step = cfg->lattice > 0 ? (double)cfg->n / (double)cfg->lattice : 0.0;
  const int rc = pm_guarded(c, [&] {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw Error(CUBE_ERR_NO_DEVICE, "no HIP device");
    CUBE_HIP(hipGetDevice(&c->device));
    (void)scratch_address(c->device);  // the allocation-free entry point's scratch, once per device
    const int n = cfg->n, B = cfg->batch;
    c->n3 = (int64_t)n * n * n;
    c->fc.init(n);
    const int64_t nk = (int64_t)n * n * (n / 2 + 1);
    c->G.ensure((size_t)nk);
    green_device(n, cfg->apm_soft, 0.0f, 1.0f, 1.0, c->G.p, nullptr);
    c->mesh.ensure((size_t)(B * c->n3));
    c->Fs.ensure((size_t)(3 * B * c->n3));
    c->Fbar.ensure((size_t)(3 * B * c->n3));
    c->work.ensure((size_t)(B * nk));
    c->acc.ensure((size_t)(3 * B * c->n3));
    c->gmax.ensure(1);
    hipLaunchKernelGGL(k_zero_u64, dim3(nblk(3 * B * c->n3)), dim3(256), 0, nullptr, c->acc.p, 3 * B * c->n3);
    CUBE_HIP(hipGetLastError());
    CUBE_HIP(hipDeviceSynchronize());
  });
  if (rc != CUBE_OK) {
    delete c;
    return rc;
  }
  *out = c;
  return CUBE_OK;
}

int cube_pm_destroy(cube_pm_ctx* ctx) {
  if (!ctx) return CUBE_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipDeviceSynchronize();
  ctx->fc.release();
  delete ctx;
  return CUBE_OK;
}

const char* cube_pm_last_error(const cube_pm_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int cube_pm_green(cube_pm_ctx* ctx, float* G_host) {
  return pm_guarded(ctx, [&] {
    if (!G_host) throw Error(CUBE_ERR_ARG, "null output");
    const int n = ctx->cfg.n;
    CUBE_HIP(hipMemcpy(G_host, ctx->G.p, sizeof(float) * (size_t)n * n * (n / 2 + 1), hipMemcpyDeviceToHost));
  });
}

int cube_pm_force(cube_pm_ctx* ctx, const float* x, int64_t np, float* F, float* acc, void* stream) {
  return pm_guarded(ctx, [&] {
    ctx->check_np(np);
    if (!x || !acc) throw Error(CUBE_ERR_ARG, "null pointer");
    const hipStream_t s = (hipStream_t)stream;
    float* Fm = F ? F : ctx->Fs.p;
    ctx->force_meshes(x, np, Fm, s);
    const int64_t ntot = np * ctx->cfg.batch;
    hipLaunchKernelGGL(k_tsc_interp, dim3(nblk(ntot)), dim3(256), 0, s, x, np, ntot, ctx->cfg.n, ctx->lat, Fm, acc);
    CUBE_HIP(hipGetLastError());
  });
}

int cube_pm_force_vjp(cube_pm_ctx* ctx, const float* x, int64_t np, const float* F, const float* g_acc, float* g_x,
                      void* stream) {
  return pm_guarded(ctx, [&] {
    ctx->check_np(np);
    if (!x || !g_acc || !g_x) throw Error(CUBE_ERR_ARG, "null pointer");
    const hipStream_t s = (hipStream_t)stream;
    const int n = ctx->cfg.n, B = ctx->cfg.batch;
    const int64_t ntot = np * B, n3 = ctx->n3;
    if (!F) {  // recompute the force meshes of x (checkpoint-and-recompute callers); bins x when tiled
      ctx->force_meshes(x, np, ctx->Fs.p, s);
      F = ctx->Fs.p;
    } else if (ctx->tiled()) {
      ctx->bin(x, np, s);
    }
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, ctx->gmax.p, 1);
    hipLaunchKernelGGL(k_absmax, dim3(std::min<int64_t>(nblk(3 * ntot), kMaxBlocks)), dim3(256), 0, s, g_acc,
                       3 * ntot, ctx->gmax.p);
    if (ctx->tiled()) {
      hipLaunchKernelGGL((k_tile_deposit<3>), dim3((unsigned)ctx->ntiles()), dim3(256), 0, s, x, g_acc, np, n,
                         ctx->lat, ctx->tstart.p, ctx->tents.p, ctx->gmax.p, ctx->Fbar.p);
    } else {
      hipLaunchKernelGGL(k_tsc_deposit_vec, dim3(nblk(ntot)), dim3(256), 0, s, x, g_acc, np, ntot, n, ctx->lat,
                         ctx->gmax.p, ctx->acc.p);
      hipLaunchKernelGGL(k_sfix_to_float, dim3(nblk(3 * B * n3)), dim3(256), 0, s, ctx->acc.p, 3 * B * n3,
                         ctx->gmax.p, ctx->Fbar.p);
    }
    hipLaunchKernelGGL(k_fd_adjoint, dim3(nblk(B * n3)), dim3(256), 0, s, ctx->Fbar.p, B, n, ctx->mesh.p);
    fft_convolve<float>(ctx->fc, ctx->mesh.p, ctx->work.p, ctx->G.p, B, (float)n3, 0.0f, s);
    hipLaunchKernelGGL(k_tsc_vjp_gather, dim3(nblk(ntot)), dim3(256), 0, s, x, np, ntot, n, ctx->lat, F, g_acc,
                       ctx->mesh.p, g_x);
    CUBE_HIP(hipGetLastError());
  });
}

int cube_pm_cic(cube_pm_ctx* ctx, const float* x, int64_t np, int32_t nw, float scale, float* rho, void* stream) {
  return pm_guarded(ctx, [&] {
    ctx->check_np(np);
    if (!x || !rho || nw < 2) throw Error(CUBE_ERR_ARG, "bad argument");
    const hipStream_t s = (hipStream_t)stream;
    const int B = ctx->cfg.batch;
    const int64_t ntot = np * B, nw3 = (int64_t)nw * nw * nw;
    if (ctx->nw_alloc != nw) {
      ctx->accw.release();
      ctx->accw.ensure((size_t)(B * nw3));
      hipLaunchKernelGGL(k_zero_u64, dim3(nblk(B * nw3)), dim3(256), 0, s, ctx->accw.p, B * nw3);
      ctx->nw_alloc = nw;
    }
    hipLaunchKernelGGL(k_cic_deposit, dim3(nblk(ntot)), dim3(256), 0, s, x, np, ntot, nw, scale, ctx->lat,
                       ctx->accw.p);
    hipLaunchKernelGGL(k_fix_to_float, dim3(nblk(B * nw3)), dim3(256), 0, s, ctx->accw.p, B * nw3, 1.0 / 2147483648.0,
                       rho);
    CUBE_HIP(hipGetLastError());
  });
}

int cube_pm_cic_noalloc(cube_pm_ctx* ctx, const float* x, int64_t np, int32_t nw, float scale, float* rho,
                        void* stream) {
  return pm_guarded(ctx, [&] {
    ctx->check_np(np);
    if (!x || !rho || nw < 2) throw Error(CUBE_ERR_ARG, "bad argument");
    const int B = ctx->cfg.batch;
    const int64_t nw3 = (int64_t)nw * nw * nw;
    if (nw > 4096 || B * nw3 > kScratchCells)
      throw Error(CUBE_ERR_ARG, "cube_pm_cic_noalloc: batch*nw^3 = " + std::to_string(B * nw3) +
                                    " cells exceeds the fixed scratch of " + std::to_string(kScratchCells) +
                                    " cells (64 MiB); use cube_pm_cic");
    unsigned long long* acc = scratch_address(ctx->device);
    const hipStream_t s = (hipStream_t)stream;
    const int64_t ntot = np * B;
    // the array keeps whatever the last call on this GPU left: clear it here
    hipLaunchKernelGGL(k_zero_u64, dim3(nblk(B * nw3)), dim3(256), 0, s, acc, B * nw3);
    hipLaunchKernelGGL(k_cic_deposit, dim3(nblk(ntot)), dim3(256), 0, s, x, np, ntot, nw, scale, ctx->lat, acc);
    hipLaunchKernelGGL(k_fix_to_float, dim3(nblk(B * nw3)), dim3(256), 0, s, acc, B * nw3, 1.0 / 2147483648.0, rho);
    CUBE_HIP(hipGetLastError());
  });
}

int cube_pm_cic_vjp(cube_pm_ctx* ctx, const float* x, int64_t np, int32_t nw, float scale, const float* g_rho,
                    float* g_x, void* stream) {
  return pm_guarded(ctx, [&] {
    ctx->check_np(np);
    if (!x || !g_rho || !g_x || nw < 2) throw Error(CUBE_ERR_ARG, "bad argument");
    const hipStream_t s = (hipStream_t)stream;
    const int64_t ntot = np * ctx->cfg.batch;
    hipLaunchKernelGGL(k_cic_vjp, dim3(nblk(ntot)), dim3(256), 0, s, x, np, ntot, nw, scale, ctx->lat, g_rho, g_x);
    CUBE_HIP(hipGetLastError());
  });
}

}  // extern "C"
