// common.h -- shared host/device definitions for the CUBE P3M HIP library.
//
// Geometry follows parameters.f90:57-82 and variables.f90:24-26 (nn = 1).
// Device arithmetic helpers reproduce the reference's rounding order; the
// library is compiled with -ffp-contract=off so no FMA is formed unless a
// kernel asks for one explicitly.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace cube {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define CUBE_HIP(call)                                                                    \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      throw ::cube::Error(2, std::string(#call) + ": " + hipGetErrorString(e_) + " at " + \
                                 __FILE__ + ":" + std::to_string(__LINE__));              \
  } while (0)

constexpr double kPi = 3.14159265358979323846;  // parameters.f90:94
constexpr int kRatioSF[8] = {0, 1, 2, 4, 6, 8, 12, 16};  // parameters.f90:26, 1-based

// Derived sizes of one image.
struct Geo {
  int64_t ng, nnt, nns, rcs, ncb, ngb, izipx, B;
  int64_t nc, nt, ntt, ngp, ngt, nte;
  int64_t sf[8], nfp[8], nfb[8], nft[8];
  int64_t ntiles, ncell_tile, ncells_ext, ns3, np_image, np_image_max, np_tile_max;

  void init(int64_t ng_, int64_t nnt_, int64_t nns_, int64_t rcs_, int64_t ncb_, int64_t ngb_, int64_t izipx_,
            float image_buffer, float tile_buffer) {
    ng = ng_; nnt = nnt_; nns = nns_; rcs = rcs_; ncb = ncb_; ngb = ngb_; izipx = izipx_;
    B = 8 * izipx;
    nc = ng / rcs; nt = nc / nnt; ntt = nt / nns; ngp = ng / nnt;
    ngt = ngp + 2 * ngb; nte = nt + 2 * ncb;
    for (int i = 1; i <= 7; ++i) {
      sf[i] = kRatioSF[i];
      nfp[i] = (ngp / nns) * sf[i];
      nfb[i] = rcs * sf[i];
      nft[i] = nfp[i] + 2 * nfb[i];
    }
    ntiles = nnt * nnt * nnt;
    ncell_tile = nte * nte * nte;
    ncells_ext = ntiles * ncell_tile;
    ns3 = ntiles * nns * nns * nns;
    np_image = ng * ng * ng;  // variables.f90:24 with np_nc = ratio_cs
    float r = (float)nte / (float)nt;
    // variables.f90:25-26 (evaluated in default real like the parameter statements)
    np_image_max = (int64_t)((float)np_image * (r * r * r) * image_buffer);
    np_tile_max = (int64_t)((float)(np_image / ntiles) * (r * r * r) * tile_buffer);
  }
};

// Flat index of rhoc(i,j,k,T) with Fortran cell indices i,j,k in 1-ncb..nt+ncb
// and T the 0-based linear tile index (itx fastest).
__host__ __device__ inline int64_t rc_index(int64_t i, int64_t j, int64_t k, int64_t T, int64_t nte, int64_t ncb) {
  return ((T * nte + (k - 1 + ncb)) * nte + (j - 1 + ncb)) * nte + (i - 1 + ncb);
}

// Compressed coordinate decode (parameters.f90:91-92, particle_mesh.f90:36):
// (int(xp+ishift,izipx)+rshift)*x_resolution == (u + 0.5) * 2^-B.
template <int IZ>
struct XP;
template <>
struct XP<1> {
  using T = int8_t;
  __device__ static double off(const int8_t* xp, int64_t ip, int d) {
    return ((double)(uint32_t)(uint8_t)xp[3 * ip + d] + 0.5) * (1.0 / 256.0);
  }
  __device__ static uint32_t raw(const int8_t* xp, int64_t ip, int d) { return (uint8_t)xp[3 * ip + d]; }
  static constexpr int B = 8;
};
template <>
struct XP<2> {
  using T = int16_t;
  __device__ static double off(const int16_t* xp, int64_t ip, int d) {
    return ((double)(uint32_t)(uint16_t)xp[3 * ip + d] + 0.5) * (1.0 / 65536.0);
  }
  __device__ static uint32_t raw(const int16_t* xp, int64_t ip, int d) { return (uint16_t)xp[3 * ip + d]; }
  static constexpr int B = 16;
};

// TSC weights (particle_mesh.f90:56-62) in the working precision R.
// FP32: l3 = real(4)(xpos - floor(xpos)); dx1 = (1-l3)**2/2; dx3 = l3**2/2;
// dx2 = 1-dx1-dx3.  Returns floor(xpos) (the Fortran idx(:,1)).
template <typename R>
__device__ inline int tsc_weights(double xpos, R w[3]) {
  double fl = floor(xpos);
  R l3 = (R)(xpos - fl);
  R a = (R)1 - l3;
  w[0] = (a * a) / (R)2;
  w[2] = (l3 * l3) / (R)2;
  w[1] = ((R)1 - w[0]) - w[2];
  return (int)fl;
}

// The same weights from l3 = xpos - floor(xpos) given directly (exact when
// xpos is a multiple of a power of two, as for power-of-two ratio_cs/gridsize).
template <typename R>
__device__ inline void tsc_weights_frac(R l3, R w[3]) {
  R a = (R)1 - l3;
  w[0] = (a * a) / (R)2;
  w[2] = (l3 * l3) / (R)2;
  w[1] = ((R)1 - w[0]) - w[2];
}

// FD weights [(alpha-1)/4, -alpha/2, alpha/2, (1-alpha)/4], alpha=4./3
// (parameters.f90:53-54), evaluated in R like the reference's real(4).
template <typename R>
struct FDWeights {
  R w0, w1, w2, w3;
  __host__ __device__ FDWeights() {
    R alpha = (R)4 / (R)3;
    w0 = (alpha - (R)1) / (R)4;
    w1 = -alpha / (R)2;
    w2 = alpha / (R)2;
    w3 = ((R)1 - alpha) / (R)4;
  }
};

// Kick update (particle_mesh.f90:205, kick.f90:797): vp = vp + dv*a_mid*dt/6/pi.
// The coefficient kc = a_mid*dt/(6 pi) is formed on the host in FP64 (the
// reference rounds a_mid*dt to real(4) first; here the kick and drift
// coefficients stay FP64), and the product and add run in FP64 before the
// store to the FP32 velocity.
__device__ inline float kick_update(float vp, float dv, double kc) { return (float)((double)vp + (double)dv * kc); }
__device__ inline double kick_update(double vp, double dv, double kc) { return vp + dv * kc; }

// F_ra_inline (kick.f90:360-374) with integer powers expanded as products.
// c must equal (1./35.)/apm**2 evaluated in R (f_ra_coef), passed in so that
// kernels compute it once.
template <typename R>
__host__ __device__ inline R f_ra_coef(R apm) {
  return ((R)1 / (R)35) / (apm * apm);
}
template <typename R>
__host__ __device__ inline R f_ra(R r, R apm, R c) {
  R ep = ((R)2 * r) / apm;
  if (apm == (R)0 || ep > (R)2) return (R)1 / (r * r);
  R e2 = ep * ep, e3 = e2 * ep, e4 = e2 * e2, e5 = e4 * ep, e6 = e3 * e3;
  if (ep > (R)1)
    return c * (((((((R)12 * ((R)1 / e2) - (R)224) + (R)896 * ep) - (R)840 * e2) + (R)224 * e3) + (R)70 * e4 -
                 (R)48 * e5) +
                (R)7 * e6);
  if (ep > (R)0) return c * (((((R)224 * ep - (R)224 * e3) + (R)70 * e4) + (R)48 * e5) - (R)21 * e6);
  return (R)0;
}

// Workgroups are dispatched round-robin over the 8 XCDs of the chip, each
// with its own L2.  Kernels whose neighbouring blocks share input (halos,
// neighbour cells) map the dispatch index i to a logical block so that each
// XCD runs one contiguous range of logical blocks: XCD x = i % 8 takes
// blocks [start(x), start(x) + count(x)).  A bijection on [0, n) for any n.
__device__ inline int64_t xcd_block(int64_t i, int64_t n) {
  const int64_t q = n >> 3, r = n & 7, x = i & 7, k = i >> 3;
  return x * q + (x < r ? x : r) + k;
}

// Order-preserving float max for non-negative values through integer atomics
// (deterministic regardless of launch order).  The atomic is issued only when
// the value exceeds what the location already holds: a running maximum
// saturates after a few blocks, so almost every block only reads, instead of
// serialising on one address.
__device__ inline void atomic_max_nonneg(float* addr, float v) {
  if (v > *(volatile float*)addr) atomicMax((int*)addr, __float_as_int(v));
}
__device__ inline void atomic_max_nonneg(double* addr, double v) {
  if (v > *(volatile double*)addr) atomicMax((unsigned long long*)addr, (unsigned long long)__double_as_longlong(v));
}
__device__ inline void atomic_max_int(int* addr, int v) {
  if (v > *(volatile int*)addr) atomicMax(addr, v);
}

// Integer sums spread over kSumSlots addresses (selected by block index) so
// that block-level atomics do not serialise on one address; the host adds
// the slots.
constexpr int kSumSlots = 64;

template <typename R>
struct Cplx;
template <>
struct Cplx<float> {
  using T = float2;
};
template <>
struct Cplx<double> {
  using T = double2;
};

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Zero device memory with a fill kernel (graph-capture safe, unlike
// hipMemsetAsync nodes).
void zero_async(void* p, size_t bytes, hipStream_t s);

class KTimer;

}  // namespace cube
