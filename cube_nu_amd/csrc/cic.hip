// cic.hip -- CIC density contrast and auto power spectrum of the physical
// particles (utilities/cicpower.f90:67-179; utilities/powerspectrum.f90:
// 140-259 auto_power with CIC deconvolution and pk_correction), nn = 1.
//
// Deposit: one wavefront per physical coarse cell, positions and weights in
// the reference's FP32 (cicpower.f90:78-88), sums in 2^-31 fixed point with
// 64-bit integer atomics on the periodic mesh (the reference's one-cell halo
// folded in by the index wrap), so the mesh is exact and independent of
// order.  Mode loop: one workgroup per chunk of half-spectrum modes, per-bin
// sums in LDS (FP64) flushed with FP64 atomics; per-mode quantities (k, |δ_k|²,
// C1k, Dk) in the reference's precision.  The alias correction's sum over
// images (n_int = 3, sinc^4) is a kernel over the reference's i >= j >= k
// octant.
#include <cmath>
#include <vector>

#include "cic.h"
#include "fft.h"

namespace cube {
namespace {

constexpr double kFix31 = 2147483648.0;  // 2^31: CIC weight products are <= 1

template <int IZ>
using XT = typename XP<IZ>::T;

template <int IZ>
__global__ void __launch_bounds__(256) k_cic_deposit(int nt, int nnt, int nte, int ncb, int nc, int nw,
                                                     const XT<IZ>* __restrict__ xp, const int32_t* __restrict__ rhoc,
                                                     const int64_t* __restrict__ off,
                                                     unsigned long long* __restrict__ acc) {
  const int64_t wv = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const int64_t nt3 = (int64_t)nt * nt * nt;
  if (wv >= nt3 * nnt * nnt * nnt) return;
  const int T = (int)(wv / nt3);
  const int q = (int)(wv - (int64_t)T * nt3);
  const int ijk[3] = {q % nt + 1, (q / nt) % nt + 1, q / (nt * nt) + 1};
  const int it[3] = {T % nnt, (T / nnt) % nnt, T / (nnt * nnt)};
  const int64_t cell = rc_index(ijk[0], ijk[1], ijk[2], T, nte, ncb);
  const int cnt = rhoc[cell];
  const int64_t base = off[cell];
  for (int l = lane; l < cnt; l += 64) {
    int i1[3], i2[3];
    float dx1[3], dx2[3];
    for (int d = 0; d < 3; ++d) {
      float pos = (float)((double)(nt * it[d] + (ijk[d] - 1)) + XP<IZ>::off(xp, base + l, d));
      pos = (pos * (float)nw) / (float)nc - 0.5f;  // cicpower.f90:79
      const int f = (int)floorf(pos) + 1;             // Fortran index into rho_grid(0:nw+1)
      dx1[d] = (float)f - pos;
      dx2[d] = 1.0f - dx1[d];
      // periodic cell of Fortran indices f and f+1 (the halo sums of cicpower.f90:108-156)
      i1[d] = f == 0 ? nw - 1 : f - 1;
      i2[d] = f == nw ? 0 : f;
    }
    auto add = [&](int a, int b, int c, float w) {
      atomicAdd(&acc[((int64_t)c * nw + b) * nw + a], (unsigned long long)__float2uint_rn((float)((double)w * kFix31)));
    };
    add(i1[0], i1[1], i1[2], (dx1[0] * dx1[1]) * dx1[2]);
    add(i2[0], i1[1], i1[2], (dx2[0] * dx1[1]) * dx1[2]);
    add(i1[0], i2[1], i1[2], (dx1[0] * dx2[1]) * dx1[2]);
    add(i1[0], i1[1], i2[2], (dx1[0] * dx1[1]) * dx2[2]);
    add(i1[0], i2[1], i2[2], (dx1[0] * dx2[1]) * dx2[2]);
    add(i2[0], i1[1], i2[2], (dx2[0] * dx1[1]) * dx2[2]);
    add(i2[0], i2[1], i1[2], (dx2[0] * dx2[1]) * dx1[2]);
    add(i2[0], i2[1], i2[2], (dx2[0] * dx2[1]) * dx2[2]);
  }
}

// Exact total of the fixed-point mesh into kSumSlots partial sums.
__global__ void k_cic_sum(const unsigned long long* __restrict__ acc, int64_t n, unsigned long long* __restrict__ slots) {
  __shared__ unsigned long long red[4];
  unsigned long long s = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    s += acc[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(&slots[blockIdx.x % kSumSlots], red[0] + red[1] + red[2] + red[3]);
}

// delta = rho / mean - 1 (cicpower.f90:176-179): rho as FP32, the division and
// the subtraction in FP64.
__global__ void k_cic_delta(const unsigned long long* __restrict__ acc, int64_t n, double mean,
                            float* __restrict__ delta) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float rho = (float)((double)acc[i] * (1.0 / kFix31));
    delta[i] = (float)((double)rho / mean - 1.0);
  }
}

constexpr int kModesPerBlock = 4096;

// auto_power mode loop (powerspectrum.f90:158-192) over the half spectrum
// ck[z][y][kx] (unnormalised FFT of delta).
__global__ void __launch_bounds__(256) k_cic_modes(const float2* __restrict__ ck, int nw, int nbin, int64_t npart,
                                                   double* __restrict__ bins) {
  extern __shared__ double lb[];  // [4][nbin+1]
  const int nh = nw / 2 + 1, nyq = nw / 2;
  const int nb = nbin + 1;
  for (int t = threadIdx.x; t < 4 * nb; t += blockDim.x) lb[t] = 0;
  __syncthreads();
  const int64_t nmode = (int64_t)nw * nw * nh;
  const int64_t m0 = (int64_t)blockIdx.x * kModesPerBlock;
  const float fnw = (float)nw;
  for (int r = threadIdx.x; r < kModesPerBlock; r += blockDim.x) {
    const int64_t mi = m0 + r;
    if (mi >= nmode) break;
    const int ig = (int)(mi % nh) + 1, jg = (int)((mi / nh) % nw) + 1, kg = (int)(mi / ((int64_t)nh * nw)) + 1;
    if (ig == 1 && jg == 1 && kg == 1) continue;  // zero frequency
    if ((ig == 1 || ig == nyq + 1) && jg > nyq + 1) continue;
    if ((ig == 1 || ig == nyq + 1) && (jg == 1 || jg == nyq + 1) && kg > nyq + 1) continue;
    const int kk[3] = {((ig + nyq - 1) % nw) - nyq, ((jg + nyq - 1) % nw) - nyq, ((kg + nyq - 1) % nw) - nyq};
    const float kx[3] = {(float)kk[0], (float)kk[1], (float)kk[2]};
    const float kr = sqrtf((kx[0] * kx[0] + kx[1] * kx[1]) + kx[2] * kx[2]);
    const int ib = (int)roundf(kr);  // nint: half away from zero
    if (ib > nbin) continue;
    const float2 c = ck[mi];
    const float re = ((c.x / fnw) / fnw) / fnw, im = ((c.y / fnw) / fnw) / fnw;
    const float amp = re * re + im * im;
    float c1[3];
    for (int d = 0; d < 3; ++d) {
      const double sn = sin(kPi * (double)kx[d] / (double)nw);
      c1[d] = (float)(1.0 - (double)(2.0f / 3.0f) * (sn * sn));  // CIC (n_interp = 2)
    }
    const float Dk = ((c1[0] * c1[1]) * c1[2]) / (float)npart;
    atomicAdd(&lb[ib], 1.0);
    atomicAdd(&lb[nb + ib], (double)kr);
    atomicAdd(&lb[2 * nb + ib], (double)amp);
    atomicAdd(&lb[3 * nb + ib], (double)(amp - Dk));
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
    if (lb[b] != 0)
      for (int c = 0; c < 4; ++c) atomicAdd(&bins[c * nb + b], lb[c * nb + b]);
}

// pk_correction integral (powerspectrum.f90:226-247) over i >= j >= k, in FP32
// like the reference: per bin, the count and the sum of W(k)/P(k) with
// W(k) = sum over images n in [-3,3]^3 of sinc(pi (k + n nw)/nw)^4 |k+n nw|^alpha
// and P(k) = |k|^alpha.
__global__ void __launch_bounds__(256) k_cic_alias(int nw, int nbin, float alpha, double* __restrict__ cb) {
  extern __shared__ double lc[];  // [2][nbin+1]
  const int nb = nbin + 1, nyq = nw / 2;
  for (int t = threadIdx.x; t < 2 * nb; t += blockDim.x) lc[t] = 0;
  __syncthreads();
  const int n1 = nyq + 1;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < (int64_t)n1 * n1 * n1) {
    const int i = (int)(p % n1) + 1, j = (int)((p / n1) % n1) + 1, k = (int)(p / ((int64_t)n1 * n1)) + 1;
    if (j <= i && k <= j) {
      const float kv[3] = {(float)(i - 1), (float)(j - 1), (float)(k - 1)};
      const float kmag = sqrtf((kv[0] * kv[0] + kv[1] * kv[1]) + kv[2] * kv[2]);
      const int ib = (int)roundf(kmag);
      if (ib >= 1 && ib <= nbin) {
        float s4[3][7];
        for (int d = 0; d < 3; ++d)
          for (int n = -3; n <= 3; ++n) {
            const float ks = (float)kPi * (kv[d] + (float)(n * nw)) / (float)nw;
            const float sc = ks == 0.0f ? 1.0f : sinf(ks) / ks;
            const float s2 = sc * sc;
            s4[d][n + 3] = s2 * s2;
          }
        float w = 0;
        for (int in = -3; in <= 3; ++in)
          for (int jn = -3; jn <= 3; ++jn)
            for (int kn = -3; kn <= 3; ++kn) {
              const float a = kv[0] + (float)(in * nw), b = kv[1] + (float)(jn * nw), c = kv[2] + (float)(kn * nw);
              const float kmn = sqrtf((a * a + b * b) + c * c);
              w += ((s4[0][in + 3] * s4[1][jn + 3]) * s4[2][kn + 3]) * powf(kmn, alpha);
            }
        atomicAdd(&lc[ib], 1.0);
        atomicAdd(&lc[nb + ib], (double)(w / powf(kmag, alpha)));
      }
    }
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
    if (lc[b] != 0) {
      atomicAdd(&cb[b], lc[b]);
      atomicAdd(&cb[nb + b], lc[nb + b]);
    }
}

template <typename T>
struct DevMem {
  T* p = nullptr;
  explicit DevMem(size_t n) { CUBE_HIP(hipMalloc(&p, n * sizeof(T))); }
  ~DevMem() {
    if (p) (void)hipFree(p);
  }
};

// interp1 (powerspectrum.f90:261-275) on xdata(1:nbin) = bins 0..nbin-1.
double interp1(const std::vector<double>& x, const std::vector<double>& y, int nbin, double xq) {
  int i1 = 1, i2 = nbin;
  while (i2 - i1 > 1) {
    const int im = (i1 + i2) / 2;
    if (xq > x[im - 1]) i1 = im;
    else i2 = im;
  }
  return y[i1 - 1] + (xq - x[i1 - 1]) / (x[i2 - 1] - x[i1 - 1]) * (y[i2 - 1] - y[i1 - 1]);
}

}  // namespace

void cic_power(const Geo& g, int nw, float box, const void* xp, const int32_t* rhoc, const int64_t* off, int64_t npart,
               float* xi, int nbin, float* delta, double* mean_out, hipStream_t s) {
  if (nw < 8 || nw % 2 != 0) throw Error(1, "cic power: nw must be even and >= 8");
  if (nbin != (int)std::lround((double)(nw / 2) * (double)std::sqrt(3.0f)))
    throw Error(1, "cic power: nbin must be nint(nw/2*sqrt(3.)) (powerspectrum.f90:14)");
  if (4 * (nbin + 1) * sizeof(double) > 64 * 1024) throw Error(1, "cic power: nw too large for the LDS bin table");
  const int64_t n3 = (int64_t)nw * nw * nw;
  const int64_t nmode = (int64_t)nw * nw * (nw / 2 + 1);
  DevMem<unsigned long long> acc((size_t)std::max<int64_t>(n3, nmode));  // reused for the spectrum
  DevMem<float> dl((size_t)n3);
  DevMem<unsigned long long> slots(kSumSlots);
  DevMem<double> bins((size_t)4 * (nbin + 1)), cb((size_t)2 * (nbin + 1));
  zero_async(acc.p, sizeof(unsigned long long) * n3, s);
  zero_async(slots.p, sizeof(unsigned long long) * kSumSlots, s);
  const int64_t ncell = (int64_t)g.nt * g.nt * g.nt * g.ntiles;
  const unsigned gd = (unsigned)ceil_div(ncell * 64, 256);
  if (g.izipx == 2)
    hipLaunchKernelGGL(k_cic_deposit<2>, dim3(gd), dim3(256), 0, s, (int)g.nt, (int)g.nnt, (int)g.nte, (int)g.ncb,
                       (int)g.nc, nw, (const int16_t*)xp, rhoc, off, acc.p);
  else
    hipLaunchKernelGGL(k_cic_deposit<1>, dim3(gd), dim3(256), 0, s, (int)g.nt, (int)g.nnt, (int)g.nte, (int)g.ncb,
                       (int)g.nc, nw, (const int8_t*)xp, rhoc, off, acc.p);
  CUBE_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_cic_sum, dim3(1024), dim3(256), 0, s, acc.p, n3, slots.p);
  CUBE_HIP(hipGetLastError());
  std::vector<unsigned long long> sl(kSumSlots);
  CUBE_HIP(hipMemcpyAsync(sl.data(), slots.p, sizeof(unsigned long long) * kSumSlots, hipMemcpyDeviceToHost, s));
  CUBE_HIP(hipStreamSynchronize(s));
  unsigned long long tot = 0;
  for (auto v : sl) tot += v;
  const double mean = ((double)tot / kFix31) / (double)nw / (double)nw / (double)nw;
  if (mean_out) *mean_out = mean;
  hipLaunchKernelGGL(k_cic_delta, dim3(2048), dim3(256), 0, s, acc.p, n3, mean, dl.p);
  CUBE_HIP(hipGetLastError());
  if (delta) CUBE_HIP(hipMemcpyAsync(delta, dl.p, sizeof(float) * n3, hipMemcpyDeviceToHost, s));
  // auto_power
  FFTCube<float> fc;
  fc.init(nw);
  float2* ck = (float2*)acc.p;
  fft_r2c3d<float>(fc, dl.p, ck, 1, s);
  zero_async(bins.p, sizeof(double) * 4 * (nbin + 1), s);
  const size_t lds = sizeof(double) * 4 * (nbin + 1);
  hipLaunchKernelGGL(k_cic_modes, dim3((unsigned)ceil_div(nmode, kModesPerBlock)), dim3(256), lds, s, ck, nw, nbin,
                     npart, bins.p);
  CUBE_HIP(hipGetLastError());
  const int nb = nbin + 1;
  std::vector<double> hb(4 * nb);
  CUBE_HIP(hipMemcpyAsync(hb.data(), bins.p, sizeof(double) * 4 * nb, hipMemcpyDeviceToHost, s));
  CUBE_HIP(hipStreamSynchronize(s));
  fc.release();
  std::vector<double> x1(nb), x2(nb), x3(nb), x4(nb), x5(nb);
  for (int b = 0; b < nb; ++b) {
    x1[b] = hb[b];
    x2[b] = hb[nb + b] / x1[b];
    x3[b] = hb[2 * nb + b] / x1[b];
    x4[b] = hb[3 * nb + b] / x1[b];
    x5[b] = x4[b];
  }
  // pk_correction, repeated while alpha is a number (powerspectrum.f90:201-204)
  const int nyq = nw / 2;
  const int64_t n1 = nyq + 1;
  for (int iter = 0; iter < 4; ++iter) {
    const double alpha = (std::log(interp1(x2, x5, nbin, (double)(float)nyq)) -
                          std::log(interp1(x2, x5, nbin, (double)(float)nyq / 2))) /
                         std::log(2.0);
    if (std::isnan(alpha)) break;
    zero_async(cb.p, sizeof(double) * 2 * nb, s);
    hipLaunchKernelGGL(k_cic_alias, dim3((unsigned)ceil_div(n1 * n1 * n1, 256)), dim3(256), sizeof(double) * 2 * nb,
                       s, nw, nbin, (float)alpha, cb.p);
    CUBE_HIP(hipGetLastError());
    std::vector<double> hc(2 * nb);
    CUBE_HIP(hipMemcpyAsync(hc.data(), cb.p, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, s));
    CUBE_HIP(hipStreamSynchronize(s));
    for (int b = 1; b < nb; ++b) x5[b] = x4[b] / (hc[nb + b] / hc[b]);
  }
  const double kfac = 2 * kPi / (double)box, b3 = (double)box * box * box;
  for (int b = 0; b < nb; ++b) {
    float* o = xi + (size_t)b * 10;
    for (int c = 0; c < 10; ++c) o[c] = 0;
    o[0] = (float)x1[b];
    o[1] = (float)(x2[b] * kfac);
    o[2] = (float)(x3[b] * b3);
    o[3] = (float)(x4[b] * b3);
    o[4] = (float)(x5[b] * b3);
  }
}

}  // namespace cube
