// ic.hip -- seeded Zel'dovich input maker on the GPU.
//
// Steps follow utilities/ic.f90: uniform noise on the nic*ng grid (here a
// counter-based hash of (seed, cell), ic.f90:247-280), Box-Muller on x-pairs
// (ic.f90:289-298), forward FFT, the filter sqrt(P(k) (nw/box)^3)
// (ic.f90:186, 367-384) times the potential kernel -4 pi / k~^2 with
// k~ = 2 sin(pi k/nw) (ic.f90:418-437, 489), inverse FFT, the 2x2x2 gradient
// stencil at each Lagrangian particle (ic.f90:629-631, 703-705), the
// displacement -grad(phi)/(8 pi nic ratio_cs) and velocity -grad(phi)/(8 pi
// nic) * vf (ic.f90:706, 722), the compressed encode (ic.f90:708-718) and
// PID (ic.f90:725-727).  Unlike ic.f90, particles are binned globally, so
// large displacements are kept.  P(k) = A k^ns T_BBKS(k)^2 normalised to
// sigma8 with a top-hat of 8 Mpc/h, scaled by Dgrow(a)^2 (ic.f90:195-208).
#include <cmath>
#include <vector>

#include "fft.h"
#include "ic.h"
#include "pm.h"

namespace cube {

namespace {

__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// Uniform in (0,1) keyed by (seed, index).
__device__ inline double uniform01(uint64_t seed, uint64_t idx) {
  uint64_t z = mix64(seed * 0x9e3779b97f4a7c15ull + mix64(idx + 0x632be59bd9b4e019ull));
  return ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

__global__ void k_noise(float* rho, int64_t npair, uint64_t seed) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= npair) return;
  const double u1 = uniform01(seed, 2 * m), u2 = uniform01(seed, 2 * m + 1);
  const double th = 2 * kPi * u1, r = sqrt(-2 * log(1 - u2));
  rho[2 * m] = (float)(r * cos(th));
  rho[2 * m + 1] = (float)(r * sin(th));
}

struct FilterP {
  int nw;
  double box, amp, ns, gam, growth2, norm;
};

__device__ inline double bbks_t(double k, double gam) {
  const double q = k / gam;
  if (q <= 0) return 1.0;
  const double a = log(1 + 2.34 * q) / (2.34 * q);
  const double b = 1 + 3.89 * q + pow(16.1 * q, 2) + pow(5.46 * q, 3) + pow(6.71 * q, 4);
  return a * pow(b, -0.25);
}

__global__ void k_filter(float2* ck, FilterP F) {
  const int nw = F.nw, nh = nw / 2 + 1;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nh * nw * nw) return;
  const int ix = (int)(t % nh), iy = (int)((t / nh) % nw), iz = (int)(t / ((int64_t)nh * nw));
  const int kx = ix, ky = iy <= nw / 2 ? iy : iy - nw, kz = iz <= nw / 2 ? iz : iz - nw;
  if (kx == 0 && ky == 0 && kz == 0) {
    ck[t] = make_float2(0.f, 0.f);
    return;
  }
  const double kr = 2 * kPi * sqrt((double)kx * kx + (double)ky * ky + (double)kz * kz) / F.box;
  const double tk = bbks_t(kr, F.gam);
  const double pk = F.amp * pow(kr, F.ns) * tk * tk * F.growth2;
  const double filt = sqrt(pk * pow((double)nw / F.box, 3));
  const double sx = 2 * sin(kPi * kx / nw), sy = 2 * sin(kPi * ky / nw), sz = 2 * sin(kPi * kz / nw);
  const double k2 = fmax(sx * sx + sy * sy + sz * sz, 1.0 / ((double)nw * nw));
  const double f = filt * (-4 * kPi / k2) * F.norm;
  const float2 v = ck[t];
  ck[t] = make_float2((float)(v.x * f), (float)(v.y * f));
}

struct PartP {
  int ng, nw, nic, rcs, nc, nt, nnt, izipx;
  double vf;
};

// One thread per Lagrangian particle.
__global__ void k_particles(const float* __restrict__ phi, PartP P, int32_t* __restrict__ pcell,
                            int16_t* __restrict__ xq16, float* __restrict__ vq, int32_t* __restrict__ count) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t ng = P.ng;
  if (q >= ng * ng * ng) return;
  const int64_t p[3] = {q % ng, (q / ng) % ng, q / (ng * ng)};
  const int nw = P.nw;
  auto PH = [&](int64_t i, int64_t j, int64_t k) {
    i = (i % nw + nw) % nw; j = (j % nw + nw) % nw; k = (k % nw + nw) % nw;
    return (double)phi[(k * nw + j) * nw + i];
  };
  const int64_t ii = P.nic * p[0], jj = P.nic * p[1], kk = P.nic * p[2];
  double gp[3] = {0, 0, 0};
  for (int b = 0; b < 2; ++b)
    for (int a = 0; a < 2; ++a) {
      gp[0] += 0.5 * (PH(ii, jj - 1 + a, kk - 1 + b) - PH(ii - 1, jj - 1 + a, kk - 1 + b));
      gp[1] += 0.5 * (PH(ii - 1 + a, jj, kk - 1 + b) - PH(ii - 1 + a, jj - 1, kk - 1 + b));
      gp[2] += 0.5 * (PH(ii - 1 + a, jj - 1 + b, kk) - PH(ii - 1 + a, jj - 1 + b, kk - 1));
    }
  int gc[3];
  const double res = (double)(1 << (8 * P.izipx));
  for (int d = 0; d < 3; ++d) {
    const float gpf = (float)gp[d];
    double xq = (double)p[d] / P.rcs + (0.5 / P.nic) / P.rcs - (double)gpf / (8.0 * P.nic * kPi * P.rcs);
    xq = fmod(xq, (double)P.nc);
    if (xq < 0) xq += P.nc;
    int g = (int)ceil(xq);
    if (g == 0) g = P.nc;  // exact integer position: last cell of the period
    gc[d] = g - 1;
    const long long u = (long long)floor(xq * res);
    xq16[3 * q + d] = (int16_t)(u & 0xffff);
    vq[3 * q + d] = (float)(-(double)gpf / (8.0 * kPi * P.nic) * P.vf);
  }
  const int nt = P.nt;
  const int T = ((gc[2] / nt) * P.nnt + gc[1] / nt) * P.nnt + gc[0] / nt;
  const int32_t c = (int32_t)((((int64_t)T * nt + gc[2] % nt) * nt + gc[1] % nt) * nt + gc[0] % nt);
  pcell[q] = c;
  atomicAdd(&count[c], 1);
}

__global__ void k_place(const int32_t* __restrict__ pcell, const int64_t* __restrict__ off, int32_t* __restrict__ cur,
                        int64_t n, int64_t* __restrict__ keypid, int32_t* __restrict__ srcq) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int32_t c = pcell[q];
  const int64_t pos = off[c] + atomicAdd(&cur[c], 1);
  keypid[pos] = q + 1;
  srcq[pos] = (int32_t)q;
}

// Ascending pid inside each cell (insertion sort, one thread per cell).
__global__ void k_cellsort(const int64_t* __restrict__ off, int64_t ncell, int64_t* __restrict__ key,
                           int32_t* __restrict__ srcq) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const int64_t a = off[c], b = off[c + 1];
  for (int64_t i = a + 1; i < b; ++i) {
    const int64_t k = key[i];
    const int32_t s = srcq[i];
    int64_t j = i - 1;
    while (j >= a && key[j] > k) {
      key[j + 1] = key[j];
      srcq[j + 1] = srcq[j];
      --j;
    }
    key[j + 1] = k;
    srcq[j + 1] = s;
  }
}

__global__ void k_gather(const int32_t* __restrict__ srcq, const int64_t* __restrict__ key, int64_t n,
                         const int16_t* __restrict__ xq16, const float* __restrict__ vq, int izipx, char* __restrict__ xp,
                         float* __restrict__ vp, int64_t* __restrict__ pid, float* __restrict__ vzmax) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t q = srcq[i];
  for (int d = 0; d < 3; ++d) {
    if (izipx == 2) ((int16_t*)xp)[3 * i + d] = xq16[3 * (int64_t)q + d];
    else ((int8_t*)xp)[3 * i + d] = (int8_t)(uint8_t)(xq16[3 * (int64_t)q + d] & 0xff);
    vp[3 * i + d] = vq[3 * (int64_t)q + d];
  }
  pid[i] = key[i];
  atomic_max_nonneg(vzmax, fabsf(vq[3 * (int64_t)q + 2]));
}

__global__ void k_rhoc_from_counts(Geo g, const int32_t* __restrict__ count, int32_t* __restrict__ rhoc) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.ncells_ext) return;
  const int64_t T = c / g.ncell_tile, e = c % g.ncell_tile;
  const int64_t i = e % g.nte + 1 - g.ncb, j = (e / g.nte) % g.nte + 1 - g.ncb, k = e / (g.nte * g.nte) + 1 - g.ncb;
  const bool ph = i >= 1 && i <= g.nt && j >= 1 && j <= g.nt && k >= 1 && k <= g.nt;
  rhoc[c] = ph ? count[((T * g.nt + k - 1) * g.nt + j - 1) * g.nt + i - 1] : 0;
}

template <typename T>
struct Tmp {
  T* p = nullptr;
  explicit Tmp(size_t n) { CUBE_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T))); }
  ~Tmp() { if (p) (void)hipFree(p); }
};

double tophat(double x) { return x < 1e-6 ? 1.0 : 3 * (std::sin(x) - std::cos(x) * x) / (x * x * x); }

double bbks_host(double k, double gam) {
  const double q = k / gam;
  const double a = std::log(1 + 2.34 * q) / (2.34 * q);
  const double b = 1 + 3.89 * q + std::pow(16.1 * q, 2) + std::pow(5.46 * q, 3) + std::pow(6.71 * q, 4);
  return a * std::pow(b, -0.25);
}

}  // namespace

// Dgrow (ic.f90:922-935, main.f90:65-78), Carroll-Press-Turner growth.
double dgrow(double a, double om, double ol) {
  const double hsq = om / (a * a * a) + (1 - om - ol) / (a * a) + ol;
  const double oma = om / (a * a * a * hsq), ola = ol / hsq;
  const double g = 2.5 * om / (std::pow(om, 4.0 / 7) - ol + (1 + om / 2) * (1 + ol / 70));
  const double ga = 2.5 * oma / (std::pow(oma, 4.0 / 7) - ola + (1 + oma / 2) * (1 + ola / 70));
  return a * ga / g;
}

// Amplitude A with sigma8 of A k^ns T^2 at a=1 equal to p.sigma8.
double bbks_sigma8_norm(const cube_ic_params& p) {
  const double gam = p.omega_m * p.h;
  const int N = 20000;
  const double l0 = std::log(1e-5), l1 = std::log(1e3);
  double s = 0;
  for (int i = 0; i <= N; ++i) {
    const double lk = l0 + (l1 - l0) * i / N, k = std::exp(lk);
    const double t = bbks_host(k, gam), w = tophat(8.0 * k);
    const double f = k * k * k * std::pow(k, p.n_s) * t * t * w * w / (2 * kPi * kPi);
    const double c = (i == 0 || i == N) ? 1 : (i % 2 ? 4 : 2);
    s += c * f;
  }
  s *= (l1 - l0) / N / 3;
  return p.sigma8 * p.sigma8 / s;
}

ICResult make_ic(const Geo& g, const cube_ic_params& p, void* xp, float* vp, int64_t* pid, int32_t* rhoc,
                 hipStream_t s) {
  if (p.nic < 1) throw Error(1, "nic must be >= 1");
  const int nw = (int)(p.nic * g.ng);
  const int64_t nreal = (int64_t)nw * nw * nw, ncplx = (int64_t)nw * nw * (nw / 2 + 1);
  const int64_t N = g.np_image;
  FFTCube<float> fc;
  fc.init(nw);
  ICResult res{};
  {
    Tmp<float> rho(nreal);
    Tmp<float2> ck(ncplx);
    hipLaunchKernelGGL(k_noise, dim3((unsigned)ceil_div(nreal / 2, 256)), dim3(256), 0, s, rho.p, nreal / 2, p.seed);
    CUBE_HIP(hipGetLastError());
    fft_r2c3d<float>(fc, rho.p, ck.p, 1, s);
    FilterP F;
    F.nw = nw;
    F.box = p.box;
    F.amp = bbks_sigma8_norm(p);
    F.ns = p.n_s;
    F.gam = p.omega_m * p.h;
    const double a = 1.0 / (1.0 + p.z);
    const double D = dgrow(a, p.omega_m, 1.0 - p.omega_m);
    F.growth2 = D * D;
    F.norm = 1.0 / (double)nreal;  // pencil_fft_backward's 1/nw^3
    hipLaunchKernelGGL(k_filter, dim3((unsigned)ceil_div(ncplx, 256)), dim3(256), 0, s, ck.p, F);
    CUBE_HIP(hipGetLastError());
    fft_c2r3d<float>(fc, ck.p, rho.p, 1, s);
    // particles
    PartP P;
    P.ng = (int)g.ng; P.nw = nw; P.nic = p.nic; P.rcs = (int)g.rcs; P.nc = (int)g.nc; P.nt = (int)g.nt;
    P.nnt = (int)g.nnt; P.izipx = (int)g.izipx;
    // vfactor(a) = a^2 H, H = 2/(3 sqrt(a^3)) sqrt(1 + a km + a^3 lm) (ic.f90:937-945)
    const double om = p.omega_m, ol = 1.0 - p.omega_m;
    const double lm = ol / om, km = (1 - om - ol) / om;
    const double H = 2.0 / (3.0 * std::sqrt(a * a * a)) * std::sqrt(1 + a * km + a * a * a * lm);
    P.vf = (double)(float)(a * a * H);
    const int64_t nphys_cells = g.ntiles * g.nt * g.nt * g.nt;
    Tmp<int32_t> pcell(N), count(nphys_cells), curs(nphys_cells), srcq(N);
    Tmp<int16_t> xq16(3 * N);
    Tmp<float> vq(3 * N);
    Tmp<int64_t> off(nphys_cells + 1), key(N);
    Tmp<float> vz(1);
    Tmp<char> stmp(scan_tmp_bytes(nphys_cells));
    zero_async(count.p, sizeof(int32_t) * nphys_cells, s);
    zero_async(curs.p, sizeof(int32_t) * nphys_cells, s);
    zero_async(vz.p, sizeof(float), s);
    hipLaunchKernelGGL(k_particles, dim3((unsigned)ceil_div(N, 256)), dim3(256), 0, s, rho.p, P, pcell.p, xq16.p,
                       vq.p, count.p);
    CUBE_HIP(hipGetLastError());
    scan_counts(count.p, off.p, nphys_cells, stmp.p, scan_tmp_bytes(nphys_cells), s);
    hipLaunchKernelGGL(k_place, dim3((unsigned)ceil_div(N, 256)), dim3(256), 0, s, pcell.p, off.p, curs.p, N, key.p,
                       srcq.p);
    hipLaunchKernelGGL(k_cellsort, dim3((unsigned)ceil_div(nphys_cells, 256)), dim3(256), 0, s, off.p, nphys_cells,
                       key.p, srcq.p);
    hipLaunchKernelGGL(k_gather, dim3((unsigned)ceil_div(N, 256)), dim3(256), 0, s, srcq.p, key.p, N, xq16.p, vq.p,
                       (int)g.izipx, (char*)xp, vp, pid, vz.p);
    hipLaunchKernelGGL(k_rhoc_from_counts, dim3((unsigned)ceil_div(g.ncells_ext, 256)), dim3(256), 0, s, g, count.p,
                       rhoc);
    CUBE_HIP(hipGetLastError());
    CUBE_HIP(hipMemcpyAsync(&res.vz_max, vz.p, sizeof(float), hipMemcpyDeviceToHost, s));
    CUBE_HIP(hipStreamSynchronize(s));
  }
  fc.release();
  res.nplocal = N;
  return res;
}

namespace {

constexpr int kFDTable = 8192;  // inverse-CDF points, uniform in probability
constexpr double kFDUmax = 40.0;

// u at CDF = i/kFDTable for u^2/(e^u + 1) on [0, kFDUmax] (the mass beyond
// 40 is 1e-15 of the total): Simpson integration on a fine grid in FP64,
// then linear inversion.
std::vector<double> fd_inverse_cdf() {
  const int nf = 400000;
  const double h = kFDUmax / nf;
  std::vector<double> cdf(nf + 1, 0.0);
  auto f = [](double u) { return u * u / (std::exp(u) + 1.0); };
  for (int i = 0; i < nf; ++i) {
    const double a = i * h;
    cdf[i + 1] = cdf[i] + h / 6.0 * (f(a) + 4.0 * f(a + 0.5 * h) + f(a + h));
  }
  const double tot = cdf[nf];
  std::vector<double> inv(kFDTable + 1);
  int j = 0;
  for (int i = 0; i <= kFDTable; ++i) {
    const double c = tot * (double)i / kFDTable;
    while (j < nf - 1 && cdf[j + 1] < c) ++j;
    const double d = cdf[j + 1] - cdf[j];
    inv[i] = (j + (d > 0 ? (c - cdf[j]) / d : 0.0)) * h;
  }
  inv[0] = 0.0;
  return inv;
}

__global__ void k_thermal(float* __restrict__ vp, const int64_t* __restrict__ pid, int64_t n, uint64_t seed,
                          double vscale, const double* __restrict__ icdf, float* __restrict__ vzmax) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float vz = 0.0f;
  if (i < n) {
    const uint64_t key = (uint64_t)pid[i];
    const double r1 = uniform01(seed, 3 * key), r2 = uniform01(seed, 3 * key + 1), r3 = uniform01(seed, 3 * key + 2);
    const double t = r1 * kFDTable;
    const int k = min((int)t, kFDTable - 1);
    const double u = icdf[k] + (t - k) * (icdf[k + 1] - icdf[k]);
    const double ct = 2.0 * r2 - 1.0, st = sqrt(fmax(0.0, 1.0 - ct * ct)), ph = 2.0 * kPi * r3;
    const double w = u * vscale;
    const double dv[3] = {w * st * cos(ph), w * st * sin(ph), w * ct};
    for (int d = 0; d < 3; ++d) vp[3 * i + d] = (float)((double)vp[3 * i + d] + dv[d]);
    vz = fabsf(vp[3 * i + 2]);
  }
  for (int o = 32; o > 0; o >>= 1) vz = fmaxf(vz, __shfl_down(vz, o, 64));
  if ((threadIdx.x & 63) == 0 && vz > 0) atomic_max_nonneg(vzmax, vz);
}

}  // namespace

float add_thermal_velocities(float* vp, const int64_t* pid, int64_t n, double v_scale, uint64_t seed,
                             hipStream_t s) {
  const std::vector<double> inv = fd_inverse_cdf();
  Tmp<double> tab(kFDTable + 1);
  Tmp<float> vz(1);
  CUBE_HIP(hipMemcpyAsync(tab.p, inv.data(), sizeof(double) * inv.size(), hipMemcpyHostToDevice, s));
  zero_async(vz.p, sizeof(float), s);
  if (n > 0) {
    hipLaunchKernelGGL(k_thermal, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, vp, pid, n, seed, v_scale,
                       tab.p, vz.p);
    CUBE_HIP(hipGetLastError());
  }
  float out = 0;
  CUBE_HIP(hipMemcpyAsync(&out, vz.p, sizeof(float), hipMemcpyDeviceToHost, s));
  CUBE_HIP(hipStreamSynchronize(s));
  return out;
}

}  // namespace cube
