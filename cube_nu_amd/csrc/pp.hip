// pp.hip -- PP_Force (kick.f90:376-588) on the GPU.
//
// Binning: every PP cell (1/ratio_sf of a coarse cell) is nested in exactly
// one coarse cell, so one thread per (sub-tile, coarse cell) owns the PP cells
// it fills: counts and in-cell order are deterministic without atomics, and
// the in-cell order is the reference's (coarse-cell particle order).
//
// Force: one workgroup per brick of PB^3 physical PP cells; each thread owns
// target particles and sums the pair force from the 27 neighbouring PP cells
// in a fixed order.  Both members of a physical pair evaluate the pair, so
// no cross-thread accumulation is needed and the result is deterministic.
// The in-range test uses the reference's FP32 operations on exact FP32
// positions, so the set of interacting pairs is the reference's.
#include "pm.h"
#include "pp.h"

namespace cube {
namespace {

template <int IZ>
using XT = typename XP<IZ>::T;

// PP cell of a particle, kick.f90:420-431 in FP32.
template <int IZ>
__device__ inline bool pp_cell(const PPLevel& L, const PPBox& B, int ci, int cj, int ck, const XT<IZ>* xp, int64_t ip,
                               float xf[3], int idx[3]) {
  const int ijk[3] = {ci, cj, ck};
  bool in = true;
  for (int d = 0; d < 3; ++d) {
    float xvec = (float)((double)(ijk[d] - 1) + XP<IZ>::off(xp, ip, d));
    xf[d] = xvec * (float)L.rcs;
    float xv = (float)L.sf * (xvec - (float)(B.ifs[d] * L.ntt));
    idx[d] = (int)floorf(xv) + 1;
    in = in && idx[d] >= 0 && idx[d] <= L.npgrid + 1;
  }
  return in;
}

// One thread per (box, coarse cell of the sub-tile + 1-cell ring).
template <int IZ, int FILL>
__global__ void k_pp_bin(PPLevel L, const PPBox* __restrict__ boxes, int nbox, const XT<IZ>* __restrict__ xp,
                         const int32_t* __restrict__ rhoc, const int64_t* __restrict__ off, int32_t* __restrict__ count,
                         const int64_t* __restrict__ start, float4* __restrict__ ent) {
  const int nr = L.ntt + 2;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nbox * nr * nr * nr) return;
  const int ib = (int)(t / (nr * nr * nr));
  const int q = (int)(t % (nr * nr * nr));
  const PPBox B = boxes[ib];
  const int ci = B.ifs[0] * L.ntt + (q % nr);  // nc1-1 + offset  (Fortran: nc1 = ifs*ntt+1)
  const int cj = B.ifs[1] * L.ntt + (q / nr) % nr;
  const int ck = B.ifs[2] * L.ntt + q / (nr * nr);
  const int64_t cell = rc_index(ci, cj, ck, B.T, L.nte, L.ncb);
  const int cnt = rhoc[cell];
  const int64_t base = off[cell];
  const int m = L.npgrid + 2;
  int32_t* cb = count + (int64_t)ib * m * m * m;
  for (int l = 0; l < cnt; ++l) {
    const int64_t ip = base + l;
    float xf[3];
    int idx[3];
    if (!pp_cell<IZ>(L, B, ci, cj, ck, xp, ip, xf, idx)) continue;
    const int64_t pc = ((int64_t)idx[2] * m + idx[1]) * m + idx[0];
    if (FILL) {
      const int64_t pos = start[(int64_t)ib * m * m * m + pc] + cb[pc];
      ent[pos] = make_float4(xf[0], xf[1], xf[2], __uint_as_float((unsigned)ip));
    }
    cb[pc] += 1;
  }
}

constexpr int PB = 4;  // physical PP cells per brick per dim

// Pair force of candidate o on target me (kick.f90:523-530, 554-560):
// f1 += (x2-x1)/r * (F(r,app) - F(r,r_pp)).
// FP32 form: r^-1 from the hardware reciprocal square root and the F_ra
// divisions replaced by products with precomputed reciprocals; the error
// of a pair term stays at the level of FP32 rounding (the exact-division
// form cost 1.7x the time).  FP64 form: the reference's operations in FP64.
template <typename R>
__device__ inline void pair_force(const PPLevel& L, const float4 me, const float4 o, R f1[3]);

template <>
__device__ inline void pair_force<float>(const PPLevel& L, const float4 me, const float4 o, float f1[3]) {
  const float rv0 = o.x - me.x, rv1 = o.y - me.y, rv2 = o.z - me.z;
  const float r2 = (rv0 * rv0 + rv1 * rv1) + rv2 * rv2;
  const float rinv = __builtin_amdgcn_rsqf(r2);
  const float r = r2 * rinv;
  auto F = [&](float inv_apm, float c) -> float {
    const float ep = (2.0f * r) * inv_apm;
    if (ep > 2.0f) return rinv * rinv;
    const float e2 = ep * ep, e3 = e2 * ep, e4 = e2 * e2, e5 = e4 * ep, e6 = e3 * e3;
    if (ep > 1.0f)
      return c * ((((((12.0f * __builtin_amdgcn_rcpf(e2) - 224.0f) + 896.0f * ep) - 840.0f * e2) + 224.0f * e3) +
                   70.0f * e4 - 48.0f * e5) +
                  7.0f * e6);
    return c * ((((224.0f * ep - 224.0f * e3) + 70.0f * e4) + 48.0f * e5) - 21.0f * e6);
  };
  const float s = rinv * (F(L.inv_app, L.c_app) - F(L.inv_range, L.c_range));
  f1[0] = f1[0] + rv0 * s;
  f1[1] = f1[1] + rv1 * s;
  f1[2] = f1[2] + rv2 * s;
}

template <>
__device__ inline void pair_force<double>(const PPLevel& L, const float4 me, const float4 o, double f1[3]) {
  const double rv0 = (double)o.x - me.x, rv1 = (double)o.y - me.y, rv2 = (double)o.z - me.z;
  const double r2 = (rv0 * rv0 + rv1 * rv1) + rv2 * rv2;
  const double r = sqrt(r2);
  const double fr = f_ra<double>(r, (double)L.app, f_ra_coef<double>((double)L.app)) -
                    f_ra<double>(r, (double)L.pp_range, f_ra_coef<double>((double)L.pp_range));
  f1[0] = f1[0] + (rv0 / r) * fr;
  f1[1] = f1[1] + (rv1 / r) * fr;
  f1[2] = f1[2] + (rv2 / r) * fr;
}

// One workgroup per brick of PB^3 physical PP cells; each thread owns target
// particles and sums the pair forces of the 27 neighbouring PP cells in the
// fixed (dz, dy, dx, p) order.  Neighbouring targets read the same
// candidates, which therefore come from L1.
template <typename R>
__global__ void __launch_bounds__(256) k_pp_force(PPLevel L, const int32_t* __restrict__ count,
                                                  const int64_t* __restrict__ start,
                                                  const float4* __restrict__ ent, float* __restrict__ vp,
                                                  PPOut* __restrict__ out) {
  __shared__ int pre[PB * PB * PB + 1];
  __shared__ float red[4];
  __shared__ unsigned long long cnt_red[2];
  const int nbr = (L.npgrid + PB - 1) / PB;
  int64_t b = blockIdx.x;
  const int bx = (int)(b % nbr);
  b /= nbr;
  const int by = (int)(b % nbr);
  b /= nbr;
  const int bz = (int)(b % nbr);
  const int ib = (int)(b / nbr);
  const int m = L.npgrid + 2;
  const int64_t cbase = (int64_t)ib * m * m * m;
  const int c0[3] = {1 + bx * PB, 1 + by * PB, 1 + bz * PB};
  if (threadIdx.x == 0) {
    int s = 0;
    for (int q = 0; q < PB * PB * PB; ++q) {
      pre[q] = s;
      const int cx = c0[0] + q % PB, cy = c0[1] + (q / PB) % PB, cz = c0[2] + q / (PB * PB);
      if (cx <= L.npgrid && cy <= L.npgrid && cz <= L.npgrid) s += count[cbase + ((int64_t)cz * m + cy) * m + cx];
    }
    pre[PB * PB * PB] = s;
    cnt_red[0] = 0;
    cnt_red[1] = 0;
  }
  __syncthreads();
  const int ntg = pre[PB * PB * PB];
  float f2 = 0, vm[3] = {0, 0, 0};
  unsigned long long npp = 0, npg = 0;
  for (int t = threadIdx.x; t < ntg; t += blockDim.x) {
    int lo = 0, hi = PB * PB * PB - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (pre[mid] <= t) lo = mid;
      else hi = mid - 1;
    }
    const int q = lo;
    const int cx = c0[0] + q % PB, cy = c0[1] + (q / PB) % PB, cz = c0[2] + q / (PB * PB);
    const int64_t tcell = cbase + ((int64_t)cz * m + cy) * m + cx;
    const float4 me = ent[start[tcell] + (t - pre[q])];
    R f1[3] = {0, 0, 0};
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int nx = cx + dx, ny = cy + dy, nz = cz + dz;
          const bool ghost = nx == 0 || ny == 0 || nz == 0 || nx == L.npgrid + 1 || ny == L.npgrid + 1 ||
                             nz == L.npgrid + 1;
          const int64_t nc = cbase + ((int64_t)nz * m + ny) * m + nx;
          const int64_t s0 = start[nc];
          const int nn = count[nc];
          for (int p = 0; p < nn; ++p) {
            const float4 o = ent[s0 + p];
            const float rv0 = o.x - me.x, rv1 = o.y - me.y, rv2 = o.z - me.z;
            const float r2 = (rv0 * rv0 + rv1 * rv1) + rv2 * rv2;  // exact-input FP32 test (kick.f90:525, 556)
            if (0.0f < r2 && r2 < L.pp_range2) {
              pair_force<R>(L, me, o, f1);
              if (ghost) ++npg;
              else ++npp;
            }
          }
        }
    // update v (kick.f90:576-582) in FP32 as the reference does
    const unsigned ip = __float_as_uint(me.w);
    float g[3];
    for (int d = 0; d < 3; ++d) g[d] = (float)f1[d] * L.mass_scale;
    f2 = fmaxf(f2, (g[0] * g[0] + g[1] * g[1]) + g[2] * g[2]);
    for (int d = 0; d < 3; ++d) {
      float v = vp[3 * (int64_t)ip + d] + g[d] * L.v_scale;
      vp[3 * (int64_t)ip + d] = v;
      vm[d] = fmaxf(vm[d], fabsf(v));
    }
  }
  auto bmax = [&](float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_down(v, o, 64));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    float r = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r = fmaxf(r, red[w]);
    return r;
  };
  float r = bmax(f2);
  if (threadIdx.x == 0 && r > 0) atomic_max_nonneg(&out->f2max, r);
  for (int d = 0; d < 3; ++d) {
    r = bmax(vm[d]);
    if (threadIdx.x == 0 && r > 0) atomic_max_nonneg(&out->vmax[d], r);
  }
  for (int o = 32; o > 0; o >>= 1) {
    npp += __shfl_down(npp, o, 64);
    npg += __shfl_down(npg, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&cnt_red[0], npp);
    atomicAdd(&cnt_red[1], npg);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(&out->pairs_pp[blockIdx.x % kSumSlots], cnt_red[0]);
    atomicAdd(&out->pairs_pg[blockIdx.x % kSumSlots], cnt_red[1]);
  }
}

}  // namespace

int64_t pp_bin_count(int izipx, const PPLevel& L, const PPBox* boxes, int nbox, const void* xp, const int32_t* rhoc,
                     const int64_t* off, PPScratch& S, hipStream_t s) {
  const int m = L.npgrid + 2;
  const int64_t ncell = (int64_t)nbox * m * m * m;
  zero_async(S.count, sizeof(int32_t) * ncell, s);
  const int nr = L.ntt + 2;
  const int64_t nthr = (int64_t)nbox * nr * nr * nr;
  if (izipx == 2)
    hipLaunchKernelGGL((k_pp_bin<2, 0>), dim3((unsigned)ceil_div(nthr, 256)), dim3(256), 0, s, L, boxes, nbox,
                       (const int16_t*)xp, rhoc, off, S.count, nullptr, nullptr);
  else
    hipLaunchKernelGGL((k_pp_bin<1, 0>), dim3((unsigned)ceil_div(nthr, 256)), dim3(256), 0, s, L, boxes, nbox,
                       (const int8_t*)xp, rhoc, off, S.count, nullptr, nullptr);
  CUBE_HIP(hipGetLastError());
  scan_counts(S.count, S.start, ncell, S.scan_tmp, S.scan_tmp_bytes, s);
  int64_t total = 0;
  CUBE_HIP(hipMemcpyAsync(&total, S.start + ncell, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  CUBE_HIP(hipStreamSynchronize(s));
  return total;
}

void pp_bin_fill(int izipx, const PPLevel& L, const PPBox* boxes, int nbox, const void* xp, const int32_t* rhoc,
                 const int64_t* off, PPScratch& S, hipStream_t s) {
  const int m = L.npgrid + 2;
  const int64_t ncell = (int64_t)nbox * m * m * m;
  zero_async(S.count, sizeof(int32_t) * ncell, s);
  const int nr = L.ntt + 2;
  const int64_t nthr = (int64_t)nbox * nr * nr * nr;
  if (izipx == 2)
    hipLaunchKernelGGL((k_pp_bin<2, 1>), dim3((unsigned)ceil_div(nthr, 256)), dim3(256), 0, s, L, boxes, nbox,
                       (const int16_t*)xp, rhoc, off, S.count, S.start, S.entries);
  else
    hipLaunchKernelGGL((k_pp_bin<1, 1>), dim3((unsigned)ceil_div(nthr, 256)), dim3(256), 0, s, L, boxes, nbox,
                       (const int8_t*)xp, rhoc, off, S.count, S.start, S.entries);
  CUBE_HIP(hipGetLastError());
}

void pp_force(const PPLevel& L, int precision, int nbox, float* vp, PPScratch& S, PPOut* out, hipStream_t s) {
  const int nbr = (L.npgrid + PB - 1) / PB;
  const int64_t grid = (int64_t)nbr * nbr * nbr * nbox;
  if (precision == 64)
    hipLaunchKernelGGL(k_pp_force<double>, dim3((unsigned)grid), dim3(256), 0, s, L, S.count, S.start, S.entries, vp,
                       out);
  else
    hipLaunchKernelGGL(k_pp_force<float>, dim3((unsigned)grid), dim3(256), 0, s, L, S.count, S.start, S.entries, vp,
                       out);
  CUBE_HIP(hipGetLastError());
}

}  // namespace cube
