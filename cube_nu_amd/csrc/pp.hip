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
// no cross-thread accumulation is needed; the pair term (x2-x1)/r*(F(r,app) -
// F(r,r_pp)) is bitwise the reference's because xf is exact in FP32 and the
// operation order is the same.
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

__global__ void __launch_bounds__(256) k_pp_force(PPLevel L, const PPBox* __restrict__ boxes,
                                                  const int32_t* __restrict__ count,
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
    // brick-local cell of target t
    int lo = 0, hi = PB * PB * PB - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (pre[mid] <= t) lo = mid;
      else hi = mid - 1;
    }
    const int q = lo;
    const int cx = c0[0] + q % PB, cy = c0[1] + (q / PB) % PB, cz = c0[2] + q / (PB * PB);
    const int64_t tcell = cbase + ((int64_t)cz * m + cy) * m + cx;
    const float4 me = ent[start[tcell] + (t - pre[q])];
    float f1[3] = {0, 0, 0};
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
            const float r2 = (rv0 * rv0 + rv1 * rv1) + rv2 * rv2;
            if (0.0f < r2 && r2 < L.pp_range2) {
              const float rmag = sqrtf(r2);
              const float fr = f_ra(rmag, L.app) - f_ra(rmag, L.pp_range);
              f1[0] = f1[0] + (rv0 / rmag) * fr;
              f1[1] = f1[1] + (rv1 / rmag) * fr;
              f1[2] = f1[2] + (rv2 / rmag) * fr;
              if (ghost) ++npg;
              else ++npp;
            }
          }
        }
    // update v (kick.f90:576-582)
    const unsigned ip = __float_as_uint(me.w);
    float g[3];
    for (int d = 0; d < 3; ++d) g[d] = f1[d] * L.mass_scale;
    f2 = fmaxf(f2, (g[0] * g[0] + g[1] * g[1]) + g[2] * g[2]);
    for (int d = 0; d < 3; ++d) {
      float v = vp[3 * (int64_t)ip + d] + g[d] * L.v_scale;
      vp[3 * (int64_t)ip + d] = v;
      vm[d] = fmaxf(vm[d], fabsf(v));
    }
  }
  // block reductions
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
  atomicAdd(&cnt_red[0], npp);
  atomicAdd(&cnt_red[1], npg);
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(&out->pairs_pp, cnt_red[0]);
    atomicAdd(&out->pairs_pg, cnt_red[1]);
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

void pp_force(const PPLevel& L, const PPBox* boxes, int nbox, float* vp, PPScratch& S, PPOut* out, hipStream_t s) {
  const int nbr = (L.npgrid + PB - 1) / PB;
  const int64_t grid = (int64_t)nbr * nbr * nbr * nbox;
  hipLaunchKernelGGL(k_pp_force, dim3((unsigned)grid), dim3(256), 0, s, L, boxes, S.count, S.start, S.entries, vp,
                     out);
  CUBE_HIP(hipGetLastError());
}

}  // namespace cube
