// pp.hip -- PP_Force (kick.f90:376-588) on the GPU.
//
// Binning: every PP cell (1/ratio_sf of a coarse cell) is nested in exactly
// one coarse cell, so one wavefront per (sub-tile, coarse cell) owns the PP
// cells it fills: counts and in-cell order are deterministic without
// atomics, and the in-cell order is the reference's (coarse-cell particle
// order).
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

// One wavefront per (box, coarse cell of the sub-tile + 1-cell ring), one
// particle per lane.  A PP cell lies inside one coarse cell (idx =
// floor(sf*(xvec - ifs*ntt)) + 1 with xvec < c), so the wavefront owns the
// counts of the sf^3 PP cells of its coarse cell and keeps them in LDS: for
// each distinct PP cell among 64 particles (one ballot per cell) the lanes
// get their rank from a population count, so the in-cell order is the
// reference's coarse-cell particle order.  FILL=0 writes the counts, FILL=1
// writes the entries at start + rank.
template <int IZ, int FILL>
__global__ void __launch_bounds__(256) k_pp_bin(PPLevel L, const PPBox* __restrict__ boxes, int nbox,
                                                const XT<IZ>* __restrict__ xp, const int32_t* __restrict__ rhoc,
                                                const int64_t* __restrict__ off, int32_t* __restrict__ count,
                                                const int64_t* __restrict__ start, float4* __restrict__ ent) {
  __shared__ int lcnt[4][512];  // sf^3 <= 512 (ratio_sf <= 8)
  const int nr = L.ntt + 2;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t wv = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (wv >= (int64_t)nbox * nr * nr * nr) return;
  const int ib = (int)(wv / (nr * nr * nr));
  const int q = (int)(wv % (nr * nr * nr));
  const PPBox B = boxes[ib];
  const int c[3] = {B.ifs[0] * L.ntt + (q % nr), B.ifs[1] * L.ntt + (q / nr) % nr,
                    B.ifs[2] * L.ntt + q / (nr * nr)};  // nc1-1 + offset (Fortran nc1 = ifs*ntt+1)
  const int64_t cell = rc_index(c[0], c[1], c[2], B.T, L.nte, L.ncb);
  const int cnt = rhoc[cell];
  const int64_t base = off[cell];
  const int m = L.npgrid + 2;
  const int sf = L.sf, sf3 = sf * sf * sf;
  int lo[3];  // PP index of the coarse cell's first PP cell per axis
  for (int d = 0; d < 3; ++d) lo[d] = sf * (c[d] - 1 - B.ifs[d] * L.ntt) + 1;
  int* lc = lcnt[wave];
  for (int t = lane; t < sf3; t += 64) lc[t] = 0;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int64_t cbase = (int64_t)ib * m * m * m;
  for (int l0 = 0; l0 < cnt; l0 += 64) {
    const int l = l0 + lane;
    float xf[3] = {0, 0, 0};
    int idx[3] = {0, 0, 0};
    bool in = false;
    if (l < cnt) in = pp_cell<IZ>(L, B, c[0], c[1], c[2], xp, base + l, xf, idx);
    const int ls = in ? ((idx[2] - lo[2]) * sf + (idx[1] - lo[1])) * sf + (idx[0] - lo[0]) : -1;
    bool todo = in;
    while (true) {
      const unsigned long long act = __ballot(todo);
      if (act == 0) break;
      const int leader = __ffsll((long long)act) - 1;
      const int lls = __shfl(ls, leader, 64);
      const unsigned long long mk = __ballot(todo && ls == lls);
      const int prev = lc[lls];  // wave-uniform read
      if (todo && ls == lls) {
        if (FILL) {
          const int64_t pc = cbase + ((int64_t)idx[2] * m + idx[1]) * m + idx[0];
          const int64_t pos = start[pc] + prev + (int)__popcll(mk & ((1ull << lane) - 1ull));
          ent[pos] = make_float4(xf[0], xf[1], xf[2], __uint_as_float((unsigned)(base + l)));
        }
        todo = false;
      }
      __builtin_amdgcn_wave_barrier();
      if (lane == leader) lc[lls] = prev + (int)__popcll(mk);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
  if (!FILL) {
    for (int t = lane; t < sf3; t += 64) {
      const int n = lc[t];
      if (n == 0) continue;
      const int ix = lo[0] + t % sf, iy = lo[1] + (t / sf) % sf, iz = lo[2] + t / (sf * sf);
      count[cbase + ((int64_t)iz * m + iy) * m + ix] = n;
    }
  }
}

constexpr int PB = 4;       // physical PP cells per brick per dim
constexpr int kUnroll = 8;  // candidate loads in flight per lane


// FP32 pair law: F(r, app) - F(r, r_pp) with the
// polynomials in Horner form (FMA) and both branches of F(r, r_pp) evaluated
// without divergence; r < r_pp for every queued pair, so F(r, r_pp) is one of
// the two polynomials.  F(r, app) = 1/r^2 except for r <= app (rare).
__device__ inline float pair_law_fast(const PPLevel& L, float r2, float& rinv_out) {
  const float rinv = __builtin_amdgcn_rsqf(r2);
  const float r = r2 * rinv;
  float fa = rinv * rinv;
  const float ea = (2.0f * r) * L.inv_app;
  if (ea <= 2.0f) {
    const float e2 = ea * ea;
    fa = ea > 1.0f
             ? L.c_app * (12.0f * __builtin_amdgcn_rcpf(e2) +
                          fmaf(ea, fmaf(ea, fmaf(ea, fmaf(ea, fmaf(ea, fmaf(ea, 7.0f, -48.0f), 70.0f), 224.0f),
                                                 -840.0f),
                                        896.0f),
                               -224.0f))
             : L.c_app * ea * fmaf(e2, fmaf(ea, fmaf(ea, fmaf(ea, -21.0f, 48.0f), 70.0f), -224.0f), 224.0f);
  }
  const float ep = (2.0f * r) * L.inv_range;
  const float e2 = ep * ep;
  const float outer =
      12.0f * __builtin_amdgcn_rcpf(e2) +
      fmaf(ep, fmaf(ep, fmaf(ep, fmaf(ep, fmaf(ep, fmaf(ep, 7.0f, -48.0f), 70.0f), 224.0f), -840.0f), 896.0f), -224.0f);
  const float inner = ep * fmaf(e2, fmaf(ep, fmaf(ep, fmaf(ep, -21.0f, 48.0f), 70.0f), -224.0f), 224.0f);
  const float fr = L.c_range * (ep > 1.0f ? outer : inner);
  rinv_out = rinv;
  return fa - fr;
}

// Pair force of candidate o on target me (kick.f90:523-530, 554-560):
// f1 += (x2-x1)/r * (F(r,app) - F(r,r_pp)).
// FP32 form: r^-1 from the hardware reciprocal square root, the F_ra
// divisions replaced by products with precomputed reciprocals and the
// polynomials in Horner form (pair_law_fast); the error of a pair term stays
// at the level of FP32 rounding.  FP64 form: the reference's operations in
// FP64.
template <typename R>
__device__ inline void pair_force(const PPLevel& L, const float4 me, const float4 o, R f1[3]);

template <>
__device__ inline void pair_force<float>(const PPLevel& L, const float4 me, const float4 o, float f1[3]) {
  const float rv0 = o.x - me.x, rv1 = o.y - me.y, rv2 = o.z - me.z;
  const float r2 = (rv0 * rv0 + rv1 * rv1) + rv2 * rv2;
  float rinv;
  const float s = pair_law_fast(L, r2, rinv);
  const float sc = rinv * s;
  f1[0] = f1[0] + rv0 * sc;
  f1[1] = f1[1] + rv1 * sc;
  f1[2] = f1[2] + rv2 * sc;
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
// fixed (dz, dy, dx, p) order.  The (PB+2)^3 neighbourhood's counts and
// starts are staged in LDS once; the three cells of each (dy, dz) column are
// one contiguous run of candidates, read kUnroll at a time so that many loads
// are in flight per lane (they come from L1/L2: neighbouring targets read the
// same candidates).  Candidates stay in the caches: staging them in LDS
// measured slower (64 ms vs 48 ms at 512^3).
template <typename R>
__global__ void __launch_bounds__(256) k_pp_force(PPLevel L, const int32_t* __restrict__ count,
                                                  const int64_t* __restrict__ start,
                                                  const float4* __restrict__ ent, float* __restrict__ vp,
                                                  PPOut* __restrict__ out) {
  constexpr int NB = PB + 2, NB3 = NB * NB * NB;
  __shared__ int pre[PB * PB * PB + 1];
  __shared__ int ncnt[NB3];
  __shared__ int64_t nstart[NB3];
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
  // neighbourhood cells c0-1 .. c0+PB (clipped to the PP grid 0..npgrid+1)
  for (int q = threadIdx.x; q < NB3; q += blockDim.x) {
    const int x = c0[0] - 1 + q % NB, y = c0[1] - 1 + (q / NB) % NB, z = c0[2] - 1 + q / (NB * NB);
    int c = 0;
    int64_t st = 0;
    if (x <= L.npgrid + 1 && y <= L.npgrid + 1 && z <= L.npgrid + 1) {
      const int64_t nc = cbase + ((int64_t)z * m + y) * m + x;
      c = count[nc];
      st = start[nc];
    }
    ncnt[q] = c;
    nstart[q] = st;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int q = 0; q < PB * PB * PB; ++q) {
      pre[q] = s;
      const int lx = q % PB, ly = (q / PB) % PB, lz = q / (PB * PB);
      if (c0[0] + lx <= L.npgrid && c0[1] + ly <= L.npgrid && c0[2] + lz <= L.npgrid)
        s += ncnt[((lz + 1) * NB + ly + 1) * NB + lx + 1];
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
    const int lx = lo % PB, ly = (lo / PB) % PB, lz = lo / (PB * PB);
    const int tq = ((lz + 1) * NB + ly + 1) * NB + lx + 1;
    const float4 me = ent[nstart[tq] + (t - pre[lo])];
    R f1[3] = {0, 0, 0};
    // PP cells are numbered x fastest, so the cells x-1, x, x+1 of a (dy, dz)
    // column are one contiguous run of the CSR: 9 runs per target
    const bool gx0 = c0[0] + lx - 1 == 0, gx1 = c0[0] + lx + 1 == L.npgrid + 1;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy) {
        const int ny = c0[1] + ly + dy, nz = c0[2] + lz + dz;
        const bool gyz = ny == 0 || nz == 0 || ny == L.npgrid + 1 || nz == L.npgrid + 1;
        const int q0 = tq + (dz * NB + dy) * NB - 1;  // cell x-1 of the column
        const int n0 = ncnt[q0], n1 = ncnt[q0 + 1];
        const int nn = n0 + n1 + ncnt[q0 + 2];
        const float4* run = ent + nstart[q0];
        for (int p = 0; p < nn; p += kUnroll) {
          float4 o[kUnroll];
#pragma unroll
          for (int k = 0; k < kUnroll; ++k) o[k] = run[min(p + k, nn - 1)];
#pragma unroll
          for (int k = 0; k < kUnroll; ++k) {
            const float rv0 = o[k].x - me.x, rv1 = o[k].y - me.y, rv2 = o[k].z - me.z;
            const float r2 = (rv0 * rv0 + rv1 * rv1) + rv2 * rv2;  // exact-input FP32 test (kick.f90:525, 556)
            if (p + k < nn && 0.0f < r2 && r2 < L.pp_range2) {
              pair_force<R>(L, me, o[k], f1);
              const int i = p + k;
              if (gyz || (gx0 && i < n0) || (gx1 && i >= n0 + n1)) ++npg;
              else ++npp;
            }
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
  const int64_t nthr = (int64_t)nbox * nr * nr * nr * 64;
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
  const int nr = L.ntt + 2;
  const int64_t nthr = (int64_t)nbox * nr * nr * nr * 64;
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
