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

constexpr int PB = 4;               // physical PP cells per brick per dim
constexpr int RB = PB + 2;          // brick plus the neighbour ring
constexpr int NRC = RB * RB * RB;   // region cells
constexpr int CAP = 2560;           // LDS candidate capacity (float4)
constexpr int QMAX = 32;            // per-lane in-range list length

// Pair force of candidate o on target me (kick.f90:523-530, 554-560).
__device__ inline void pair_force(const PPLevel& L, const float4 me, const float4 o, float f1[3]) {
  const float rv0 = o.x - me.x, rv1 = o.y - me.y, rv2 = o.z - me.z;
  const float r2 = (rv0 * rv0 + rv1 * rv1) + rv2 * rv2;
  const float rmag = sqrtf(r2);
  const float fr = f_ra(rmag, L.app, L.c_app) - f_ra(rmag, L.pp_range, L.c_range);
  f1[0] = f1[0] + (rv0 / rmag) * fr;
  f1[1] = f1[1] + (rv1 / rmag) * fr;
  f1[2] = f1[2] + (rv2 / rmag) * fr;
}

__device__ inline void pp_update(const PPLevel& L, const float4 me, const float f1[3], float* vp, float& f2,
                                 float vm[3]) {
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

__device__ inline void pp_reduce_out(float f2, const float vm[3], unsigned long long npp, unsigned long long npg,
                                     float* red, unsigned long long* cnt_red, PPOut* out) {
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

struct BrickGeo {
  int ib, c0[3];
  int64_t cbase;
};
__device__ inline BrickGeo brick_of(const PPLevel& L, int64_t b) {
  const int nbr = (L.npgrid + PB - 1) / PB;
  BrickGeo g;
  const int bx = (int)(b % nbr);
  b /= nbr;
  const int by = (int)(b % nbr);
  b /= nbr;
  const int bz = (int)(b % nbr);
  g.ib = (int)(b / nbr);
  const int m = L.npgrid + 2;
  g.cbase = (int64_t)g.ib * m * m * m;
  g.c0[0] = 1 + bx * PB;
  g.c0[1] = 1 + by * PB;
  g.c0[2] = 1 + bz * PB;
  return g;
}

// Main PP kernel: one workgroup per brick of PB^3 physical PP cells, whose
// particles plus the neighbour ring are staged in LDS.  Each wavefront owns
// 2x2x2-cell sub-bricks; its lanes hold the targets and walk the candidates
// of the surrounding 4x4x4 cells in lockstep (broadcast LDS reads, the cheap
// r^2 test), appending in-range candidates to a per-lane list.  The lists are
// evaluated when one fills and at the end, so the force law runs on nearly
// full wavefronts.  Candidates two cells away never pass the test (a PP cell
// is wider than r_pp), so the pair set is the reference's.  Bricks whose
// region does not fit in LDS are queued for k_pp_force_global.
__global__ void __launch_bounds__(256) k_pp_force(PPLevel L, const int32_t* __restrict__ count,
                                                  const int64_t* __restrict__ start,
                                                  const float4* __restrict__ ent, float* __restrict__ vp,
                                                  PPOut* __restrict__ out, int* __restrict__ ovf) {
  __shared__ float4 cand[CAP];
  __shared__ unsigned short qlist[QMAX][256];
  __shared__ int lst[NRC + 1];
  __shared__ int cnt[NRC];
  __shared__ int64_t gst[NRC];
  __shared__ float red[4];
  __shared__ unsigned long long cnt_red[2];
  const BrickGeo B = brick_of(L, blockIdx.x);
  const int m = L.npgrid + 2;
  for (int r = threadIdx.x; r < NRC; r += blockDim.x) {
    const int gx = B.c0[0] - 1 + r % RB, gy = B.c0[1] - 1 + (r / RB) % RB, gz = B.c0[2] - 1 + r / (RB * RB);
    const bool ok = gx <= L.npgrid + 1 && gy <= L.npgrid + 1 && gz <= L.npgrid + 1;
    const int64_t gc = B.cbase + ((int64_t)gz * m + gy) * m + gx;
    cnt[r] = ok ? count[gc] : 0;
    gst[r] = ok ? start[gc] : 0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int r = 0; r < NRC; ++r) {
      lst[r] = s;
      s += cnt[r];
    }
    lst[NRC] = s;
    cnt_red[0] = 0;
    cnt_red[1] = 0;
  }
  __syncthreads();
  const int total = lst[NRC];
  if (total > CAP) {  // uniform: hand the brick to the global-memory kernel
    if (threadIdx.x == 0) {
      const int k = atomicAdd(&ovf[0], 1);
      ovf[1 + k] = (int)blockIdx.x;
    }
    return;
  }
  for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
    int lo = 0, hi = NRC - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (lst[mid] <= idx) lo = mid;
      else hi = mid - 1;
    }
    cand[idx] = ent[gst[lo] + (idx - lst[lo])];
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  float f2 = 0, vm[3] = {0, 0, 0};
  unsigned long long npp = 0, npg = 0;
  unsigned short* myq = &qlist[0][threadIdx.x];
  for (int sb = wave; sb < 8; sb += 4) {
    const int sx = 2 * (sb & 1), sy = 2 * ((sb >> 1) & 1), sz = 2 * (sb >> 2);
    // targets: physical cells of the 2x2x2 sub-brick (region coords +1)
    int tcnt[8], tbase[8], ntg = 0;
    for (int q = 0; q < 8; ++q) {
      const int qx = sx + (q & 1), qy = sy + ((q >> 1) & 1), qz = sz + (q >> 2);
      const bool phys = B.c0[0] + qx <= L.npgrid && B.c0[1] + qy <= L.npgrid && B.c0[2] + qz <= L.npgrid;
      const int r = ((qz + 1) * RB + (qy + 1)) * RB + (qx + 1);
      tcnt[q] = phys ? cnt[r] : 0;
      tbase[q] = lst[r];
      ntg += tcnt[q];
    }
    for (int t0 = 0; t0 < ntg; t0 += 64) {
      const int t = t0 + lane;
      const bool active = t < ntg;
      float4 me = make_float4(0.f, 0.f, 0.f, 0.f);
      if (active) {
        int k = t, q = 0;
        while (k >= tcnt[q]) {
          k -= tcnt[q];
          ++q;
        }
        me = cand[tbase[q] + k];
      }
      float f1[3] = {0, 0, 0};
      int nq = 0;
      // the 4x4x4 candidate cells around the sub-brick (region coords sx..sx+3)
      for (int c = 0; c < 64; ++c) {
        const int rx = sx + (c & 3), ry = sy + ((c >> 2) & 3), rz = sz + (c >> 4);
        const int r = (rz * RB + ry) * RB + rx;
        const int gx = B.c0[0] - 1 + rx, gy = B.c0[1] - 1 + ry, gz = B.c0[2] - 1 + rz;
        const bool ghost = gx == 0 || gy == 0 || gz == 0 || gx == L.npgrid + 1 || gy == L.npgrid + 1 ||
                           gz == L.npgrid + 1;
        const int i0 = lst[r], i1 = lst[r + 1];
#pragma unroll 4
        for (int i = i0; i < i1; ++i) {
          const float4 o = cand[i];  // same address in every lane: broadcast
          const float rv0 = o.x - me.x, rv1 = o.y - me.y, rv2 = o.z - me.z;
          const float r2 = (rv0 * rv0 + rv1 * rv1) + rv2 * rv2;
          if (active && 0.0f < r2 && r2 < L.pp_range2) {
            myq[nq * 256] = (unsigned short)i;
            ++nq;
            if (ghost) ++npg;
            else ++npp;
          }
          if (__any(nq == QMAX)) {  // wave-uniform flush
#pragma unroll 1
            for (int j = 0; j < nq; ++j) pair_force(L, me, cand[myq[j * 256]], f1);
            nq = 0;
          }
        }
      }
#pragma unroll 1
      for (int j = 0; j < nq; ++j) pair_force(L, me, cand[myq[j * 256]], f1);
      if (active) pp_update(L, me, f1, vp, f2, vm);
    }
  }
  pp_reduce_out(f2, vm, npp, npg, red, cnt_red, out);
}

// Bricks whose particles do not fit in LDS: per-target loop over the 27
// neighbouring cells straight from global memory.
__global__ void __launch_bounds__(256) k_pp_force_global(PPLevel L, const int32_t* __restrict__ count,
                                                         const int64_t* __restrict__ start,
                                                         const float4* __restrict__ ent, float* __restrict__ vp,
                                                         PPOut* __restrict__ out, const int* __restrict__ ovf) {
  __shared__ int pre[PB * PB * PB + 1];
  __shared__ float red[4];
  __shared__ unsigned long long cnt_red[2];
  if ((int)blockIdx.x >= ovf[0]) return;
  const BrickGeo B = brick_of(L, ovf[1 + blockIdx.x]);
  const int m = L.npgrid + 2;
  if (threadIdx.x == 0) {
    int s = 0;
    for (int q = 0; q < PB * PB * PB; ++q) {
      pre[q] = s;
      const int cx = B.c0[0] + q % PB, cy = B.c0[1] + (q / PB) % PB, cz = B.c0[2] + q / (PB * PB);
      if (cx <= L.npgrid && cy <= L.npgrid && cz <= L.npgrid) s += count[B.cbase + ((int64_t)cz * m + cy) * m + cx];
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
    const int cx = B.c0[0] + q % PB, cy = B.c0[1] + (q / PB) % PB, cz = B.c0[2] + q / (PB * PB);
    const int64_t tcell = B.cbase + ((int64_t)cz * m + cy) * m + cx;
    const float4 me = ent[start[tcell] + (t - pre[q])];
    float f1[3] = {0, 0, 0};
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int nx = cx + dx, ny = cy + dy, nz = cz + dz;
          const bool ghost = nx == 0 || ny == 0 || nz == 0 || nx == L.npgrid + 1 || ny == L.npgrid + 1 ||
                             nz == L.npgrid + 1;
          const int64_t nc = B.cbase + ((int64_t)nz * m + ny) * m + nx;
          const int64_t s0 = start[nc];
          const int nn = count[nc];
          for (int p = 0; p < nn; ++p) {
            const float4 o = ent[s0 + p];
            const float rv0 = o.x - me.x, rv1 = o.y - me.y, rv2 = o.z - me.z;
            const float r2 = (rv0 * rv0 + rv1 * rv1) + rv2 * rv2;
            if (0.0f < r2 && r2 < L.pp_range2) {
              pair_force(L, me, o, f1);
              if (ghost) ++npg;
              else ++npp;
            }
          }
        }
    pp_update(L, me, f1, vp, f2, vm);
  }
  pp_reduce_out(f2, vm, npp, npg, red, cnt_red, out);
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
  (void)boxes;
  const int nbr = (L.npgrid + PB - 1) / PB;
  const int64_t grid = (int64_t)nbr * nbr * nbr * nbox;
  zero_async(S.ovf, sizeof(int), s);
  hipLaunchKernelGGL(k_pp_force, dim3((unsigned)grid), dim3(256), 0, s, L, S.count, S.start, S.entries, vp, out,
                     S.ovf);
  CUBE_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_pp_force_global, dim3((unsigned)grid), dim3(256), 0, s, L, S.count, S.start, S.entries, vp,
                     out, S.ovf);
  CUBE_HIP(hipGetLastError());
}

int64_t pp_bricks(const PPLevel& L, int nbox) {
  const int64_t nbr = (L.npgrid + PB - 1) / PB;
  return nbr * nbr * nbr * nbox;
}

}  // namespace cube
