// pm.hip -- TSC deposit, gradient + interpolation + kick, Green's functions.
//
// Deposit: one workgroup per BD^3 block of a box mesh.  It reads the
// particles of every coarse cell whose TSC footprint reaches the block
// (particles are cell-sorted, so these are contiguous runs), one wavefront
// per coarse cell, and accumulates (wx*wy)*wz into LDS.  The FP32 form sums
// in 2^-32 fixed point with 64-bit integer LDS atomics: each FP32 weight
// (< 0.43) becomes one 32-bit integer (v_cvt_u32_f32 after rounding), integer
// addition is associative, so the result does not depend on thread order and
// is more accurate than an FP32 running sum.  The FP64 form uses FP64 LDS
// atomics.
//
// Interpolation: one workgroup per BD^3 block of force cells.  The block
// stages phi (block + 3-cell halo) in LDS, evaluates the 4-point FD force
// (particle_mesh.f90:237-242) for block + 2 cells, takes the f2max maximum
// over its physical cells, and applies the 27-point TSC gather and kick to
// the particles whose floor(xpos) lies in the block.  The force field is
// never written to HBM.
#include <cmath>
#include <type_traits>

#include "pm.h"
#include "timing.h"

namespace cube {
namespace {

constexpr float kFixF = 4294967296.0f;  // 2^32: weights are < 0.43, so w * 2^32 fits a u32

// FP32 TSC weight product as a 2^-32 fixed-point integer (round to nearest).
__device__ inline unsigned long long fix32(float cw) { return (unsigned long long)__float2uint_rn(cw * kFixF); }

__device__ inline int floor_div(int a, int b) {
  int q = a / b;
  if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
  return q;
}

template <int IZ>
using XT = typename XP<IZ>::T;


// Exclusive block scan of cnt[0..n) into pre[0..n] (pre[n] = total), n <= 1024.
__device__ inline void block_scan(const int* cnt, int* pre, int n, int* wsum) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int b0 = 4 * threadIdx.x;
  int loc[4], s = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    loc[k] = (b0 + k < n) ? cnt[b0 + k] : 0;
    s += loc[k];
  }
  int incl = s;
  for (int o = 1; o < 64; o <<= 1) {
    const int v = __shfl_up(incl, o, 64);
    if (lane >= o) incl += v;
  }
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  int wbase = 0;
  for (int w = 0; w < wave; ++w) wbase += wsum[w];
  int run = wbase + incl - s;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (b0 + k < n) pre[b0 + k] = run;
    run += loc[k];
  }
  if (threadIdx.x == blockDim.x - 1) {
    int tot = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) tot += wsum[w];
    pre[n] = tot;
  }
  __syncthreads();
}

// Counts and offsets of the coarse cells [clo, chi]^3 of a box, staged in
// LDS with one round of independent loads.  Returns the number of cells.
__device__ inline int stage_cells(const PMLevel& L, const PMBox& B, const int clo[3], const int chi[3],
                                  const int32_t* __restrict__ rhoc, const int64_t* __restrict__ off, int* cnt,
                                  int* pre, int64_t* gof, int* wsum) {
  const int ncx = max(0, chi[0] - clo[0] + 1), ncy = max(0, chi[1] - clo[1] + 1), ncz = max(0, chi[2] - clo[2] + 1);
  const int nC = ncx * ncy * ncz;
  for (int t = threadIdx.x; t < nC; t += blockDim.x) {
    const int ci = clo[0] + t % ncx, cj = clo[1] + (t / ncx) % ncy, ck = clo[2] + t / (ncx * ncy);
    const int64_t cell = rc_index(ci, cj, ck, B.T, L.nte, L.ncb);
    cnt[t] = rhoc[cell];
    gof[t] = off[cell];
  }
  __syncthreads();
  block_scan(cnt, pre, nC, wsum);
  return nC;
}

// largest c with pre[c] <= i
__device__ inline int find_cell(const int* pre, int nC, int i) {
  int lo = 0, hi = nC - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Position of a particle in the box's mesh units, particle_mesh.f90:36-37 /
// kick.f90:628-629: ((c-1) + frac) * ratio_cs / gridsize - shift in FP64.
// When gridsize is a power of two the product with ratio_cs/gridsize is the
// same double as the division.
__device__ inline double pm_xpos(const PMLevel& L, int c, double frac, int shift) {
  const double a = (double)(c - 1) + frac;
  return (L.xdiv ? (a * (double)L.rcs) / L.gridsize : a * L.xmul) - (double)shift;
}

// Deposits one particle's 27 TSC weights (floor f, weights w) into the
// block's LDS accumulator.
//
// The accumulator of a BD^3 block stores cell (x, y, z) at row (z, y) and
// column x ^ ((5y + 9z) mod BD).  The queued particles of one wavefront come
// from a few neighbouring coarse cells, so for one tap their cells differ
// mostly in y and z; unswizzled, the 64-bit LDS bank of a cell would depend
// on x alone and the lanes' atomics would pile onto a handful of banks.
template <int BD>
__device__ inline int acc_slot(int x, int y, int z) {
  return (z * BD + y) * BD + (x ^ ((5 * y + 9 * z) & (BD - 1)));
}

template <typename R, typename Acc, int BD>
__device__ inline void deposit_particle(Acc* acc, const int f[3], const R w0[3][3], const int mlo[3], const int mhi[3]) {
  // FP32: the z weights carry the 2^32 of the fixed point; scaling by a power
  // of two is exact, so (wx*wy)*(wz*2^32) == ((wx*wy)*wz)*2^32 bit for bit
  R w[3][3];
  for (int d = 0; d < 3; ++d)
    for (int i = 0; i < 3; ++i)
      w[d][i] = (std::is_same<R, float>::value && d == 2) ? w0[d][i] * (R)kFixF : w0[d][i];
  const bool inner = f[0] >= mlo[0] && f[0] + 2 <= mhi[0] && f[1] >= mlo[1] && f[1] + 2 <= mhi[1] &&
                     f[2] >= mlo[2] && f[2] + 2 <= mhi[2];
  for (int i3 = 0; i3 < 3; ++i3) {
    const int mz = f[2] + i3;
    if (!inner && (mz < mlo[2] || mz > mhi[2])) continue;
    for (int i2 = 0; i2 < 3; ++i2) {
      const int my = f[1] + i2;
      if (!inner && (my < mlo[1] || my > mhi[1])) continue;
      const R wyz = w[1][i2];
      const int ly = my - mlo[1], lz = mz - mlo[2];
      const int row = (lz * BD + ly) * BD, sw = (5 * ly + 9 * lz) & (BD - 1);
      for (int i1 = 0; i1 < 3; ++i1) {
        const int mx = f[0] + i1;
        if (!inner && (mx < mlo[0] || mx > mhi[0])) continue;
        const R cw = (w[0][i1] * wyz) * w[2][i3];  // dx(1,i1)*dx(2,i2)*dx(3,i3)
        const int slot = row + ((mx - mlo[0]) ^ sw);
        if constexpr (std::is_same<R, float>::value)
          atomicAdd(&acc[slot], (unsigned long long)__float2uint_rn(cw));
        else atomicAdd(&acc[slot], cw);
      }
    }
  }
}

__device__ inline void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Fine-level deposit (tile and sub-tile meshes), gather form: the block owns
// BD^3 mesh cells and reads every coarse cell whose particles can reach them.
// Only a thin layer of a ring cell reaches the block (1/w of its particles per
// axis), so each wavefront loads the compressed coordinates of four cells'
// particles at a time (twelve loads in flight per lane), tests the floor
// index (integer arithmetic on the coordinate when ratio_cs/gridsize is a
// power of two), appends hits with their coordinates to a per-wavefront LDS
// queue, and deposits 64 queued particles at a time with every lane busy and
// no memory access but LDS atomics.
template <typename R, int IZ, int BD, bool CA>
__global__ void __launch_bounds__(256) k_deposit(PMLevel L, const PMBox* __restrict__ boxes,
                                                 const XT<IZ>* __restrict__ xp, const int32_t* __restrict__ rhoc,
                                                 const int64_t* __restrict__ off, R* __restrict__ mesh, float mass) {
  using Acc = typename std::conditional<std::is_same<R, float>::value, unsigned long long, double>::type;
  constexpr int kMaxFine = 216;  // (16/4 + 2)^3: fine levels have w >= 4 (checked on the host)
  constexpr int G = 4;           // cells loaded together per wavefront
  __shared__ Acc acc[BD * BD * BD];
  __shared__ int cnt[kMaxFine], pre[kMaxFine + 1], wsum[4];
  __shared__ int64_t gof[kMaxFine];
  __shared__ unsigned long long queue[4][128];
  // CA: blocks aligned to groups of BD/w coarse cells, counted down from the
  // last deposited cell nc2 + nloop (their regions then end past win_hi and
  // begin before win_lo); the block reads its own cells and the lower
  // neighbours only.  Otherwise blocks tile the window from win_lo.
  const int gcell = BD / L.w;
  const int nblk = CA ? 1 + (L.nforce / L.w - 1 + 2 * L.nloop + gcell - 1) / gcell : (L.win_hi - L.win_lo + BD) / BD;
  int64_t b = blockIdx.x;
  const int bx = (int)(b % nblk);
  b /= nblk;
  const int by = (int)(b % nblk);
  b /= nblk;
  const int bz = (int)(b % nblk);
  const int ibox = (int)(b / nblk);
  const PMBox B = boxes[ibox];
  const int bidx[3] = {bx, by, bz};
  int mlo[3], mhi[3], clo[3], chi[3];
  for (int d = 0; d < 3; ++d) {
    int m0 = L.win_lo + bidx[d] * BD;
    if (CA) m0 = L.w * (B.nc2[d] + L.nloop - (nblk - 1 - bidx[d]) * gcell - 1) - B.ushift[d];
    mlo[d] = max(m0, L.win_lo);
    mhi[d] = min(m0 + BD - 1, L.win_hi);
    // particle of coarse cell i reaches mesh cells [w(i-1)-s, wi-s+1]
    clo[d] = max(-floor_div(-(mlo[d] - 1 + B.ushift[d]), L.w), B.nc1[d] - L.nloop);
    chi[d] = min(floor_div(mhi[d] + B.ushift[d], L.w) + 1, B.nc2[d] + L.nloop);
  }
  for (int t = threadIdx.x; t < BD * BD * BD; t += blockDim.x) acc[t] = 0;
  const int nC = stage_cells(L, B, clo, chi, rhoc, off, cnt, pre, gof, wsum);
  const int ncx = max(0, chi[0] - clo[0] + 1), ncy = max(0, chi[1] - clo[1] + 1);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  unsigned long long* q = queue[wave];
  // the floor index as integer arithmetic: shift = B - log2(w)
  const int sh = L.xdiv ? -1 : XP<IZ>::B - (31 - __clz(L.w));
  const R rsh = sh >= 0 ? (R)1 / (R)(1u << sh) : (R)0;
  constexpr double kRes = 1.0 / (double)(1u << XP<IZ>::B);
  auto process = [&](int nq) {
    wave_sync_lds();
    if (lane < nq) {
      const unsigned long long e = q[lane];
      const int c = (int)(e >> 48);  // staged cell as 5-bit (x, y, z) offsets from clo
      const int ijk[3] = {clo[0] + (c & 31), clo[1] + ((c >> 5) & 31), clo[2] + ((c >> 10) & 31)};
      int f[3];
      R wt[3][3];
      for (int d = 0; d < 3; ++d) {
        const uint32_t u = (uint32_t)(e >> (16 * d)) & 0xFFFFu;
        if (sh >= 0) {
          // xpos = w(c-1) - s + (u + 1/2) 2^-sh exactly: floor and fraction
          // from the integer coordinate, the fraction exact in R
          f[d] = L.w * (ijk[d] - 1) - B.ushift[d] + (int)(u >> sh);
          tsc_weights_frac<R>(((R)(u & ((1u << sh) - 1u)) + (R)0.5) * rsh, wt[d]);
        } else {
          f[d] = tsc_weights<R>(pm_xpos(L, ijk[d], ((double)u + 0.5) * kRes, B.ushift[d]), wt[d]);
        }
      }
      deposit_particle<R, Acc, BD>(acc, f, wt, mlo, mhi);
    }
  };
  int qn = 0;  // wave-uniform queue length
  for (int c0 = wave; c0 < nC; c0 += G * nw) {
    int nmax = 0;
    for (int k = 0; k < G; ++k) {
      const int c = c0 + k * nw;
      if (c < nC) nmax = max(nmax, cnt[c]);
    }
    for (int l0 = 0; l0 < nmax; l0 += 64) {
      const int l = l0 + lane;
      uint32_t u[G][3];
      bool valid[G];
#pragma unroll
      for (int k = 0; k < G; ++k) {
        const int c = c0 + k * nw;
        valid[k] = c < nC && l < cnt[c];
        const int64_t ip = valid[k] ? gof[c] + l : 0;
        for (int d = 0; d < 3; ++d) u[k][d] = valid[k] ? XP<IZ>::raw(xp, ip, d) : 0u;
      }
#pragma unroll
      for (int k = 0; k < G; ++k) {
        const int c = c0 + k * nw;  // wave-uniform
        const int cx = c % ncx, cy = (c / ncx) % ncy, cz = c / (ncx * ncy);
        bool hit = valid[k];
        if (hit) {
          const int ijk[3] = {clo[0] + cx, clo[1] + cy, clo[2] + cz};
          for (int d = 0; d < 3; ++d) {
            int f;
            if (sh >= 0) f = L.w * (ijk[d] - 1) - B.ushift[d] + (int)(u[k][d] >> sh);
            else f = (int)floor(pm_xpos(L, ijk[d], ((double)u[k][d] + 0.5) * kRes, B.ushift[d]));
            hit = hit && f + 2 >= mlo[d] && f <= mhi[d];
          }
        }
        const unsigned long long m = __ballot(hit);
        if (hit) {
          const int pos = qn + (int)__popcll(m & ((1ull << lane) - 1ull));
          q[pos] = ((unsigned long long)(cx | (cy << 5) | (cz << 10)) << 48) | ((unsigned long long)u[k][2] << 32) |
                   ((unsigned long long)u[k][1] << 16) | (unsigned long long)u[k][0];
        }
        qn += (int)__popcll(m);
        if (qn >= 64) {
          process(64);
          // move the overflow (< 64 entries) to the front
          wave_sync_lds();
          const unsigned long long mv = (lane < qn - 64) ? q[64 + lane] : 0ull;
          wave_sync_lds();
          if (lane < qn - 64) q[lane] = mv;
          qn -= 64;
        }
      }
    }
  }
  if (qn > 0) process(qn);
  __syncthreads();
  for (int t = threadIdx.x; t < BD * BD * BD; t += blockDim.x) {
    const int lx = t % BD, ly = (t / BD) % BD, lz = t / (BD * BD);
    const int m[3] = {mlo[0] + lx, mlo[1] + ly, mlo[2] + lz};
    if (m[0] > mhi[0] || m[1] > mhi[1] || m[2] > mhi[2]) continue;
    const Acc av = acc[acc_slot<BD>(lx, ly, lz)];
    R v;
    if constexpr (std::is_same<R, float>::value) v = (float)((double)av * (1.0 / (double)kFixF));
    else v = av;
    int a[3];
    for (int d = 0; d < 3; ++d) a[d] = m[d] + B.org[d] - L.m0;
    mesh[B.moff + ((int64_t)a[2] * L.n + a[1]) * L.n + a[0]] = v * (R)mass;
  }
}

// Binned deposit, pass 1/2: one wavefront per (box, coarse cell of the box's
// deposit range).  A coarse cell's floors and taps span at most two 16^3
// blocks per axis (w <= 8), so each particle belongs to a subset of the 8
// blocks b0 + {0,1}^3; per block, a ballot counts the members (FILL=0) or
// places them at start + cursor + rank (FILL=1).  The order inside a list is
// free: the deposit sums integers, so its result does not depend on it.
constexpr int kDBD = 16;  // deposit block edge in mesh cells

__host__ __device__ inline int dep_nblk(const PMLevel& L) { return (L.win_hi - L.win_lo + kDBD) / kDBD; }

template <int IZ, int FILL>
__global__ void __launch_bounds__(256) k_dep_bin(PMLevel L, const PMBox* __restrict__ boxes, int nbox,
                                                 const XT<IZ>* __restrict__ xp, const int32_t* __restrict__ rhoc,
                                                 const int64_t* __restrict__ off, int32_t* __restrict__ count,
                                                 const int64_t* __restrict__ start,
                                                 unsigned long long* __restrict__ ents) {
  const int nr = L.nforce / L.w + 2 * L.nloop;
  const int lane = threadIdx.x & 63;
  const int64_t wv = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (wv >= (int64_t)nbox * nr * nr * nr) return;
  const int ib = (int)(wv / (nr * nr * nr)), qc = (int)(wv % (nr * nr * nr));
  const PMBox B = boxes[ib];
  const int cl[3] = {qc % nr, (qc / nr) % nr, qc / (nr * nr)};
  int c[3], lo0[3], b0[3];
  for (int d = 0; d < 3; ++d) {
    c[d] = B.nc1[d] - L.nloop + cl[d];
    lo0[d] = L.w * (c[d] - 1) - B.ushift[d];  // the cell's lowest floor
    b0[d] = (max(lo0[d], L.win_lo) - L.win_lo) / kDBD;
  }
  const int64_t cell = rc_index(c[0], c[1], c[2], B.T, L.nte, L.ncb);
  const int cnt = rhoc[cell];
  const int64_t base = off[cell];
  const int nb = dep_nblk(L);
  const int sh = XP<IZ>::B - (31 - __clz(L.w));
  const int64_t bbase = (int64_t)ib * nb * nb * nb;
  // box-local cell as three 5-bit fields (nr <= 32), so the deposit decodes
  // it without divisions
  const unsigned long long ckey = (unsigned long long)(cl[0] | (cl[1] << 5) | (cl[2] << 10)) << 48;
  const unsigned long long lt = (1ull << lane) - 1ull;
  for (int l0 = 0; l0 < cnt; l0 += 64) {
    const int l = l0 + lane;
    uint32_t u[3] = {0, 0, 0};
    unsigned mem = 0;  // bit ox + 2 oy + 4 oz: member of block b0 + (ox, oy, oz)
    if (l < cnt) {
      int rlo[3], rhi[3];
      bool any = true;
      for (int d = 0; d < 3; ++d) {
        u[d] = XP<IZ>::raw(xp, base + l, d);
        const int f = lo0[d] + (int)(u[d] >> sh);
        const int tl = max(f, L.win_lo), th = min(f + 2, L.win_hi);
        any = any && tl <= th;
        rlo[d] = (tl - L.win_lo) / kDBD - b0[d];
        rhi[d] = (th - L.win_lo) / kDBD - b0[d];
      }
      if (any)
        for (int j = 0; j < 8; ++j) {
          const int o[3] = {j & 1, (j >> 1) & 1, j >> 2};
          bool in = true;
          for (int d = 0; d < 3; ++d) in = in && o[d] >= rlo[d] && o[d] <= rhi[d];
          if (in) mem |= 1u << j;
        }
    }
    const unsigned long long e = ckey | ((unsigned long long)u[2] << 32) | ((unsigned long long)u[1] << 16) | u[0];
    // lane j < 8 owns block b0 + o(j): its member mask, one atomic for the
    // wavefront's members and (FILL) the list position; the eight atomics
    // and start loads are independent, so they overlap
    unsigned long long mine = 0;
    for (int j = 0; j < 8; ++j) {
      const unsigned long long m = __ballot((mem >> j) & 1u);
      if (lane == j) mine = m;
    }
    int64_t where = 0, lim = 0;
    if (lane < 8 && mine != 0) {
      const int64_t blk =
          bbase + ((int64_t)(b0[2] + (lane >> 2)) * nb + (b0[1] + ((lane >> 1) & 1))) * nb + b0[0] + (lane & 1);
      const int pos = atomicAdd(&count[blk], (int)__popcll(mine));
      if (FILL) {
        where = start[blk] + pos;
        lim = start[blk + 1];  // end of the block's capacity
      }
    }
    if (FILL) {
      for (int j = 0; j < 8; ++j) {
        const int64_t wj = __shfl(where, j, 64), lj = __shfl(lim, j, 64);
        const unsigned long long mj = __shfl(mine, j, 64);
        const int64_t at = wj + (int64_t)__popcll(mj & lt);
        if (((mem >> j) & 1u) && at < lj) ents[at] = e;
      }
    }
  }
}

// List capacity of each block: the particles of every coarse cell whose taps
// can reach it (the cells a gather-form block would read), summed from the
// counts, so that no pass over the particles is needed to size the lists.
__global__ void k_dep_cap(PMLevel L, const PMBox* __restrict__ boxes, int nbox, const int32_t* __restrict__ rhoc,
                          int32_t* __restrict__ cap) {
  const int nb = dep_nblk(L);
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nbox * nb * nb * nb) return;
  int64_t b = t;
  const int bi[3] = {(int)(b % nb), (int)((b / nb) % nb), (int)((b / ((int64_t)nb * nb)) % nb)};
  const PMBox B = boxes[b / ((int64_t)nb * nb * nb)];
  int clo[3], chi[3];
  for (int d = 0; d < 3; ++d) {
    const int mlo = L.win_lo + bi[d] * kDBD, mhi = min(mlo + kDBD - 1, L.win_hi);
    clo[d] = max(-floor_div(-(mlo - 1 + B.ushift[d]), L.w), B.nc1[d] - L.nloop);
    chi[d] = min(floor_div(mhi + B.ushift[d], L.w) + 1, B.nc2[d] + L.nloop);
  }
  int sum = 0;
  for (int k = clo[2]; k <= chi[2]; ++k)
    for (int j = clo[1]; j <= chi[1]; ++j)
      for (int i = clo[0]; i <= chi[0]; ++i) sum += rhoc[rc_index(i, j, k, B.T, L.nte, L.ncb)];
  cap[t] = sum;
}

// Binned deposit, pass 2/2: one workgroup per 16^3 block of one box, the
// same regions as k_deposit's mesh-aligned blocks; each thread deposits one
// listed particle at a time (taps outside the block skipped), then the block
// writes its mesh cells.
// NT threads share a block's 32 KB accumulator: more wavefronts per CU cover
// the LDS atomics' latency (512 threads measured 11.9 -> 10.9 ms against 256
// for the 512^3 PM3 deposit); NT in {256, 512, 1024} is picked by the
// engine's autotuner (integer sums: every NT gives the same mesh).
template <typename R, int IZ, int NT>
__global__ void __launch_bounds__(NT) k_deposit_lists(PMLevel L, const PMBox* __restrict__ boxes,
                                                       const int64_t* __restrict__ start,
                                                       const int32_t* __restrict__ used,
                                                       const unsigned long long* __restrict__ ents,
                                                       R* __restrict__ mesh, float mass) {
  using Acc = typename std::conditional<std::is_same<R, float>::value, unsigned long long, double>::type;
  constexpr int BD = kDBD;
  __shared__ Acc acc[BD * BD * BD];
  const int nb = dep_nblk(L);
  const int64_t blk = blockIdx.x;
  int64_t b = blk;
  const int bx = (int)(b % nb);
  b /= nb;
  const int by = (int)(b % nb);
  b /= nb;
  const int bz = (int)(b % nb);
  const int ibox = (int)(b / nb);
  const PMBox B = boxes[ibox];
  const int bidx[3] = {bx, by, bz};
  int mlo[3], mhi[3];
  for (int d = 0; d < 3; ++d) {
    mlo[d] = L.win_lo + bidx[d] * BD;
    mhi[d] = min(mlo[d] + BD - 1, L.win_hi);
  }
  for (int t = threadIdx.x; t < BD * BD * BD; t += blockDim.x) acc[t] = 0;
  __syncthreads();
  const int sh = XP<IZ>::B - (31 - __clz(L.w));
  const R rsh = (R)1 / (R)(1u << sh);
  const int64_t e0 = start[blk], e1 = e0 + used[blk];
  for (int64_t i = e0 + threadIdx.x; i < e1; i += blockDim.x) {
    const unsigned long long e = ents[i];
    const int qc = (int)(e >> 48);
    const int cl[3] = {qc & 31, (qc >> 5) & 31, (qc >> 10) & 31};
    int f[3];
    R wt[3][3];
    for (int d = 0; d < 3; ++d) {
      const uint32_t u = (uint32_t)(e >> (16 * d)) & 0xFFFFu;
      const int c = B.nc1[d] - L.nloop + cl[d];
      f[d] = L.w * (c - 1) - B.ushift[d] + (int)(u >> sh);
      tsc_weights_frac<R>(((R)(u & ((1u << sh) - 1u)) + (R)0.5) * rsh, wt[d]);
    }
    deposit_particle<R, Acc, BD>(acc, f, wt, mlo, mhi);
  }
  __syncthreads();
  for (int t = threadIdx.x; t < BD * BD * BD; t += blockDim.x) {
    const int lx = t % BD, ly = (t / BD) % BD, lz = t / (BD * BD);
    const int m[3] = {mlo[0] + lx, mlo[1] + ly, mlo[2] + lz};
    if (m[0] > mhi[0] || m[1] > mhi[1] || m[2] > mhi[2]) continue;
    const Acc av = acc[acc_slot<BD>(lx, ly, lz)];
    R v;
    if constexpr (std::is_same<R, float>::value) v = (float)((double)av * (1.0 / (double)kFixF));
    else v = av;
    int a[3];
    for (int d = 0; d < 3; ++d) a[d] = m[d] + B.org[d] - L.m0;
    mesh[B.moff + ((int64_t)a[2] * L.n + a[1]) * L.n + a[0]] = v * (R)mass;
  }
}

// Coarse-level deposit (PM1), step 1.  Every particle of coarse cell c has
// floor(xpos) = c-1 and deposits into the same 27 mesh cells c-1..c+1
// (particle_mesh.f90:56-58), so one wavefront sums the 27 tap weights of a
// cell over its particles (exact fixed point, or FP64), transposes the lane
// partials through LDS and writes the 27 sums: stencil[tap][cell].  Each
// particle is read once.
template <typename R, int IZ>
__global__ void __launch_bounds__(128) k_pm1_stencil(PMLevel L, const XT<IZ>* __restrict__ xp,
                                                     const int32_t* __restrict__ rhoc,
                                                     const int64_t* __restrict__ off, void* __restrict__ stencil) {
  using Acc = typename std::conditional<std::is_same<R, float>::value, unsigned long long, double>::type;
  // the 27 tap sums leave through LDS in two rounds of at most 14 taps: half
  // the LDS of one round, so the register limit (five wavefronts per SIMD)
  // rather than LDS sets the occupancy
  constexpr int kRound = 14;
  __shared__ Acc red[2][kRound][65];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t ncell = (int64_t)L.n * L.n * L.n;
  const int64_t g = (int64_t)blockIdx.x * 2 + wave;
  if (g >= ncell) return;
  const int nt = L.nt, nnt = L.n / L.nt;
  const int gx = (int)(g % L.n), gy = (int)((g / L.n) % L.n), gz = (int)(g / ((int64_t)L.n * L.n));
  const int T = ((gz / nt) * nnt + gy / nt) * nnt + gx / nt;
  const int ijk[3] = {gx % nt + 1, gy % nt + 1, gz % nt + 1};
  const int64_t cell = rc_index(ijk[0], ijk[1], ijk[2], T, L.nte, L.ncb);
  const int n = rhoc[cell];
  const int64_t base = off[cell];
  Acc a27[27];
#pragma unroll
  for (int q = 0; q < 27; ++q) a27[q] = 0;
  // one mesh cell per coarse cell: xpos - floor(xpos) = (u + 1/2) 2^-B exactly
  const bool unit = !L.xdiv && L.xmul == 1.0;
  constexpr R kRes = (R)1 / (R)(1u << XP<IZ>::B);
  for (int l = lane; l < n; l += 64) {
    R w[3][3];
    for (int d = 0; d < 3; ++d) {
      if (unit) tsc_weights_frac<R>(((R)XP<IZ>::raw(xp, base + l, d) + (R)0.5) * kRes, w[d]);
      else tsc_weights<R>(pm_xpos(L, ijk[d], XP<IZ>::off(xp, base + l, d), 0), w[d]);
    }
#pragma unroll
    for (int q = 0; q < 27; ++q) {
      const R cw = (w[0][q % 3] * w[1][(q / 3) % 3]) * w[2][q / 9];
      if constexpr (std::is_same<R, float>::value) a27[q] += fix32(cw);
      else a27[q] += cw;
    }
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int q0 = h * kRound, nq = h ? 27 - kRound : kRound;
#pragma unroll
    for (int q = 0; q < 27; ++q)
      if (q >= q0 && q < q0 + nq) red[wave][q - q0][lane] = a27[q];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // 4 lanes per tap, each summing 16 lane partials, then the four parts
    // combined through two shuffles: every sum in a fixed order
    const int tq = lane % kRound, part = lane / kRound;
    Acc s = 0;
    if (part < 4 && tq < nq)
      for (int l = 16 * part; l < 16 * part + 16; ++l) s += red[wave][tq][l];
    s += __shfl(s, (lane + 2 * kRound) & 63, 64);
    s += __shfl(s, (lane + kRound) & 63, 64);
    if (lane < nq) ((Acc*)stencil)[(int64_t)(q0 + lane) * ncell + g] = s;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// Coarse-level deposit, step 2: mesh cell M (global, periodic) is the sum of
// tap t of the cells M+1-t, t in {0,1,2}^3, in a fixed order; then the
// reference's rho1 = rho * mass_p_cdm (particle_mesh.f90:102-105) and the
// exact testrho sum (particle_mesh.f90:106).
template <typename R>
__global__ void __launch_bounds__(256) k_pm1_gather(PMLevel L, const void* __restrict__ stencil, R* __restrict__ mesh,
                                                    float mass, PMOut* __restrict__ out) {
  using Acc = typename std::conditional<std::is_same<R, float>::value, unsigned long long, double>::type;
  __shared__ Acc red[4];
  const int n = L.n;
  const int64_t ncell = (int64_t)n * n * n;
  const int64_t M = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  Acc v = 0;
  if (M < ncell) {
    const int mx = (int)(M % n), my = (int)((M / n) % n), mz = (int)(M / ((int64_t)n * n));
    const Acc* st = (const Acc*)stencil;
#pragma unroll
    for (int q = 0; q < 27; ++q) {
      const int tx = q % 3, ty = (q / 3) % 3, tz = q / 9;
      int cx = mx + 1 - tx, cy = my + 1 - ty, cz = mz + 1 - tz;
      cx = cx < 0 ? cx + n : (cx >= n ? cx - n : cx);
      cy = cy < 0 ? cy + n : (cy >= n ? cy - n : cy);
      cz = cz < 0 ? cz + n : (cz >= n ? cz - n : cz);
      v += st[(int64_t)q * ncell + ((int64_t)cz * n + cy) * n + cx];
    }
    R r;
    if constexpr (std::is_same<R, float>::value) r = (float)((double)v * (1.0 / (double)kFixF));
    else r = v;
    mesh[M] = r * (R)mass;
  }
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    Acc s = 0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) s += red[q];
    if constexpr (std::is_same<R, float>::value) atomicAdd(&out->testrho_fx[blockIdx.x % kSumSlots], s);
    else atomicAdd(&out->testrho_d[blockIdx.x % kSumSlots], s);
  }
}

template <typename R>
__device__ inline R block_max(R v, R* red) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_down(v, o, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  R m = 0;
  for (int q = 0; q < (int)(blockDim.x >> 6); ++q) m = max(m, red[q]);
  return m;
}

template <typename R, int IZ, int BD, int MAXC>
__global__ void __launch_bounds__(256) k_interp(PMLevel L, const PMBox* __restrict__ boxes,
                                                const XT<IZ>* __restrict__ xp, float* __restrict__ vp,
                                                const int32_t* __restrict__ rhoc, const int64_t* __restrict__ off,
                                                const R* __restrict__ phi, double kc,
                                                PMOut* __restrict__ out) {
  constexpr int FS = BD + 2, PS = BD + 6;
  constexpr int NF = (FS * FS * FS + 255) / 256;
  constexpr int LDSN = (PS * PS * PS > 3 * FS * FS * FS) ? PS * PS * PS : 3 * FS * FS * FS;
  __shared__ R lds[LDSN];
  __shared__ R red[4];
  __shared__ int cnt[MAXC], pre[MAXC + 1], wsum[4];
  __shared__ int64_t gof[MAXC];
  if (L.kdev) kc = *L.kdev;
  const int nblk = (L.nforce + BD - 1) / BD;
  int64_t b = blockIdx.x;
  const int bx = (int)(b % nblk);
  b /= nblk;
  const int by = (int)(b % nblk);
  b /= nblk;
  const int bz = (int)(b % nblk);
  const int ibox = (int)(b / nblk);
  const PMBox B = boxes[ibox];
  const int f0[3] = {bx * BD, by * BD, bz * BD};
  // coarse cells whose particles can have floor(xpos) in this block; a cell i
  // covers floor in [w(i-1)-s, wi-s-1]
  int clo[3], chi[3];
  for (int d = 0; d < 3; ++d) {
    clo[d] = max(B.nc1[d], floor_div(f0[d] + B.ushift[d], L.w) + 1);
    chi[d] = min(B.nc2[d], floor_div(f0[d] + BD - 1 + B.ushift[d], L.w) + 1);
  }
  // 1. phi(f0-2 .. f0+BD+3) and the cell metadata: every load is issued
  //    before the first LDS store, so the block waits for one round trip.
  constexpr int NPH = (PS * PS * PS + 255) / 256;
  R ph[NPH];
#pragma unroll
  for (int q = 0; q < NPH; ++q) {
    const int t = threadIdx.x + q * 256;
    ph[q] = (R)0;
    if (t < PS * PS * PS) {
      const int l[3] = {t % PS, (t / PS) % PS, t / (PS * PS)};
      int a[3];
      bool ok = true;
      for (int d = 0; d < 3; ++d) {
        a[d] = f0[d] - 2 + l[d] + B.org[d] - L.m0;
        if (L.periodic) a[d] = a[d] < 0 ? a[d] + L.n : (a[d] >= L.n ? a[d] - L.n : a[d]);
        else ok = ok && a[d] >= 0 && a[d] < L.n;
      }
      if (ok) ph[q] = phi[B.moff + ((int64_t)a[2] * L.n + a[1]) * L.n + a[0]];
    }
  }
  const int ncx = max(0, chi[0] - clo[0] + 1), ncy = max(0, chi[1] - clo[1] + 1),
            ncz = max(0, chi[2] - clo[2] + 1);
  const int nC = ncx * ncy * ncz;
  for (int t = threadIdx.x; t < nC; t += blockDim.x) {
    const int64_t cell = rc_index(clo[0] + t % ncx, clo[1] + (t / ncx) % ncy, clo[2] + t / (ncx * ncy), B.T, L.nte,
                                  L.ncb);
    cnt[t] = rhoc[cell];
    gof[t] = off[cell];
  }
#pragma unroll
  for (int q = 0; q < NPH; ++q) {
    const int t = threadIdx.x + q * 256;
    if (t < PS * PS * PS) lds[t] = ph[q];
  }
  __syncthreads();
  block_scan(cnt, pre, nC, wsum);
  // first two particles of this thread: coordinates and velocities loaded
  // now, consumed after the force computation
  const int total = pre[nC];
  int pc[2];
  uint32_t pu[2][3];
  float pv[2][3];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int i = threadIdx.x + u * 256;
    pc[u] = -1;
    if (i < total) {
      pc[u] = find_cell(pre, nC, i);
      const int64_t ip = gof[pc[u]] + (i - pre[pc[u]]);
      for (int d = 0; d < 3; ++d) {
        pu[u][d] = XP<IZ>::raw(xp, ip, d);
        pv[u][d] = vp[3 * ip + d];
      }
    }
  }
  // 2. force(d, f0 .. f0+BD+1) = sum(phi(m + [-2,-1,1,2]) * weight)
  const FDWeights<R> W;
  R fr[NF][3];
  R f2 = 0;
#pragma unroll
  for (int q = 0; q < NF; ++q) {
    const int t = threadIdx.x + q * 256;
    if (t < FS * FS * FS) {
      const int lx = t % FS, ly = (t / FS) % FS, lz = t / (FS * FS);
      const int px = lx + 2, py = ly + 2, pz = lz + 2;
      auto P = [&](int x, int y, int z) { return lds[(z * PS + y) * PS + x]; };
      R sx = 0, sy = 0, sz = 0;
      sx = sx + P(px - 2, py, pz) * W.w0;
      sx = sx + P(px - 1, py, pz) * W.w1;
      sx = sx + P(px + 1, py, pz) * W.w2;
      sx = sx + P(px + 2, py, pz) * W.w3;
      sy = sy + P(px, py - 2, pz) * W.w0;
      sy = sy + P(px, py - 1, pz) * W.w1;
      sy = sy + P(px, py + 1, pz) * W.w2;
      sy = sy + P(px, py + 2, pz) * W.w3;
      sz = sz + P(px, py, pz - 2) * W.w0;
      sz = sz + P(px, py, pz - 1) * W.w1;
      sz = sz + P(px, py, pz + 1) * W.w2;
      sz = sz + P(px, py, pz + 2) * W.w3;
      fr[q][0] = sx;
      fr[q][1] = sy;
      fr[q][2] = sz;
      // f2max over force cells 1..nforce (kick.f90:717); cell f0+BD is the
      // next block's first cell and counting it twice is harmless.
      const int m[3] = {f0[0] + lx, f0[1] + ly, f0[2] + lz};
      bool in = true;
      for (int d = 0; d < 3; ++d) in = in && (m[d] >= 1 && m[d] <= L.nforce && m[d] <= f0[d] + BD);
      if (in) f2 = max(f2, (sx * sx + sy * sy) + sz * sz);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NF; ++q) {
    const int t = threadIdx.x + q * 256;
    if (t < FS * FS * FS) {
      lds[t] = fr[q][0];
      lds[FS * FS * FS + t] = fr[q][1];
      lds[2 * FS * FS * FS + t] = fr[q][2];
    }
  }
  __syncthreads();
  // 3. particles whose floor(xpos) lies in this block
  constexpr double kRes = 1.0 / (double)(1u << XP<IZ>::B);
  const bool unit = !L.xdiv && L.xmul == 1.0;  // one mesh cell per coarse cell (PM1)
  R vm[3] = {0, 0, 0};
  for (int i0 = threadIdx.x; i0 < total; i0 += 2 * 256) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int i = i0 + u * 256;
      if (i >= total) continue;
      int c;
      uint32_t uu[3];
      float v0[3];
      if (i0 == (int)threadIdx.x) {
        c = pc[u];
        for (int d = 0; d < 3; ++d) {
          uu[d] = pu[u][d];
          v0[d] = pv[u][d];
        }
      } else {
        c = find_cell(pre, nC, i);
        const int64_t ipl = gof[c] + (i - pre[c]);
        for (int d = 0; d < 3; ++d) {
          uu[d] = XP<IZ>::raw(xp, ipl, d);
          v0[d] = vp[3 * ipl + d];
        }
      }
      const int ijk[3] = {clo[0] + c % ncx, clo[1] + (c / ncx) % ncy, clo[2] + c / (ncx * ncy)};
      const int64_t ip = gof[c] + (i - pre[c]);
      R w[3][3];
      int f[3];
      bool mine = true;
      for (int d = 0; d < 3; ++d) {
        if (unit) {
          // xpos = (c-1) + (u + 1/2) 2^-B - s exactly: floor c-1-s, fraction
          // (u + 1/2) 2^-B (exact in R), the FP64 path's values
          f[d] = ijk[d] - 1 - B.ushift[d] - f0[d];
          tsc_weights_frac<R>(((R)uu[d] + (R)0.5) * (R)kRes, w[d]);
        } else {
          const double xpos = pm_xpos(L, ijk[d], ((double)uu[d] + 0.5) * kRes, B.ushift[d]);
          f[d] = tsc_weights<R>(xpos, w[d]) - f0[d];
        }
        mine = mine && f[d] >= 0 && f[d] < BD;
      }
      if (!mine) continue;
      R dv[3] = {0, 0, 0};
      for (int i3 = 0; i3 < 3; ++i3)
        for (int i2 = 0; i2 < 3; ++i2)
          for (int i1 = 0; i1 < 3; ++i1) {
            const int li = ((f[2] + i3) * FS + (f[1] + i2)) * FS + (f[0] + i1);
            for (int d = 0; d < 3; ++d)
              dv[d] = dv[d] + ((lds[d * FS * FS * FS + li] * w[0][i1]) * w[1][i2]) * w[2][i3];
          }
      for (int d = 0; d < 3; ++d) {
        const float v = (float)kick_update((R)v0[d], dv[d], kc);
        vp[3 * ip + d] = v;
        vm[d] = max(vm[d], (R)fabsf(v));
      }
    }
  }
  R bm = block_max(f2, red);
  if (threadIdx.x == 0 && bm > 0) {
    if constexpr (std::is_same<R, float>::value) atomic_max_nonneg(&out->f2max, bm);
    else atomic_max_nonneg(&out->f2max_d, bm);
  }
  for (int d = 0; d < 3; ++d) {
    R m = block_max(vm[d], red);
    if (threadIdx.x == 0 && m > 0) {
      if constexpr (std::is_same<R, float>::value) atomic_max_nonneg(&out->vmax[d], m);
      else atomic_max_nonneg(&out->vmax_d[d], m);
    }
  }
}

// Interpolation for the tile and sub-tile levels with w = 4 or 8 mesh cells per
// coarse cell and a power-of-two ratio_cs/gridsize: one workgroup per group
// of coarse cells (2x2x4 or 1x1x2) whose particles' floor(xpos) fill an
// BX x BY x BZ block of mesh cells.  For such levels xpos of a particle
// of coarse cell i is w(i-1) - shift + (u + 1/2) 2^-sh exactly (sh = B -
// log2 w), so floor and fraction come from the integer coordinate and the
// block's particles are exactly those of its cells: each is read once, unlike a mesh-aligned block, which reads
// every coarse cell its cells touch.  The staging, FD force, f2max and the
// 27-point gather are those of k_interp.
// Block shapes (BX, BY, BZ, threads), picked per level by the engine's
// autotuner (every shape gives the same result): a larger block stages a
// smaller share of halo (phi halo 4.2 -> 3.3 values per cell, FD force on 1.76
// -> 1.58 cells per cell) with twice the threads.
template <typename R, int IZ, int BX, int BY, int BZ, int NT>
__global__ void __launch_bounds__(NT) k_interp_cells(PMLevel L, const PMBox* __restrict__ boxes,
                                                      const XT<IZ>* __restrict__ xp, float* __restrict__ vp,
                                                      const int32_t* __restrict__ rhoc,
                                                      const int64_t* __restrict__ off, const R* __restrict__ phi,
                                                      double kc, PMOut* __restrict__ out) {
  if (L.kdev) kc = *L.kdev;
  constexpr int FX = BX + 2, FY = BY + 2, FZ = BZ + 2, FN = FX * FY * FZ;
  constexpr int PX = BX + 6, PY = BY + 6, PZ = BZ + 6, PN = PX * PY * PZ;
  constexpr int NF = (FN + NT - 1) / NT, NPH = (PN + NT - 1) / NT;
  constexpr int LDSN = PN > 3 * FN ? PN : 3 * FN;
  constexpr int MAXC = (BX / 4) * (BY / 4) * (BZ / 4);  // coarse cells of a block at w >= 4
  constexpr int NW = NT / 64;
  __shared__ R lds[LDSN];
  __shared__ R red[NW];
  __shared__ int cnt[MAXC], pre[MAXC + 1], wsum[NW], cxyz[MAXC];
  __shared__ int64_t gof[MAXC];
  const int w = L.w;
  const int g[3] = {BX / w, BY / w, BZ / w};  // coarse cells per block per axis
  const int ncell = L.nforce / w;                // physical coarse cells of a box per axis
  const int nb[3] = {(ncell + g[0] - 1) / g[0], (ncell + g[1] - 1) / g[1], (ncell + g[2] - 1) / g[2]};
  // neighbouring blocks share phi halos: keep them on one XCD's L2
  int64_t b = xcd_block(blockIdx.x, gridDim.x);
  const int bx = (int)(b % nb[0]);
  b /= nb[0];
  const int by = (int)(b % nb[1]);
  b /= nb[1];
  const int bz = (int)(b % nb[2]);
  const int ibox = (int)(b / nb[2]);
  const PMBox B = boxes[ibox];
  const int bi[3] = {bx, by, bz};
  int clo[3], chi[3], f0[3], fend[3];
  for (int d = 0; d < 3; ++d) {
    clo[d] = B.nc1[d] + bi[d] * g[d];
    chi[d] = min(clo[d] + g[d] - 1, B.nc2[d]);
    f0[d] = w * (clo[d] - 1) - B.ushift[d];
    fend[d] = f0[d] + w * (chi[d] - clo[d] + 1);  // last force cell this block counts for f2max
  }
  // 1. phi(f0-2 .. f0+IB+3) and the cell metadata, all loads before the
  //    first LDS store
  R ph[NPH];
#pragma unroll
  for (int q = 0; q < NPH; ++q) {
    const int t = threadIdx.x + q * NT;
    ph[q] = (R)0;
    if (t < PN) {
      const int l[3] = {t % PX, (t / PX) % PY, t / (PX * PY)};
      int a[3];
      bool ok = true;
      for (int d = 0; d < 3; ++d) {
        a[d] = f0[d] - 2 + l[d] + B.org[d] - L.m0;
        ok = ok && a[d] >= 0 && a[d] < L.n;
      }
      if (ok) ph[q] = phi[B.moff + ((int64_t)a[2] * L.n + a[1]) * L.n + a[0]];
    }
  }
  const int ncx = chi[0] - clo[0] + 1, ncy = chi[1] - clo[1] + 1, ncz = chi[2] - clo[2] + 1;
  const int nC = ncx * ncy * ncz;
  for (int t = threadIdx.x; t < nC; t += blockDim.x) {
    const int cx = t % ncx, cy = (t / ncx) % ncy, cz = t / (ncx * ncy);
    const int64_t cell = rc_index(clo[0] + cx, clo[1] + cy, clo[2] + cz, B.T, L.nte, L.ncb);
    cnt[t] = rhoc[cell];
    gof[t] = off[cell];
    cxyz[t] = cx | (cy << 5) | (cz << 10);
  }
#pragma unroll
  for (int q = 0; q < NPH; ++q) {
    const int t = threadIdx.x + q * NT;
    if (t < PN) lds[t] = ph[q];
  }
  __syncthreads();
  block_scan(cnt, pre, nC, wsum);
  const int total = pre[nC];
  // 2. force(d, f0 .. f0+IB+1) = sum(phi(m + [-2,-1,1,2]) * weight), fused
  //    multiply-adds (one rounding per term; the kick tolerance is 1e-5)
  const FDWeights<R> W;
  R fr[NF][3];
  R f2 = 0;
#pragma unroll
  for (int q = 0; q < NF; ++q) {
    const int t = threadIdx.x + q * NT;
    if (t < FN) {
      const int lx = t % FX, ly = (t / FX) % FY, lz = t / (FX * FY);
      const int px = lx + 2, py = ly + 2, pz = lz + 2;
      auto P = [&](int x, int y, int z) { return lds[(z * PY + y) * PX + x]; };
      R sx = 0, sy = 0, sz = 0;
      sx = fma(P(px - 2, py, pz), W.w0, sx);
      sx = fma(P(px - 1, py, pz), W.w1, sx);
      sx = fma(P(px + 1, py, pz), W.w2, sx);
      sx = fma(P(px + 2, py, pz), W.w3, sx);
      sy = fma(P(px, py - 2, pz), W.w0, sy);
      sy = fma(P(px, py - 1, pz), W.w1, sy);
      sy = fma(P(px, py + 1, pz), W.w2, sy);
      sy = fma(P(px, py + 2, pz), W.w3, sy);
      sz = fma(P(px, py, pz - 2), W.w0, sz);
      sz = fma(P(px, py, pz - 1), W.w1, sz);
      sz = fma(P(px, py, pz + 1), W.w2, sz);
      sz = fma(P(px, py, pz + 2), W.w3, sz);
      fr[q][0] = sx;
      fr[q][1] = sy;
      fr[q][2] = sz;
      // f2max over force cells 1..nforce (kick.f90:717); neighbouring blocks
      // share their boundary cell, which is harmless for a maximum
      const int m[3] = {f0[0] + lx, f0[1] + ly, f0[2] + lz};
      bool in = true;
      for (int d = 0; d < 3; ++d) in = in && (m[d] >= 1 && m[d] <= L.nforce && m[d] <= fend[d]);
      if (in) f2 = max(f2, (sx * sx + sy * sy) + sz * sz);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NF; ++q) {
    const int t = threadIdx.x + q * NT;
    if (t < FN) {
      lds[t] = fr[q][0];
      lds[FN + t] = fr[q][1];
      lds[2 * FN + t] = fr[q][2];
    }
  }
  __syncthreads();
  // 3. every particle of the block's cells
  const int sh = XP<IZ>::B - (31 - __clz(w));
  const R rsh = (R)1 / (R)(1u << sh);
  R vm[3] = {0, 0, 0};
  for (int i = threadIdx.x; i < total; i += NT) {
    const int c = find_cell(pre, nC, i);
    const int64_t ip = gof[c] + (i - pre[c]);
    uint32_t uu[3];
    float v0[3];
    for (int d = 0; d < 3; ++d) {
      uu[d] = XP<IZ>::raw(xp, ip, d);
      v0[d] = vp[3 * ip + d];
    }
    const int cc = cxyz[c];
    const int ijk[3] = {clo[0] + (cc & 31), clo[1] + ((cc >> 5) & 31), clo[2] + ((cc >> 10) & 31)};
    R wt[3][3];
    int f[3];
    bool ok = true;
    for (int d = 0; d < 3; ++d) {
      // xpos = w(c-1) - s + (u + 1/2) 2^-sh exactly (sh = B - log2 w)
      f[d] = L.w * (ijk[d] - 1) - B.ushift[d] + (int)(uu[d] >> sh) - f0[d];
      tsc_weights_frac<R>(((R)(uu[d] & ((1u << sh) - 1u)) + (R)0.5) * rsh, wt[d]);
      ok = ok && f[d] >= 0 && f[d] < (d == 0 ? BX : (d == 1 ? BY : BZ));
    }
    if (!ok) continue;  // unreachable for a power-of-two xmul; guards the LDS reads
    R dv[3] = {0, 0, 0};
    for (int i3 = 0; i3 < 3; ++i3)
      for (int i2 = 0; i2 < 3; ++i2)
        for (int i1 = 0; i1 < 3; ++i1) {
          const int li = ((f[2] + i3) * FY + (f[1] + i2)) * FX + (f[0] + i1);
          for (int d = 0; d < 3; ++d)
            dv[d] = dv[d] + ((lds[d * FN + li] * wt[0][i1]) * wt[1][i2]) * wt[2][i3];
        }
    for (int d = 0; d < 3; ++d) {
      const float v = (float)kick_update((R)v0[d], dv[d], kc);
      vp[3 * ip + d] = v;
      vm[d] = max(vm[d], (R)fabsf(v));
    }
  }
  R bm = block_max(f2, red);
  if (threadIdx.x == 0 && bm > 0) {
    if constexpr (std::is_same<R, float>::value) atomic_max_nonneg(&out->f2max, bm);
    else atomic_max_nonneg(&out->f2max_d, bm);
  }
  for (int d = 0; d < 3; ++d) {
    R m = block_max(vm[d], red);
    if (threadIdx.x == 0 && m > 0) {
      if constexpr (std::is_same<R, float>::value) atomic_max_nonneg(&out->vmax[d], m);
      else atomic_max_nonneg(&out->vmax_d[d], m);
    }
  }
}

// Green_3D, Green/Green.f90:22-71 at nn=1.
__global__ void k_green(int n, float apm_soft, float apm_range, float gridsize, double factor, float* out) {
  const int nh = n / 2 + 1;
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nh * n * n) return;
  const int k1 = (int)(t % nh), k2 = (int)((t / nh) % n), k3 = (int)(t / ((int64_t)nh * n));
  const float inv = 1.0f / (float)n;  // (1./nglobal) in default real
  auto fsp = [&](int k) {  // fftspace(k+1)
    int m = ((k + 1 + n / 2 - 1) % n) - n / 2;
    return (2.0 * kPi * (double)inv) * (double)m;
  };
  const double kv[3] = {fsp(k1), fsp(k2), fsp(k3)};
  const double alpha = (double)(4.0f / 3.0f), oma = (double)(1.0f - 4.0f / 3.0f), c215 = (double)(2.0f / 15.0f);
  double D[3], ur[3] = {0, 0, 0};
  for (int d = 0; d < 3; ++d) D[d] = alpha * sin(kv[d]) + oma * sin(2 * kv[d]) / 2;
  for (int i3 = -2; i3 <= 2; ++i3)
    for (int i2 = -2; i2 <= 2; ++i2)
      for (int i1 = -2; i1 <= 2; ++i1) {
        const double kn[3] = {kv[0] + 2 * kPi * i1, kv[1] + 2 * kPi * i2, kv[2] + 2 * kPi * i3};
        const double kmag = sqrt(kn[0] * kn[0] + kn[1] * kn[1] + kn[2] * kn[2]);
        if (kmag == 0) continue;  // 0/0 at k=0 only; that mode is zeroed
        double U = 1;
        for (int d = 0; d < 3; ++d) U *= (kn[d] == 0) ? 1.0 : sin(kn[d] / 2) / (kn[d] / 2);
        U = U * U * U * U * U * U;  // **(2p+2), p=2
        double y = (double)apm_soft * kmag / 2;
        double y2 = y * y;
        double Ss = 12 * (1.0 / (y2 * y2)) * (2 - 2 * cos(y) - y * sin(y));
        double Sr = 0;
        if (apm_range != 0) {
          y = (double)apm_range * kmag / 2;
          y2 = y * y;
          Sr = 12 * (1.0 / (y2 * y2)) * (2 - 2 * cos(y) - y * sin(y));
        }
        for (int d = 0; d < 3; ++d) ur[d] += U * (kn[d] * (Ss * Ss - Sr * Sr) / (kmag * kmag));
      }
  double u2n = 1;
  for (int d = 0; d < 3; ++d) {
    double s = sin(kv[d] / 2);
    u2n *= 1 - s * s + c215 * (s * s) * (s * s);
  }
  const double sDu = D[0] * ur[0] + D[1] * ur[1] + D[2] * ur[2];
  const double sDD = D[0] * D[0] + D[1] * D[1] + D[2] * D[2];
  const float gs2 = gridsize * gridsize;
  double G = 4 * kPi * sDu / sDD / (u2n * u2n) / (double)gs2;
  const bool corner = (k1 == 0 || k1 == n / 2) && (k2 == 0 || k2 == n / 2) && (k3 == 0 || k3 == n / 2);
  if (corner) G = 0;  // Green.f90:71
  float g32 = (float)G;
  out[t] = (factor == 1.0) ? g32 : (float)(factor * (double)g32);
}

// kick.f90:110-133: npc_max = maxval(rhoc(sub-tile))/64.
__global__ void k_classify(Geo g, int ires_fixed, const int32_t* __restrict__ rhoc, int32_t* __restrict__ ires,
                           int32_t* __restrict__ napm, int32_t* __restrict__ cmax) {
  __shared__ int32_t red[256];
  const int64_t it = blockIdx.x;
  const int64_t nns = g.nns;
  const int64_t ifs[3] = {it % nns, (it / nns) % nns, (it / (nns * nns)) % nns};
  const int64_t T = it / (nns * nns * nns);
  const int64_t ntt = g.ntt;
  int32_t mx = 0;
  for (int64_t q = threadIdx.x; q < ntt * ntt * ntt; q += blockDim.x) {
    int64_t i = ifs[0] * ntt + q % ntt + 1, j = ifs[1] * ntt + (q / ntt) % ntt + 1, k = ifs[2] * ntt + q / (ntt * ntt) + 1;
    mx = max(mx, rhoc[rc_index(i, j, k, T, g.nte, g.ncb)]);
  }
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    float npc = (float)red[0] / 64.0f;
    int r = npc < 24.0f ? 2 : (npc < 114.0f ? 3 : (npc < 320.0f ? 4 : 5));
    if (ires_fixed >= 2) r = ires_fixed;  // kick.f90:131 (commented there)
    ires[it] = r;
    cmax[it] = red[0];
    atomicAdd(&napm[r - 1], 1);
  }
}

// Three-phase exclusive scan (block sums, scan of sums, block scan + add).
constexpr int kScanBlock = 1024;

__global__ void k_scan_sums(const int32_t* __restrict__ in, int64_t n, int64_t* __restrict__ sums) {
  __shared__ int64_t red[256];
  const int64_t b0 = (int64_t)blockIdx.x * kScanBlock;
  int64_t s = 0;
  for (int q = threadIdx.x; q < kScanBlock; q += blockDim.x)
    if (b0 + q < n) s += in[b0 + q];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = blockDim.x / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = red[0];
}

__global__ void k_scan_top(int64_t* sums, int64_t nb) {
  // single workgroup sequential-over-chunks scan of the block sums
  __shared__ int64_t buf[1024];
  __shared__ int64_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t c0 = 0; c0 < nb; c0 += 1024) {
    int64_t v = (c0 + threadIdx.x < nb) ? sums[c0 + threadIdx.x] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      int64_t a = (threadIdx.x >= (unsigned)o) ? buf[threadIdx.x - o] : 0;
      __syncthreads();
      buf[threadIdx.x] += a;
      __syncthreads();
    }
    if (c0 + threadIdx.x < nb) sums[c0 + threadIdx.x] = carry + buf[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += buf[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[nb] = carry;
}

__global__ void k_scan_apply(const int32_t* __restrict__ in, int64_t n, const int64_t* __restrict__ sums,
                             int64_t* __restrict__ out, int64_t nb) {
  __shared__ int64_t buf[kScanBlock];
  const int64_t b0 = (int64_t)blockIdx.x * kScanBlock;
  for (int q = threadIdx.x; q < kScanBlock; q += blockDim.x) buf[q] = (b0 + q < n) ? in[b0 + q] : 0;
  __syncthreads();
  for (int o = 1; o < kScanBlock; o <<= 1) {
    int64_t a[kScanBlock / 256];
    for (int r = 0; r < kScanBlock / 256; ++r) {
      int q = threadIdx.x + r * 256;
      a[r] = q >= o ? buf[q - o] : 0;
    }
    __syncthreads();
    for (int r = 0; r < kScanBlock / 256; ++r) buf[threadIdx.x + r * 256] += a[r];
    __syncthreads();
  }
  const int64_t base = sums[blockIdx.x];
  for (int q = threadIdx.x; q < kScanBlock; q += blockDim.x)
    if (b0 + q < n) out[b0 + q] = base + buf[q] - in[b0 + q];
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = sums[nb];
}

// Deposit block edge: 16 fine cells for the tile/sub-tile levels (FP32), 8
// for the coarse level (one mesh cell per coarse cell: 8^3 cells already hold
// ~32k particles) and for the FP64 form (LDS size).  Interpolation blocks
// are 8^3 force cells (LDS for phi + force of 14^3 and 3*10^3 values).
template <typename R, int IZ, int BD>
void launch_deposit_bd(const PMLevel& L, const PMBox* boxes, int nbox, const void* xp, const int32_t* rhoc,
                       const int64_t* off, R* mesh, float mass, hipStream_t s) {
  // Cell-aligned blocks when w divides BD, w < BD/2 and the window is the
  // deposited cells' floors plus one cell (win_lo = 1 - nloop w,
  // win_hi = nforce + nloop w): the regions, counted down from the last
  // deposited cell, then cover it.
  const bool ca = !L.xdiv && 2 * L.w < BD && BD % L.w == 0 && L.nforce % L.w == 0 &&
                  L.win_lo == 1 - L.nloop * L.w && L.win_hi == L.nforce + L.nloop * L.w;
  const int nblk =
      ca ? 1 + (L.nforce / L.w - 1 + 2 * L.nloop + BD / L.w - 1) / (BD / L.w) : (L.win_hi - L.win_lo + BD) / BD;
  const int64_t grid = (int64_t)nblk * nblk * nblk * nbox;
  if (ca)
    hipLaunchKernelGGL((k_deposit<R, IZ, BD, true>), dim3((unsigned)grid), dim3(256), 0, s, L, boxes,
                       (const XT<IZ>*)xp, rhoc, off, mesh, mass);
  else
    hipLaunchKernelGGL((k_deposit<R, IZ, BD, false>), dim3((unsigned)grid), dim3(256), 0, s, L, boxes,
                       (const XT<IZ>*)xp, rhoc, off, mesh, mass);
  CUBE_HIP(hipGetLastError());
}

template <typename R, int IZ>
void launch_deposit(const PMLevel& L, const PMBox* boxes, int nbox, const void* xp, const int32_t* rhoc,
                    const int64_t* off, R* mesh, float mass, hipStream_t s) {
  // a block reads at most (BD/w + 2)^3 <= 216 coarse cells for w >= 4
  if (L.w < 4) throw Error(1, "fine-level deposit needs >= 4 mesh cells per coarse cell");
  if (std::is_same<R, float>::value)
    launch_deposit_bd<R, IZ, 16>(L, boxes, nbox, xp, rhoc, off, mesh, mass, s);
  else
    launch_deposit_bd<R, IZ, 8>(L, boxes, nbox, xp, rhoc, off, mesh, mass, s);
}

template <typename R, int IZ, int BD>
void launch_interp_bd(const PMLevel& L, const PMBox* boxes, int nbox, const void* xp, float* vp, const int32_t* rhoc,
                      const int64_t* off, const R* phi, double kc, PMOut* out, hipStream_t s) {
  const int nblk = (L.nforce + BD - 1) / BD;
  const int64_t grid = (int64_t)nblk * nblk * nblk * nbox;
  // a block reads at most (BD/w + 1)^3 coarse cells
  if (L.w >= 4)
    hipLaunchKernelGGL((k_interp<R, IZ, BD, (BD / 4 + 1) * (BD / 4 + 1) * (BD / 4 + 1)>), dim3((unsigned)grid),
                       dim3(256), 0, s, L, boxes, (const XT<IZ>*)xp, vp, rhoc, off, phi, kc, out);
  else
    hipLaunchKernelGGL((k_interp<R, IZ, BD, (BD + 1) * (BD + 1) * (BD + 1)>), dim3((unsigned)grid), dim3(256), 0,
                       s, L, boxes, (const XT<IZ>*)xp, vp, rhoc, off, phi, kc, out);
  CUBE_HIP(hipGetLastError());
}

template <typename R, int IZ, int BX, int BY, int BZ, int NT>
void launch_interp_cells(const PMLevel& L, const PMBox* boxes, int nbox, const void* xp, float* vp,
                         const int32_t* rhoc, const int64_t* off, const R* phi, double kc, PMOut* out,
                         hipStream_t s) {
  const int ncell = L.nforce / L.w;
  int64_t grid = nbox;
  const int g[3] = {BX / L.w, BY / L.w, BZ / L.w};
  for (int d = 0; d < 3; ++d) grid *= (ncell + g[d] - 1) / g[d];
  hipLaunchKernelGGL((k_interp_cells<R, IZ, BX, BY, BZ, NT>), dim3((unsigned)grid), dim3(NT), 0, s, L, boxes,
                     (const XT<IZ>*)xp, vp, rhoc, off, phi, kc, out);
  CUBE_HIP(hipGetLastError());
}

template <typename R, int IZ>
void launch_interp(const PMLevel& L, const PMBox* boxes, int nbox, const void* xp, float* vp, const int32_t* rhoc,
                   const int64_t* off, const R* phi, double kc, PMOut* out, hipStream_t s, int shape) {
  if (pm_interp_cells_form(L)) {
    // the larger shapes' LDS (FP32: 38.9 KB) exceeds the static limit in FP64
    if constexpr (std::is_same<R, float>::value) {
      if (shape == 1) {
        launch_interp_cells<R, IZ, 8, 16, 16, 512>(L, boxes, nbox, xp, vp, rhoc, off, phi, kc, out, s);
        return;
      }
      if (shape == 2) {
        launch_interp_cells<R, IZ, 16, 16, 8, 512>(L, boxes, nbox, xp, vp, rhoc, off, phi, kc, out, s);
        return;
      }
    }
    launch_interp_cells<R, IZ, 8, 8, 16, 256>(L, boxes, nbox, xp, vp, rhoc, off, phi, kc, out, s);
    return;
  }
  launch_interp_bd<R, IZ, 8>(L, boxes, nbox, xp, vp, rhoc, off, phi, kc, out, s);
}

}  // namespace

template <typename R>
void pm_deposit(int izipx, const PMLevel& L, const PMBox* boxes, int nbox, const void* xp, const int32_t* rhoc,
                const int64_t* off, R* mesh, float mass, hipStream_t s) {
  if (izipx == 2) launch_deposit<R, 2>(L, boxes, nbox, xp, rhoc, off, mesh, mass, s);
  else launch_deposit<R, 1>(L, boxes, nbox, xp, rhoc, off, mesh, mass, s);
}

bool pm_deposit_binned_applies(const PMLevel& L) {
  // w = 8 (two blocks per axis per coarse cell), power-of-two xmul, and the
  // window of the deposited cells' taps; box-local cell coordinates < 32
  const int nr = L.nforce / L.w + 2 * L.nloop;
  return !L.periodic && !L.xdiv && L.w == 8 && L.nforce % L.w == 0 && L.win_lo == 1 - L.nloop * L.w &&
         L.win_hi == L.nforce + L.nloop * L.w && nr <= 32;
}

int64_t pm_dep_blocks(const PMLevel& L) {
  const int64_t nb = dep_nblk(L);
  return nb * nb * nb;
}

namespace {
template <int FILL>
void launch_dep_bin(int izipx, const PMLevel& L, const PMBox* boxes, int nbox, const void* xp, const int32_t* rhoc,
                    const int64_t* off, DepBins& S, hipStream_t s) {
  const int nr = L.nforce / L.w + 2 * L.nloop;
  const int64_t nthr = (int64_t)nbox * nr * nr * nr * 64;
  if (izipx == 2)
    hipLaunchKernelGGL((k_dep_bin<2, FILL>), dim3((unsigned)ceil_div(nthr, 256)), dim3(256), 0, s, L, boxes, nbox,
                       (const int16_t*)xp, rhoc, off, S.count, S.start, S.ents);
  else
    hipLaunchKernelGGL((k_dep_bin<1, FILL>), dim3((unsigned)ceil_div(nthr, 256)), dim3(256), 0, s, L, boxes, nbox,
                       (const int8_t*)xp, rhoc, off, S.count, S.start, S.ents);
  CUBE_HIP(hipGetLastError());
}
}  // namespace

int64_t pm_dep_bin_count(int izipx, const PMLevel& L, const PMBox* boxes, int nbox, const void* xp,
                         const int32_t* rhoc, const int64_t* off, DepBins& S, hipStream_t s) {
  (void)izipx;
  (void)xp;
  (void)off;
  const int64_t nblk = (int64_t)nbox * pm_dep_blocks(L);
  hipLaunchKernelGGL(k_dep_cap, dim3((unsigned)ceil_div(nblk, 256)), dim3(256), 0, s, L, boxes, nbox, rhoc, S.count);
  CUBE_HIP(hipGetLastError());
  scan_counts(S.count, S.start, nblk, S.scan_tmp, S.scan_tmp_bytes, s);
  int64_t total = 0;
  CUBE_HIP(hipMemcpyAsync(&total, S.start + nblk, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  CUBE_HIP(hipStreamSynchronize(s));
  return total;
}

void pm_dep_bin_fill(int izipx, const PMLevel& L, const PMBox* boxes, int nbox, const void* xp, const int32_t* rhoc,
                     const int64_t* off, DepBins& S, hipStream_t s) {
  zero_async(S.count, sizeof(int32_t) * nbox * pm_dep_blocks(L), s);
  launch_dep_bin<1>(izipx, L, boxes, nbox, xp, rhoc, off, S, s);
}

template <typename R, int IZ>
static void launch_deposit_lists(int threads, int64_t grid, const PMLevel& L, const PMBox* boxes, const DepBins& S,
                                 R* mesh, float mass, hipStream_t s) {
  if (threads == 256)
    hipLaunchKernelGGL((k_deposit_lists<R, IZ, 256>), dim3((unsigned)grid), dim3(256), 0, s, L, boxes, S.start,
                       S.count, S.ents, mesh, mass);
  else if (threads == 1024)
    hipLaunchKernelGGL((k_deposit_lists<R, IZ, 1024>), dim3((unsigned)grid), dim3(1024), 0, s, L, boxes, S.start,
                       S.count, S.ents, mesh, mass);
  else
    hipLaunchKernelGGL((k_deposit_lists<R, IZ, 512>), dim3((unsigned)grid), dim3(512), 0, s, L, boxes, S.start,
                       S.count, S.ents, mesh, mass);
}

template <typename R>
void pm_deposit_lists(int izipx, const PMLevel& L, const PMBox* boxes, int nbox, const DepBins& S, R* mesh,
                      float mass, hipStream_t s, int threads) {
  const int64_t grid = (int64_t)nbox * pm_dep_blocks(L);
  if (izipx == 2) launch_deposit_lists<R, 2>(threads, grid, L, boxes, S, mesh, mass, s);
  else launch_deposit_lists<R, 1>(threads, grid, L, boxes, S, mesh, mass, s);
  CUBE_HIP(hipGetLastError());
}
template void pm_deposit_lists<float>(int, const PMLevel&, const PMBox*, int, const DepBins&, float*, float,
                                      hipStream_t, int);
template void pm_deposit_lists<double>(int, const PMLevel&, const PMBox*, int, const DepBins&, double*, float,
                                       hipStream_t, int);

void pm_level_scale(PMLevel& L) {
  int e = 0;
  const double m = std::frexp(L.gridsize, &e);
  L.xdiv = (m == 0.5) ? 0 : 1;
  L.xmul = (double)L.rcs / L.gridsize;
}

size_t pm1_stencil_bytes(const PMLevel& L) { return (size_t)27 * L.n * L.n * L.n * 8; }

template <typename R>
void pm1_deposit(int izipx, const PMLevel& L, const void* xp, const int32_t* rhoc, const int64_t* off,
                 void* stencil, R* mesh, float mass, PMOut* out, hipStream_t s) {
  if (L.w != 1 || !L.periodic || L.n % L.nt != 0) throw Error(1, "coarse deposit needs the periodic w=1 level");
  const int64_t ncell = (int64_t)L.n * L.n * L.n;
  const unsigned g1 = (unsigned)ceil_div(ncell, 2);
  if (izipx == 2)
    hipLaunchKernelGGL((k_pm1_stencil<R, 2>), dim3(g1), dim3(128), 0, s, L, (const int16_t*)xp, rhoc, off, stencil);
  else
    hipLaunchKernelGGL((k_pm1_stencil<R, 1>), dim3(g1), dim3(128), 0, s, L, (const int8_t*)xp, rhoc, off, stencil);
  CUBE_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_pm1_gather<R>, dim3((unsigned)ceil_div(ncell, 256)), dim3(256), 0, s, L,
                     (const void*)stencil, mesh, mass, out);
  CUBE_HIP(hipGetLastError());
}

template <typename R>
void pm_interp(int izipx, const PMLevel& L, const PMBox* boxes, int nbox, const void* xp, float* vp,
               const int32_t* rhoc, const int64_t* off, const R* phi, double kc, PMOut* out,
               hipStream_t s, int shape) {
  if (izipx == 2) launch_interp<R, 2>(L, boxes, nbox, xp, vp, rhoc, off, phi, kc, out, s, shape);
  else launch_interp<R, 1>(L, boxes, nbox, xp, vp, rhoc, off, phi, kc, out, s, shape);
}

bool pm_interp_cells_form(const PMLevel& L) {
  return !L.periodic && !L.xdiv && (L.w == 4 || L.w == 8) && L.nforce % L.w == 0;
}

template void pm_deposit<float>(int, const PMLevel&, const PMBox*, int, const void*, const int32_t*, const int64_t*,
                                float*, float, hipStream_t);
template void pm_deposit<double>(int, const PMLevel&, const PMBox*, int, const void*, const int32_t*, const int64_t*,
                                 double*, float, hipStream_t);
template void pm1_deposit<float>(int, const PMLevel&, const void*, const int32_t*, const int64_t*, void*, float*, float,
                                 PMOut*, hipStream_t);
template void pm1_deposit<double>(int, const PMLevel&, const void*, const int32_t*, const int64_t*, void*, double*,
                                  float, PMOut*, hipStream_t);
template void pm_interp<float>(int, const PMLevel&, const PMBox*, int, const void*, float*, const int32_t*,
                               const int64_t*, const float*, double, PMOut*, hipStream_t, int);
template void pm_interp<double>(int, const PMLevel&, const PMBox*, int, const void*, float*, const int32_t*,
                                const int64_t*, const double*, double, PMOut*, hipStream_t, int);

void green_device(int n, float apm_soft, float apm_range, float gridsize, double factor, float* out, hipStream_t s) {
  int64_t tot = (int64_t)(n / 2 + 1) * n * n;
  hipLaunchKernelGGL(k_green, dim3((unsigned)ceil_div(tot, 256)), dim3(256), 0, s, n, apm_soft, apm_range, gridsize,
                     factor, out);
  CUBE_HIP(hipGetLastError());
}

__global__ void k_zero(uint32_t* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0;
}
__global__ void k_zero_bytes(unsigned char* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0;
}

// Order-independent 64-bit checksum of a buffer of 32-bit words: the sum of
// word * odd(index), so equal buffers give equal sums and a changed word
// changes the sum.
__global__ void k_checksum(const uint32_t* __restrict__ w, int64_t n, unsigned long long* __restrict__ out) {
  unsigned long long acc = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    acc += (unsigned long long)w[i] * (2ull * (unsigned long long)i + 1ull);
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(out, acc);
}

uint64_t device_checksum(const void* p, size_t bytes, void* scratch8, hipStream_t s) {
  zero_async(scratch8, 8, s);
  const int64_t n = (int64_t)(bytes / 4);
  hipLaunchKernelGGL(k_checksum, dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 8192))),
                     dim3(256), 0, s, (const uint32_t*)p, n, (unsigned long long*)scratch8);
  CUBE_HIP(hipGetLastError());
  uint64_t v = 0;
  CUBE_HIP(hipMemcpyAsync(&v, scratch8, 8, hipMemcpyDeviceToHost, s));
  CUBE_HIP(hipStreamSynchronize(s));
  return v;
}

namespace {
template <typename R>
__global__ void k_mesh_add(R* __restrict__ a, const R* __restrict__ b, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = a[i] + b[i];
}
}  // namespace

namespace {
__global__ void k_fill_pattern(float* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t z = (uint64_t)i * 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 31)) * 0xbf58476d1ce4e5b9ull;
    z ^= z >> 29;
    p[i] = (float)(z >> 40) * (1.0f / 16777216.0f) - 0.5f;
  }
}
}  // namespace

void fill_pattern(float* p, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  const int64_t blocks = std::min<int64_t>(ceil_div(n, 256), 65536);
  hipLaunchKernelGGL(k_fill_pattern, dim3((unsigned)blocks), dim3(256), 0, s, p, n);
  CUBE_HIP(hipGetLastError());
}

template <typename R>
void mesh_add(R* a, const R* b, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  const int64_t blocks = std::min<int64_t>(ceil_div(n, 256), 65536);
  hipLaunchKernelGGL(k_mesh_add<R>, dim3((unsigned)blocks), dim3(256), 0, s, a, b, n);
  CUBE_HIP(hipGetLastError());
}
template void mesh_add<float>(float*, const float*, int64_t, hipStream_t);
template void mesh_add<double>(double*, const double*, int64_t, hipStream_t);

void zero_async(void* p, size_t bytes, hipStream_t s) {
  if (bytes == 0) return;
  if (((uintptr_t)p % 4) == 0 && bytes % 4 == 0) {
    int64_t n = (int64_t)(bytes / 4);
    int64_t blocks = std::min<int64_t>(ceil_div(n, 256), 65536);
    hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, s, (uint32_t*)p, n);
  } else {
    int64_t n = (int64_t)bytes;
    int64_t blocks = std::min<int64_t>(ceil_div(n, 256), 65536);
    hipLaunchKernelGGL(k_zero_bytes, dim3((unsigned)blocks), dim3(256), 0, s, (unsigned char*)p, n);
  }
  CUBE_HIP(hipGetLastError());
}

void classify_subtiles(const Geo& g, int ires_fixed, const int32_t* rhoc, int32_t* ires, int32_t* napm,
                       int32_t* cmax, hipStream_t s) {
  zero_async(napm, sizeof(int32_t) * 7, s);
  hipLaunchKernelGGL(k_classify, dim3((unsigned)g.ns3), dim3(256), 0, s, g, ires_fixed, rhoc, ires, napm, cmax);
  CUBE_HIP(hipGetLastError());
}

size_t scan_tmp_bytes(int64_t n) { return sizeof(int64_t) * (size_t)(ceil_div(n, kScanBlock) + 2); }

void scan_counts(const int32_t* in, int64_t* out, int64_t n, void* tmp, size_t tmp_bytes, hipStream_t s) {
  int64_t nb = ceil_div(n, kScanBlock);
  if (tmp_bytes < scan_tmp_bytes(n)) throw Error(1, "scan scratch too small");
  int64_t* sums = (int64_t*)tmp;
  if (nb == 0) {
    zero_async(out, sizeof(int64_t), s);
    return;
  }
  hipLaunchKernelGGL(k_scan_sums, dim3((unsigned)nb), dim3(256), 0, s, in, n, sums);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, sums, nb);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(256), 0, s, in, n, sums, out, nb);
  CUBE_HIP(hipGetLastError());
}

}  // namespace cube
