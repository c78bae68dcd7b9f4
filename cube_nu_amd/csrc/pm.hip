// pm.hip -- TSC deposit, gradient + interpolation + kick, Green's functions.
//
// Deposit: one workgroup per BD^3 block of a box mesh.  It reads the
// particles of every coarse cell whose TSC footprint reaches the block
// (particles are cell-sorted, so these are contiguous runs), one wavefront
// per coarse cell, and accumulates (wx*wy)*wz into LDS.  The FP32 form sums
// in 2^-40 fixed point with 64-bit integer LDS atomics: integer addition is
// associative, so the result does not depend on thread order and is more
// accurate than an FP32 running sum.  The FP64 form uses FP64 LDS atomics.
//
// Interpolation: one workgroup per BD^3 block of force cells.  The block
// stages phi (block + 3-cell halo) in LDS, evaluates the 4-point FD force
// (particle_mesh.f90:237-242) for block + 2 cells, takes the f2max maximum
// over its physical cells, and applies the 27-point TSC gather and kick to
// the particles whose floor(xpos) lies in the block.  The force field is
// never written to HBM.
#include <type_traits>

#include "pm.h"
#include "timing.h"

namespace cube {
namespace {

constexpr double kFix = 1099511627776.0;  // 2^40

__device__ inline int floor_div(int a, int b) {
  int q = a / b;
  if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
  return q;
}

template <int IZ>
using XT = typename XP<IZ>::T;

constexpr int kMaxCells = 1000;  // coarse cells one block reads: (8+2)^3 on the coarse level

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

template <typename R, int IZ, int BD>
__global__ void __launch_bounds__(256) k_deposit(PMLevel L, const PMBox* __restrict__ boxes,
                                                 const XT<IZ>* __restrict__ xp, const int32_t* __restrict__ rhoc,
                                                 const int64_t* __restrict__ off, R* __restrict__ mesh, float mass,
                                                 PMOut* __restrict__ out, int testrho) {
  using Acc = typename std::conditional<std::is_same<R, float>::value, unsigned long long, double>::type;
  __shared__ Acc acc[BD * BD * BD];
  __shared__ Acc red[256 / 64];
  __shared__ int cnt[kMaxCells], pre[kMaxCells + 1], wsum[4];
  __shared__ int64_t gof[kMaxCells];
  const int nblk = (L.win_hi - L.win_lo + 1 + BD - 1) / BD;
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
    mlo[d] = L.win_lo + bidx[d] * BD;
    mhi[d] = min(mlo[d] + BD - 1, L.win_hi);
    // particle of coarse cell i reaches mesh cells [w(i-1)-s, wi-s+1]
    clo[d] = max(-floor_div(-(mlo[d] - 1 + B.ushift[d]), L.w), B.nc1[d] - L.nloop);
    chi[d] = min(floor_div(mhi[d] + B.ushift[d], L.w) + 1, B.nc2[d] + L.nloop);
  }
  for (int t = threadIdx.x; t < BD * BD * BD; t += blockDim.x) acc[t] = 0;
  const int nC = stage_cells(L, B, clo, chi, rhoc, off, cnt, pre, gof, wsum);
  const int ncx = max(0, chi[0] - clo[0] + 1), ncy = max(0, chi[1] - clo[1] + 1);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  if (L.w == 1) {
    // Coarse mesh: every particle of cell (ci,cj,ck) has floor(xpos) = ci-1
    // etc., so all of them hit the same 27 mesh cells.  Lanes keep 27 partial
    // sums, the wavefront reduces them in a fixed tree, and one lane adds
    // each to LDS.
    for (int c = wave; c < nC; c += nw) {
      const int n = cnt[c];
      if (n == 0) continue;
      const int ci = clo[0] + c % ncx, cj = clo[1] + (c / ncx) % ncy, ck = clo[2] + c / (ncx * ncy);
      const int64_t base = gof[c];
      const int ijk[3] = {ci, cj, ck};
      Acc a27[27];
#pragma unroll
      for (int q = 0; q < 27; ++q) a27[q] = 0;
      for (int l = lane; l < n; l += 64) {
        R w[3][3];
        for (int d = 0; d < 3; ++d) {
          double xpos = ((double)(ijk[d] - 1) + XP<IZ>::off(xp, base + l, d)) * (double)L.rcs / L.gridsize -
                        (double)B.ushift[d];
          tsc_weights<R>(xpos, w[d]);
        }
#pragma unroll
        for (int q = 0; q < 27; ++q) {
          const R cw = (w[0][q % 3] * w[1][(q / 3) % 3]) * w[2][q / 9];
          if constexpr (std::is_same<R, float>::value) a27[q] += (Acc)__double2ull_rn((double)cw * kFix);
          else a27[q] += cw;
        }
      }
#pragma unroll
      for (int q = 0; q < 27; ++q) {
        Acc v = a27[q];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
        // taps floor(xpos) + {0,1,2} = ci-1, ci, ci+1 (particle_mesh.f90:56-58)
        const int mx = ci - 1 + q % 3, my = cj - 1 + (q / 3) % 3, mz = ck - 1 + q / 9;
        if (lane == 0 && mx >= mlo[0] && mx <= mhi[0] && my >= mlo[1] && my <= mhi[1] && mz >= mlo[2] &&
            mz <= mhi[2])
          atomicAdd(&acc[((mz - mlo[2]) * BD + (my - mlo[1])) * BD + (mx - mlo[0])], v);
      }
    }
  } else {
    // Fine meshes: particles of all the block's cells in one flat loop,
    // four per thread in flight.
    const int total = pre[nC];
    for (int i0 = threadIdx.x; i0 < total; i0 += 4 * 256) {
      int ijk[4][3];
      int64_t ip[4];
      double xo[4][3];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i = i0 + u * 256;
        ip[u] = -1;
        if (i < total) {
          const int c = find_cell(pre, nC, i);
          ijk[u][0] = clo[0] + c % ncx;
          ijk[u][1] = clo[1] + (c / ncx) % ncy;
          ijk[u][2] = clo[2] + c / (ncx * ncy);
          ip[u] = gof[c] + (i - pre[c]);
#pragma unroll
          for (int d = 0; d < 3; ++d) xo[u][d] = XP<IZ>::off(xp, ip[u], d);
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (ip[u] < 0) continue;
        R w[3][3];
        int f[3];
        for (int d = 0; d < 3; ++d) {
          double xpos = ((double)(ijk[u][d] - 1) + xo[u][d]) * (double)L.rcs / L.gridsize - (double)B.ushift[d];
          f[d] = tsc_weights<R>(xpos, w[d]);
        }
        for (int i3 = 0; i3 < 3; ++i3) {
          const int mz = f[2] + i3;
          if (mz < mlo[2] || mz > mhi[2]) continue;
          for (int i2 = 0; i2 < 3; ++i2) {
            const int my = f[1] + i2;
            if (my < mlo[1] || my > mhi[1]) continue;
            const R wyz = w[1][i2];
            for (int i1 = 0; i1 < 3; ++i1) {
              const int mx = f[0] + i1;
              if (mx < mlo[0] || mx > mhi[0]) continue;
              const R cw = (w[0][i1] * wyz) * w[2][i3];  // dx(1,i1)*dx(2,i2)*dx(3,i3)
              const int li = ((mz - mlo[2]) * BD + (my - mlo[1])) * BD + (mx - mlo[0]);
              if constexpr (std::is_same<R, float>::value)
                atomicAdd(&acc[li], (unsigned long long)__double2ull_rn((double)cw * kFix));
              else
                atomicAdd(&acc[li], cw);
            }
          }
        }
      }
    }
  }
  __syncthreads();
  Acc tsum = 0;
  for (int t = threadIdx.x; t < BD * BD * BD; t += blockDim.x) {
    const int lx = t % BD, ly = (t / BD) % BD, lz = t / (BD * BD);
    const int m[3] = {mlo[0] + lx, mlo[1] + ly, mlo[2] + lz};
    if (m[0] > mhi[0] || m[1] > mhi[1] || m[2] > mhi[2]) continue;
    R v;
    if constexpr (std::is_same<R, float>::value) v = (float)((double)acc[t] * (1.0 / kFix));
    else v = acc[t];
    tsum += acc[t];
    int a[3];
    for (int d = 0; d < 3; ++d) a[d] = m[d] + B.org[d] - L.m0;
    mesh[B.moff + ((int64_t)a[2] * L.n + a[1]) * L.n + a[0]] = v * (R)mass;
  }
  if (testrho) {
    // wave then block reduction of the (exact, for FP32) integer sums
    for (int o = 32; o > 0; o >>= 1) tsum += __shfl_down(tsum, o, 64);
    if (lane == 0) red[wave] = tsum;
    __syncthreads();
    if (threadIdx.x == 0) {
      Acc s = 0;
      for (int q = 0; q < nw; ++q) s += red[q];
      if constexpr (std::is_same<R, float>::value) atomicAdd(&out->testrho_fx[blockIdx.x % kSumSlots], s);
      else atomicAdd(&out->testrho_d[blockIdx.x % kSumSlots], s);
    }
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

template <typename R, int IZ, int BD>
__global__ void __launch_bounds__(256) k_interp(PMLevel L, const PMBox* __restrict__ boxes,
                                                const XT<IZ>* __restrict__ xp, float* __restrict__ vp,
                                                const int32_t* __restrict__ rhoc, const int64_t* __restrict__ off,
                                                const R* __restrict__ phi, float a_mid, float dt,
                                                PMOut* __restrict__ out) {
  constexpr int FS = BD + 2, PS = BD + 6;
  constexpr int NF = (FS * FS * FS + 255) / 256;
  constexpr int LDSN = (PS * PS * PS > 3 * FS * FS * FS) ? PS * PS * PS : 3 * FS * FS * FS;
  constexpr int MAXC = kMaxCells;
  __shared__ R lds[LDSN];
  __shared__ R red[4];
  __shared__ int cnt[MAXC], pre[MAXC + 1], wsum[4];
  __shared__ int64_t gof[MAXC];
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
  // 1. phi(f0-2 .. f0+BD+3) into LDS, cell metadata alongside
  for (int t = threadIdx.x; t < PS * PS * PS; t += blockDim.x) {
    const int l[3] = {t % PS, (t / PS) % PS, t / (PS * PS)};
    int a[3];
    bool ok = true;
    for (int d = 0; d < 3; ++d) {
      a[d] = f0[d] - 2 + l[d] + B.org[d] - L.m0;
      if (L.periodic) a[d] = ((a[d] % L.n) + L.n) % L.n;
      else if (a[d] < 0 || a[d] >= L.n) ok = false;
    }
    lds[t] = ok ? phi[B.moff + ((int64_t)a[2] * L.n + a[1]) * L.n + a[0]] : (R)0;
  }
  const int nC = stage_cells(L, B, clo, chi, rhoc, off, cnt, pre, gof, wsum);  // contains barriers
  const int ncx = max(0, chi[0] - clo[0] + 1), ncy = max(0, chi[1] - clo[1] + 1);
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
  // 3. particles whose floor(xpos) lies in this block, two per thread in flight
  R vm[3] = {0, 0, 0};
  const int total = pre[nC];
  for (int i0 = threadIdx.x; i0 < total; i0 += 2 * 256) {
    int ijk[2][3];
    int64_t ip[2];
    double xo[2][3];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int i = i0 + u * 256;
      ip[u] = -1;
      if (i < total) {
        const int c = find_cell(pre, nC, i);
        ijk[u][0] = clo[0] + c % ncx;
        ijk[u][1] = clo[1] + (c / ncx) % ncy;
        ijk[u][2] = clo[2] + c / (ncx * ncy);
        ip[u] = gof[c] + (i - pre[c]);
#pragma unroll
        for (int d = 0; d < 3; ++d) xo[u][d] = XP<IZ>::off(xp, ip[u], d);
      }
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      if (ip[u] < 0) continue;
      R w[3][3];
      int f[3];
      bool mine = true;
      for (int d = 0; d < 3; ++d) {
        double xpos = ((double)(ijk[u][d] - 1) + xo[u][d]) * (double)L.rcs / L.gridsize - (double)B.ushift[d];
        f[d] = tsc_weights<R>(xpos, w[d]) - f0[d];
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
        float v = vp[3 * ip[u] + d];
        v = (float)kick_update((R)v, dv[d], a_mid, dt);
        vp[3 * ip[u] + d] = v;
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
__global__ void k_classify(Geo g, const int32_t* __restrict__ rhoc, int32_t* __restrict__ ires,
                           int32_t* __restrict__ napm) {
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
    ires[it] = r;
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
                       const int64_t* off, R* mesh, float mass, PMOut* out, int testrho, hipStream_t s) {
  const int nblk = (L.win_hi - L.win_lo + 1 + BD - 1) / BD;
  const int64_t grid = (int64_t)nblk * nblk * nblk * nbox;
  hipLaunchKernelGGL((k_deposit<R, IZ, BD>), dim3((unsigned)grid), dim3(256), 0, s, L, boxes,
                     (const XT<IZ>*)xp, rhoc, off, mesh, mass, out, testrho);
  CUBE_HIP(hipGetLastError());
}

template <typename R, int IZ>
void launch_deposit(const PMLevel& L, const PMBox* boxes, int nbox, const void* xp, const int32_t* rhoc,
                    const int64_t* off, R* mesh, float mass, PMOut* out, int testrho, hipStream_t s) {
  if (std::is_same<R, float>::value && L.w > 1)
    launch_deposit_bd<R, IZ, 16>(L, boxes, nbox, xp, rhoc, off, mesh, mass, out, testrho, s);
  else
    launch_deposit_bd<R, IZ, 8>(L, boxes, nbox, xp, rhoc, off, mesh, mass, out, testrho, s);
}

template <typename R, int IZ, int BD>
void launch_interp_bd(const PMLevel& L, const PMBox* boxes, int nbox, const void* xp, float* vp, const int32_t* rhoc,
                      const int64_t* off, const R* phi, float a_mid, float dt, PMOut* out, hipStream_t s) {
  const int nblk = (L.nforce + BD - 1) / BD;
  const int64_t grid = (int64_t)nblk * nblk * nblk * nbox;
  hipLaunchKernelGGL((k_interp<R, IZ, BD>), dim3((unsigned)grid), dim3(256), 0, s, L, boxes,
                     (const XT<IZ>*)xp, vp, rhoc, off, phi, a_mid, dt, out);
  CUBE_HIP(hipGetLastError());
}

template <typename R, int IZ>
void launch_interp(const PMLevel& L, const PMBox* boxes, int nbox, const void* xp, float* vp, const int32_t* rhoc,
                   const int64_t* off, const R* phi, float a_mid, float dt, PMOut* out, hipStream_t s) {
  launch_interp_bd<R, IZ, 8>(L, boxes, nbox, xp, vp, rhoc, off, phi, a_mid, dt, out, s);
}

}  // namespace

template <typename R>
void pm_deposit(int izipx, const PMLevel& L, const PMBox* boxes, int nbox, const void* xp, const int32_t* rhoc,
                const int64_t* off, R* mesh, float mass, PMOut* out, hipStream_t s) {
  const int testrho = L.periodic;  // only the coarse level reports testrho
  if (izipx == 2) launch_deposit<R, 2>(L, boxes, nbox, xp, rhoc, off, mesh, mass, out, testrho, s);
  else launch_deposit<R, 1>(L, boxes, nbox, xp, rhoc, off, mesh, mass, out, testrho, s);
}

template <typename R>
void pm_interp(int izipx, const PMLevel& L, const PMBox* boxes, int nbox, const void* xp, float* vp,
               const int32_t* rhoc, const int64_t* off, const R* phi, float a_mid, float dt, PMOut* out,
               hipStream_t s) {
  if (izipx == 2) launch_interp<R, 2>(L, boxes, nbox, xp, vp, rhoc, off, phi, a_mid, dt, out, s);
  else launch_interp<R, 1>(L, boxes, nbox, xp, vp, rhoc, off, phi, a_mid, dt, out, s);
}

template void pm_deposit<float>(int, const PMLevel&, const PMBox*, int, const void*, const int32_t*, const int64_t*,
                                float*, float, PMOut*, hipStream_t);
template void pm_deposit<double>(int, const PMLevel&, const PMBox*, int, const void*, const int32_t*, const int64_t*,
                                 double*, float, PMOut*, hipStream_t);
template void pm_interp<float>(int, const PMLevel&, const PMBox*, int, const void*, float*, const int32_t*,
                               const int64_t*, const float*, float, float, PMOut*, hipStream_t);
template void pm_interp<double>(int, const PMLevel&, const PMBox*, int, const void*, float*, const int32_t*,
                                const int64_t*, const double*, float, float, PMOut*, hipStream_t);

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

void classify_subtiles(const Geo& g, const int32_t* rhoc, int32_t* ires, int32_t* napm, hipStream_t s) {
  zero_async(napm, sizeof(int32_t) * 7, s);
  hipLaunchKernelGGL(k_classify, dim3((unsigned)g.ns3), dim3(256), 0, s, g, rhoc, ires, napm);
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
