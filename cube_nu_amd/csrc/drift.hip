// drift.hip -- drift.f90 on the GPU, bit-exact.
//
// The reference fixes the order of particles inside a destination cell by
// its plane-colour OpenMP passes (drift.f90:46-49, 96-136): source planes are
// visited in passes ilayer = (k-(1-ncb)) mod nlayer, each plane by one thread
// in (j, i, l) order.  With the colour count nlayer = 2*ceil(dt_mid*vz_max/4)
// + 2, a destination cell receives from at most one plane per pass, so the
// order is the key (pass(k_src), k_src, j_src, i_src, l).
//
// GPU form (no atomics on any order):
//  1. shift pass: destination of every particle, range check, maximum cell
//     shift per axis, per-destination integer counts (atomic adds of
//     integers are order-independent);
//  2. rank pass (one wavefront per source cell): for each destination
//     direction d, the rank of each particle among the cell's earlier
//     particles with the same d (ballot + popcount), and the per-(cell, d)
//     count;
//  3. offset pass (one thread per destination cell): walks its neighbouring
//     source cells in key order and turns the (cell, d) counts into offsets;
//  4. scatter pass: out = base(dest) + offset(cell, d) + rank, with the
//     re-encode xp' = xp + nint(dt_mid*v/(x_resolution*ratio_cs)) wrapped to
//     izipx bytes (drift.f90:117).
#include "drift.h"

namespace cube {
namespace {

template <int IZ>
using XT = typename XP<IZ>::T;

}  // namespace

namespace {

template <int IZ>
__device__ inline void dest_cell(const DriftParams& P, const XT<IZ>* xp, const float* vp, int64_t ip,
                                 const int ijk[3], int g[3]) {
  for (int d = 0; d < 3; ++d) {
    const double wv = P.dt_mid * (double)vp[3 * ip + d];        // FP64 coefficient
    const double dx = wv * P.inv_rcs;                            // drift.f90:57 (exact: rcs = 2^k)
    const double xq = ((double)ijk[d] - 1.0) + XP<IZ>::off(xp, ip, d);  // drift.f90:59
    g[d] = (int)ceil(xq + dx);                                   // drift.f90:63
  }
}

__device__ inline void cell_of(const DriftParams& P, int64_t cglob, int* T, int ijk[3]) {
  // extended cell counts stay below 2^31 (64 tiles of nte^3 cells)
  const uint32_t nct = (uint32_t)(P.nte * P.nte * P.nte), cg = (uint32_t)cglob;
  *T = (int)(cg / nct);
  const int e = (int)(cg - (uint32_t)*T * nct);
  ijk[0] = e % P.nte + 1 - P.ncb;
  ijk[1] = (e / P.nte) % P.nte + 1 - P.ncb;
  ijk[2] = e / (P.nte * P.nte) + 1 - P.ncb;
}

__device__ inline bool physical(const DriftParams& P, const int g[3]) {
  return g[0] >= 1 && g[0] <= P.nt && g[1] >= 1 && g[1] <= P.nt && g[2] >= 1 && g[2] <= P.nt;
}

__device__ inline int dcode(const DriftParams& P, const int g[3], const int s[3]) {
  return ((g[2] - s[2] + P.M[2]) * P.D[1] + (g[1] - s[1] + P.M[1])) * P.D[0] + (g[0] - s[0] + P.M[0]);
}

__device__ inline int64_t pcell(const DriftParams& P, int T, const int g[3]) {
  return (((int64_t)T * P.nt + (g[2] - 1)) * P.nt + (g[1] - 1)) * P.nt + (g[0] - 1);
}

template <int IZ>
__global__ void k_drift_shift(DriftParams P, const XT<IZ>* __restrict__ xp, const float* __restrict__ vp,
                              const int32_t* __restrict__ rhoc, const int64_t* __restrict__ off, int64_t ncells,
                              int32_t* __restrict__ flags, int32_t* __restrict__ newcount) {
  const int64_t c = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (c >= ncells) return;
  int T, ijk[3];
  cell_of(P, c, &T, ijk);
  const int cnt = rhoc[c];
  const int64_t base = off[c];
  const int lo = 1 - 2 * P.ncb, hi = P.nt + 2 * P.ncb;
  int err = 0, ms[3] = {0, 0, 0};
  for (int l0 = 0; l0 < cnt; l0 += 64) {
    const int l = l0 + lane;
    int64_t key = -1;
    if (l < cnt) {
      int g[3];
      dest_cell<IZ>(P, xp, vp, base + l, ijk, g);
      for (int d = 0; d < 3; ++d)
        if (g[d] < lo || g[d] > hi) err = 1;
      if (physical(P, g)) {
        for (int d = 0; d < 3; ++d) ms[d] = max(ms[d], abs(g[d] - ijk[d]));
        key = pcell(P, T, g);
      }
    }
    // one integer atomic per distinct destination in the wavefront
    unsigned long long active = __ballot(key >= 0);
    while (active) {
      const int leader = __ffsll((long long)active) - 1;
      const int64_t lk = __shfl(key, leader, 64);
      const unsigned long long peers = __ballot(key == lk);
      if (lane == leader) atomicAdd(&newcount[lk], (int)__popcll(peers));
      active &= ~peers;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    err |= __shfl_down(err, o, 64);
    for (int d = 0; d < 3; ++d) ms[d] = max(ms[d], __shfl_down(ms[d], o, 64));
  }
  if (lane == 0 && cnt > 0) {
    if (err) atomicOr(&flags[0], DRIFT_OUT_OF_RANGE);
    for (int d = 0; d < 3; ++d) atomic_max_int(&flags[1 + d], ms[d]);
  }
}

// Rank pass.  COUNT = true also does the shift pass's work in the same read
// of the particles (range check, destination counts, largest shift), with
// P.M a bound given by the caller: a particle whose shift exceeds it takes no
// table entry and sets DRIFT_SHIFT_BOUND, after which the caller redoes the
// drift with the measured bound.
template <int IZ, bool COUNT>
__global__ void k_drift_rank(DriftParams P, const XT<IZ>* __restrict__ xp, const float* __restrict__ vp,
                             const int32_t* __restrict__ rhoc, const int64_t* __restrict__ off, int64_t c0,
                             int64_t ncells, int32_t* __restrict__ rank, int32_t* __restrict__ table,
                             int32_t* __restrict__ flags, int32_t* __restrict__ newcount) {
  const int64_t cl = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (cl >= ncells) return;
  const int64_t c = c0 + cl;
  int T, ijk[3];
  cell_of(P, c, &T, ijk);
  const int cnt = rhoc[c];
  const int64_t base = off[c];
  const int Dt = P.D[0] * P.D[1] * P.D[2];
  int32_t* tbl = table + cl * Dt;
  const int lo = 1 - 2 * P.ncb, hi = P.nt + 2 * P.ncb;
  int err = 0, ms[3] = {0, 0, 0};
  for (int l0 = 0; l0 < cnt; l0 += 64) {
    const int l = l0 + lane;
    int key = -1;
    int64_t pkey = -1;
    if (l < cnt) {
      int g[3];
      dest_cell<IZ>(P, xp, vp, base + l, ijk, g);
      if (COUNT)
        for (int d = 0; d < 3; ++d)
          if (g[d] < lo || g[d] > hi) err |= DRIFT_OUT_OF_RANGE;
      if (physical(P, g)) {
        bool fits = true;
        if (COUNT) {
          for (int d = 0; d < 3; ++d) {
            const int sd = abs(g[d] - ijk[d]);
            ms[d] = max(ms[d], sd);
            fits = fits && sd <= P.M[d];
          }
          pkey = pcell(P, T, g);
        }
        if (fits) key = dcode(P, g, ijk);
        else err |= DRIFT_SHIFT_BOUND;
      }
    }
    if (COUNT) {
      // one integer atomic per distinct destination in the wavefront
      unsigned long long act = __ballot(pkey >= 0);
      while (act) {
        const int leader = __ffsll((long long)act) - 1;
        const int64_t lk = __shfl(pkey, leader, 64);
        const unsigned long long peers = __ballot(pkey == lk);
        if (lane == leader) atomicAdd(&newcount[lk], (int)__popcll(peers));
        act &= ~peers;
      }
    }
    unsigned long long active = __ballot(key >= 0);
    int myrank = -1;
    while (active) {
      const int leader = __ffsll((long long)active) - 1;
      const int lk = __shfl(key, leader, 64);
      const unsigned long long peers = __ballot(key == lk);
      int bc = 0;
      if (lane == leader) bc = tbl[lk];
      bc = __shfl(bc, leader, 64);
      if (key == lk) myrank = bc + __popcll(peers & ((1ull << lane) - 1ull));
      if (lane == leader) tbl[lk] = bc + __popcll(peers);
      active &= ~peers;
    }
    if (l < cnt) rank[base + l] = myrank;
  }
  if (COUNT) {
    for (int o = 32; o > 0; o >>= 1) {
      err |= __shfl_down(err, o, 64);
      for (int d = 0; d < 3; ++d) ms[d] = max(ms[d], __shfl_down(ms[d], o, 64));
    }
    if (lane == 0 && cnt > 0) {
      if (err) atomicOr(&flags[0], err);
      for (int d = 0; d < 3; ++d) atomic_max_int(&flags[1 + d], ms[d]);
    }
  }
}

__global__ void k_drift_offsets(DriftParams P, int t0, int t1, int32_t* __restrict__ table,
                                const int32_t* __restrict__ newcount, int32_t* __restrict__ flags) {
  const int64_t nt3 = (int64_t)P.nt * P.nt * P.nt;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)(t1 - t0) * nt3) return;
  const int T = t0 + (int)(t / nt3);
  const int q = (int)(t % nt3);
  const int g[3] = {q % P.nt + 1, (q / P.nt) % P.nt + 1, q / (P.nt * P.nt) + 1};
  const int lo = 1 - P.ncb, hi = P.nt + P.ncb;
  const int Dt = P.D[0] * P.D[1] * P.D[2];
  // source planes in pass order (pass(k) = (k - (1-ncb)) mod nlayer), then k
  int ks[2 * kMaxShift + 1];
  int nk = 0;
  for (int k = max(lo, g[2] - P.M[2]); k <= min(hi, g[2] + P.M[2]); ++k) ks[nk++] = k;
  for (int a = 1; a < nk; ++a) {
    int v = ks[a], pv = (v - lo) % P.nlayer, b = a - 1;
    while (b >= 0) {
      int pb = (ks[b] - lo) % P.nlayer;
      if (pb > pv || (pb == pv && ks[b] > v)) {
        ks[b + 1] = ks[b];
        --b;
      } else
        break;
    }
    ks[b + 1] = v;
  }
  const int64_t nct = (int64_t)P.nte * P.nte * P.nte;
  int running = 0;
  for (int a = 0; a < nk; ++a) {
    const int k = ks[a];
    for (int j = max(lo, g[1] - P.M[1]); j <= min(hi, g[1] + P.M[1]); ++j)
      for (int i = max(lo, g[0] - P.M[0]); i <= min(hi, g[0] + P.M[0]); ++i) {
        const int s[3] = {i, j, k};
        const int64_t cl = (int64_t)(T - t0) * nct + rc_index(i, j, k, 0, P.nte, P.ncb);
        const int64_t idx = cl * Dt + dcode(P, g, s);
        const int c = table[idx];
        table[idx] = running;
        running += c;
      }
  }
  if (running != newcount[pcell(P, T, g)]) atomicOr(&flags[0], 4);
}

template <int IZ>
__global__ void k_drift_scatter(DriftParams P, const XT<IZ>* __restrict__ xp, const float* __restrict__ vp,
                                const int64_t* __restrict__ pid, const int32_t* __restrict__ rhoc,
                                const int64_t* __restrict__ off, int64_t c0, int64_t ncells,
                                const int32_t* __restrict__ rank, const int32_t* __restrict__ table,
                                const int64_t* __restrict__ base_out, XT<IZ>* __restrict__ xo,
                                float* __restrict__ vo, int64_t* __restrict__ po, int32_t* __restrict__ flags) {
  const int64_t cl = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (cl >= ncells) return;
  const int64_t c = c0 + cl;
  int T, ijk[3];
  cell_of(P, c, &T, ijk);
  const int cnt = rhoc[c];
  const int64_t base = off[c];
  const int Dt = P.D[0] * P.D[1] * P.D[2];
  // 1 / (x_resolution rcs): both powers of two, so wv * shs == wv / (xr rcs)
  const double shs = P.inv_rcs * (double)(1 << XP<IZ>::B);
  for (int l = lane; l < cnt; l += 64) {
    const int64_t ip = base + l;
    const int r = rank[ip];
    if (r < 0) continue;
    int g[3];
    dest_cell<IZ>(P, xp, vp, ip, ijk, g);
    const int64_t pos = base_out[pcell(P, T, g)] + table[cl * Dt + dcode(P, g, ijk)] + r;
    for (int d = 0; d < 3; ++d) {
      const double wv = P.dt_mid * (double)vp[3 * ip + d];
      const double sh = wv * shs;  // drift.f90:117
      const long long dn = (long long)round(sh);  // nint: half away from zero
      const long long nv = (long long)xp[3 * ip + d] + dn;
      if (IZ == 2) xo[3 * pos + d] = (int16_t)(uint16_t)(nv & 0xffff);
      else xo[3 * pos + d] = (int8_t)(uint8_t)(nv & 0xff);
      vo[3 * pos + d] = vp[3 * ip + d];
    }
    const int64_t pv = pid ? pid[ip] : 1;
    if (po) po[pos] = pv;
    if (pv <= 0) atomicOr(&flags[0], DRIFT_PID_ZERO);
  }
}

__global__ void k_drift_rhoc(DriftParams P, int64_t ncells, const int32_t* __restrict__ newcount,
                             int32_t* __restrict__ rhoc) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  int T, ijk[3];
  cell_of(P, c, &T, ijk);
  rhoc[c] = physical(P, ijk) ? newcount[pcell(P, T, ijk)] : 0;
}

__global__ void k_tile_check(const int64_t* __restrict__ off, int64_t nct, int ntiles, int64_t cap,
                             int32_t* __restrict__ flags) {
  for (int T = threadIdx.x; T < ntiles; T += blockDim.x) {
    const int64_t n = off[(int64_t)(T + 1) * nct] - off[(int64_t)T * nct];
    atomic_max_int(&flags[4], (int32_t)(n < (int64_t)INT32_MAX ? n : (int64_t)INT32_MAX));
    if (n > cap) atomicOr(&flags[0], DRIFT_TILE_OVERFLOW);
  }
}

__global__ void k_buffer_rhoc(Geo g, const int32_t* __restrict__ rp, int32_t* __restrict__ re,
                              int32_t* __restrict__ physcount) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.ncells_ext) return;
  const int64_t T = c / g.ncell_tile, e = c % g.ncell_tile;
  const int64_t it[3] = {T % g.nnt, (T / g.nnt) % g.nnt, T / (g.nnt * g.nnt)};
  const int64_t ijk[3] = {e % g.nte + 1 - g.ncb, (e / g.nte) % g.nte + 1 - g.ncb, e / (g.nte * g.nte) + 1 - g.ncb};
  int64_t st[3], sc[3];
  for (int d = 0; d < 3; ++d) {
    int64_t gc = it[d] * g.nt + ijk[d] - 1;
    gc = ((gc % g.nc) + g.nc) % g.nc;
    st[d] = gc / g.nt;
    sc[d] = gc % g.nt + 1;
  }
  const int64_t Ts = (st[2] * g.nnt + st[1]) * g.nnt + st[0];
  const int32_t v = rp[rc_index(sc[0], sc[1], sc[2], Ts, g.nte, g.ncb)];
  re[c] = v;
  const bool phys = ijk[0] >= 1 && ijk[0] <= g.nt && ijk[1] >= 1 && ijk[1] <= g.nt && ijk[2] >= 1 && ijk[2] <= g.nt;
  if (phys) physcount[((T * g.nt + ijk[2] - 1) * g.nt + ijk[1] - 1) * g.nt + ijk[0] - 1] = v;
}


}  // namespace

void drift_tile_check(const int64_t* off, int64_t ncell_tile, int ntiles, int64_t np_tile_max, int32_t* flags,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_tile_check, dim3(1), dim3(256), 0, s, off, ncell_tile, ntiles, np_tile_max, flags);
  CUBE_HIP(hipGetLastError());
}

void drift_shift(const DriftParams& P, const void* xp, const float* vp, const int32_t* rhoc, const int64_t* off,
                 int64_t ncells, DriftScratch& S, hipStream_t s) {
  const int64_t nthr = ncells * 64;
  if (P.izipx == 2)
    hipLaunchKernelGGL(k_drift_shift<2>, dim3((unsigned)ceil_div(nthr, 256)), dim3(256), 0, s, P,
                       (const int16_t*)xp, vp, rhoc, off, ncells, S.flags, S.newcount);
  else
    hipLaunchKernelGGL(k_drift_shift<1>, dim3((unsigned)ceil_div(nthr, 256)), dim3(256), 0, s, P, (const int8_t*)xp,
                       vp, rhoc, off, ncells, S.flags, S.newcount);
  CUBE_HIP(hipGetLastError());
}

void drift_rank(const DriftParams& P, const void* xp, const float* vp, const int32_t* rhoc, const int64_t* off,
                int t0, int t1, DriftScratch& S, hipStream_t s, bool count) {
  const int64_t nct = (int64_t)P.nte * P.nte * P.nte;
  const int64_t ncells = (int64_t)(t1 - t0) * nct;
  const dim3 grid((unsigned)ceil_div(ncells * 64, 256));
  const int64_t c0 = (int64_t)t0 * nct;
  if (P.izipx == 2) {
    if (count)
      hipLaunchKernelGGL((k_drift_rank<2, true>), grid, dim3(256), 0, s, P, (const int16_t*)xp, vp, rhoc, off, c0,
                         ncells, S.rank, S.table, S.flags, S.newcount);
    else
      hipLaunchKernelGGL((k_drift_rank<2, false>), grid, dim3(256), 0, s, P, (const int16_t*)xp, vp, rhoc, off, c0,
                         ncells, S.rank, S.table, S.flags, S.newcount);
  } else {
    if (count)
      hipLaunchKernelGGL((k_drift_rank<1, true>), grid, dim3(256), 0, s, P, (const int8_t*)xp, vp, rhoc, off, c0,
                         ncells, S.rank, S.table, S.flags, S.newcount);
    else
      hipLaunchKernelGGL((k_drift_rank<1, false>), grid, dim3(256), 0, s, P, (const int8_t*)xp, vp, rhoc, off, c0,
                         ncells, S.rank, S.table, S.flags, S.newcount);
  }
  CUBE_HIP(hipGetLastError());
}

void drift_offsets(const DriftParams& P, int t0, int t1, DriftScratch& S, hipStream_t s) {
  const int64_t n = (int64_t)(t1 - t0) * P.nt * P.nt * P.nt;
  hipLaunchKernelGGL(k_drift_offsets, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, P, t0, t1, S.table,
                     S.newcount, S.flags);
  CUBE_HIP(hipGetLastError());
}

void drift_scatter(const DriftParams& P, const void* xp, const float* vp, const int64_t* pid, const int32_t* rhoc,
                   const int64_t* off, int t0, int t1, DriftScratch& S, void* xp_out, float* vp_out,
                   int64_t* pid_out, hipStream_t s) {
  const int64_t nct = (int64_t)P.nte * P.nte * P.nte;
  const int64_t ncells = (int64_t)(t1 - t0) * nct;
  const int64_t nthr = ncells * 64;
  if (P.izipx == 2)
    hipLaunchKernelGGL(k_drift_scatter<2>, dim3((unsigned)ceil_div(nthr, 256)), dim3(256), 0, s, P,
                       (const int16_t*)xp, vp, pid, rhoc, off, (int64_t)t0 * nct, ncells, S.rank, S.table, S.base,
                       (int16_t*)xp_out, vp_out, pid_out, S.flags);
  else
    hipLaunchKernelGGL(k_drift_scatter<1>, dim3((unsigned)ceil_div(nthr, 256)), dim3(256), 0, s, P,
                       (const int8_t*)xp, vp, pid, rhoc, off, (int64_t)t0 * nct, ncells, S.rank, S.table, S.base,
                       (int8_t*)xp_out, vp_out, pid_out, S.flags);
  CUBE_HIP(hipGetLastError());
}

void drift_write_rhoc(const DriftParams& P, const DriftScratch& S, int32_t* rhoc_out, hipStream_t s) {
  const int64_t nct = (int64_t)P.nte * P.nte * P.nte;
  const int64_t ncells = (int64_t)P.nnt * P.nnt * P.nnt * nct;
  hipLaunchKernelGGL(k_drift_rhoc, dim3((unsigned)ceil_div(ncells, 256)), dim3(256), 0, s, P, ncells, S.newcount,
                     rhoc_out);
  CUBE_HIP(hipGetLastError());
}

void buffer_rhoc(const Geo& g, const int32_t* rhoc_phys, int32_t* rhoc_ext, int32_t* physcount, hipStream_t s) {
  hipLaunchKernelGGL(k_buffer_rhoc, dim3((unsigned)ceil_div(g.ncells_ext, 256)), dim3(256), 0, s, g, rhoc_phys,
                     rhoc_ext, physcount);
  CUBE_HIP(hipGetLastError());
}

namespace {

// nn=1 buffer copy by rows: an x-row of an extended tile is three runs of
// consecutive cells, each a contiguous run of particles in the source
// layout (cells 1..nt of the tile, the last ncb cells of the periodic left
// neighbour, the first ncb of the right one) and in the extended layout.
// One workgroup per (tile, z, y) row copies the three runs with coalesced,
// independent loads (xp as its 16- or 8-bit elements, vp as floats, pid as
// 64-bit words), so the copy streams instead of waiting once per cell.
template <int IZ>
__global__ void __launch_bounds__(256) k_buffer_rows(Geo g, const int64_t* __restrict__ eoff,
                                                     const int64_t* __restrict__ poff, const XT<IZ>* __restrict__ xp,
                                                     const float* __restrict__ vp, const int64_t* __restrict__ pid,
                                                     XT<IZ>* __restrict__ xo, float* __restrict__ vo,
                                                     int64_t* __restrict__ po) {
  const int nte = (int)g.nte, nt = (int)g.nt, nc = (int)g.nc, nnt = (int)g.nnt, ncb = (int)g.ncb;
  const int64_t r = blockIdx.x;  // (T, ze, ye) with ye fastest
  const int ye = (int)(r % nte), ze = (int)((r / nte) % nte), T = (int)(r / ((int64_t)nte * nte));
  const int it[3] = {T % nnt, (T / nnt) % nnt, T / (nnt * nnt)};
  // source tile and cell of the row's y and z (periodic)
  int sy, sz, ty, tz;
  {
    int gy = it[1] * nt + (ye + 1 - ncb) - 1;
    gy = gy < 0 ? gy + nc : (gy >= nc ? gy - nc : gy);
    ty = gy / nt;
    sy = gy - ty * nt + 1;
    int gz = it[2] * nt + (ze + 1 - ncb) - 1;
    gz = gz < 0 ? gz + nc : (gz >= nc ? gz - nc : gz);
    tz = gz / nt;
    sz = gz - tz * nt + 1;
  }
  const int64_t erow = ((int64_t)T * nte + ze) * nte + ye;  // ext cell index of x_e = 0 is erow * nte
  for (int seg = 0; seg < 3; ++seg) {
    // ext x range [xe0, xe1) and its source tile along x
    const int xe0 = seg == 0 ? 0 : (seg == 1 ? ncb : ncb + nt);
    const int xe1 = seg == 0 ? ncb : (seg == 1 ? ncb + nt : nte);
    int gx = it[0] * nt + (xe0 + 1 - ncb) - 1;
    gx = gx < 0 ? gx + nc : (gx >= nc ? gx - nc : gx);
    const int tx = gx / nt, sx0 = gx - tx * nt + 1;  // first source cell (1-based)
    const int Ts = (tz * nnt + ty) * nnt + tx;
    const int64_t c0 = (((int64_t)Ts * nt + sz - 1) * nt + sy - 1) * nt + sx0 - 1;
    const int64_t src = poff[c0];
    const int64_t n = poff[c0 + (xe1 - xe0)] - src;
    const int64_t dst = eoff[erow * nte + xe0];
    for (int64_t q = threadIdx.x; q < 3 * n; q += 256) {
      xo[3 * dst + q] = xp[3 * src + q];
      vo[3 * dst + q] = vp[3 * src + q];
    }
    if (pid)
      for (int64_t q = threadIdx.x; q < n; q += 256) po[dst + q] = pid[src + q];
  }
}

// buffer_v at nn=1 (buffer_particle.f90:11-27, buffer_xvp(.false., .true.)):
// in the extended layout, the velocities (and IDs) of every buffer cell are
// overwritten by those of the physical cell it is the periodic image of.  The
// same three runs per (tile, z, y) row as k_buffer_rows, read from the
// physical cells of the extended array itself: sources are physical cells and
// destinations buffer cells, so no run is both.  A run copies the smaller of
// its source and destination counts (equal in a layout made by the buffer
// step), so every access stays inside both runs.
template <bool PID>
__global__ void __launch_bounds__(256) k_buffer_v_rows(Geo g, const int64_t* __restrict__ eoff, float* vp,
                                                       int64_t* pid) {
  const int nte = (int)g.nte, nt = (int)g.nt, nc = (int)g.nc, nnt = (int)g.nnt, ncb = (int)g.ncb;
  const int64_t r = blockIdx.x;  // (T, ze, ye) with ye fastest
  const int ye = (int)(r % nte), ze = (int)((r / nte) % nte), T = (int)(r / ((int64_t)nte * nte));
  const int it[3] = {T % nnt, (T / nnt) % nnt, T / (nnt * nnt)};
  const bool yz_phys = ye >= ncb && ye < ncb + nt && ze >= ncb && ze < ncb + nt;
  int sy, sz, ty, tz;
  {
    int gy = it[1] * nt + (ye + 1 - ncb) - 1;
    gy = gy < 0 ? gy + nc : (gy >= nc ? gy - nc : gy);
    ty = gy / nt;
    sy = gy - ty * nt + 1;
    int gz = it[2] * nt + (ze + 1 - ncb) - 1;
    gz = gz < 0 ? gz + nc : (gz >= nc ? gz - nc : gz);
    tz = gz / nt;
    sz = gz - tz * nt + 1;
  }
  const int64_t erow = ((int64_t)T * nte + ze) * nte + ye;
  for (int seg = 0; seg < 3; ++seg) {
    if (seg == 1 && yz_phys) continue;  // the row's physical cells are their own source
    const int xe0 = seg == 0 ? 0 : (seg == 1 ? ncb : ncb + nt);
    const int xe1 = seg == 0 ? ncb : (seg == 1 ? ncb + nt : nte);
    int gx = it[0] * nt + (xe0 + 1 - ncb) - 1;
    gx = gx < 0 ? gx + nc : (gx >= nc ? gx - nc : gx);
    const int tx = gx / nt, sx0 = gx - tx * nt + 1;
    const int Ts = (tz * nnt + ty) * nnt + tx;
    const int64_t s0 = (((int64_t)Ts * nte + (sz - 1 + ncb)) * nte + (sy - 1 + ncb)) * nte + (sx0 - 1 + ncb);
    const int64_t src = eoff[s0], nsrc = eoff[s0 + (xe1 - xe0)] - src;
    const int64_t dst = eoff[erow * nte + xe0], ndst = eoff[erow * nte + xe1] - dst;
    const int64_t n = nsrc < ndst ? nsrc : ndst;
    for (int64_t q = threadIdx.x; q < 3 * n; q += 256) vp[3 * dst + q] = vp[3 * src + q];
    if (PID)
      for (int64_t q = threadIdx.x; q < n; q += 256) pid[dst + q] = pid[src + q];
  }
}

}  // namespace

void buffer_v(const Geo& g, const int64_t* ext_off, float* vp, int64_t* pid, hipStream_t s) {
  const int64_t rows = g.ntiles * g.nte * g.nte;
  if (pid)
    hipLaunchKernelGGL(k_buffer_v_rows<true>, dim3((unsigned)rows), dim3(256), 0, s, g, ext_off, vp, pid);
  else
    hipLaunchKernelGGL(k_buffer_v_rows<false>, dim3((unsigned)rows), dim3(256), 0, s, g, ext_off, vp, pid);
  CUBE_HIP(hipGetLastError());
}

void buffer_copy(const Geo& g, const int32_t* rhoc_ext, const int64_t* ext_off, const int64_t* phys_off,
                 const void* xp, const float* vp, const int64_t* pid, void* xp_out, float* vp_out, int64_t* pid_out,
                 hipStream_t s) {
  (void)rhoc_ext;
  const int64_t rows = g.ntiles * g.nte * g.nte;
  if (g.izipx == 2)
    hipLaunchKernelGGL(k_buffer_rows<2>, dim3((unsigned)rows), dim3(256), 0, s, g, ext_off, phys_off,
                       (const int16_t*)xp, vp, pid, (int16_t*)xp_out, vp_out, pid_out);
  else
    hipLaunchKernelGGL(k_buffer_rows<1>, dim3((unsigned)rows), dim3(256), 0, s, g, ext_off, phys_off,
                       (const int8_t*)xp, vp, pid, (int8_t*)xp_out, vp_out, pid_out);
  CUBE_HIP(hipGetLastError());
}

}  // namespace cube
