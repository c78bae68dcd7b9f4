// pp.hip -- PP_Force (kick.f90:376-588) on the GPU.
//
// Binning: every PP cell (1/ratio_sf of a coarse cell) is nested in exactly
// one coarse cell, so one wavefront per (sub-tile, coarse cell) owns the PP
// cells it fills: counts and in-cell order are deterministic without
// atomics, and the in-cell order is the reference's (coarse-cell particle
// order).
//
// Force: a persistent kernel whose workgroups claim bricks of PB^3 physical
// PP cells from a work queue; each thread owns target particles and sums the
// pair force from the 27 neighbouring PP cells in a fixed order.  Both
// members of a physical pair evaluate the pair, so no cross-thread
// accumulation is needed and the result is deterministic.
// The in-range test uses the reference's FP32 operations on exact FP32
// positions, so the set of interacting pairs is the reference's.
#include "pm.h"
#include "pp.h"

namespace cube {
namespace {

template <int IZ>
using XT = typename XP<IZ>::T;

// PP cell of a particle, kick.f90:420-431 in FP32, and its x sub-bin: the
// PP cell idx covers xv in [idx-1, idx), and sub-bin q = floor(kSX * frac(xv))
// (exact: xv - floor(xv) and the power-of-two product are exact in FP32).
template <int IZ>
__device__ inline bool pp_cell(const PPLevel& L, const PPBox& B, int ci, int cj, int ck, const XT<IZ>* xp, int64_t ip,
                               float xf[3], int idx[3], int& q) {
  const int ijk[3] = {ci, cj, ck};
  bool in = true;
  for (int d = 0; d < 3; ++d) {
    float xvec = (float)((double)(ijk[d] - 1) + XP<IZ>::off(xp, ip, d));
    xf[d] = xvec * (float)L.rcs;
    float xv = (float)L.sf * (xvec - (float)(B.ifs[d] * L.ntt));
    const float fl = floorf(xv);
    idx[d] = (int)fl + 1;
    if (d == 0) q = min((int)((xv - fl) * (float)kSX), kSX - 1);
    in = in && idx[d] >= 0 && idx[d] <= L.npgrid + 1;
  }
  return in;
}

// One wavefront per (box, coarse cell of the sub-tile + 1-cell ring), one
// particle per lane.  A PP cell lies inside one coarse cell (idx =
// floor(sf*(xvec - ifs*ntt)) + 1 with xvec < c), so the wavefront owns the
// counts of the nsb = sf^3 * kSX sub-bins of its coarse cell and keeps them in
// LDS (dynamic, nsb ints per wavefront).  The lanes holding the same sub-bin
// key are found with one ballot per key bit; a lane's rank among them is a
// population count, so within a sub-bin the order is the coarse-cell particle
// order, and the lowest such lane advances the sub-bin's LDS counter (one
// writer per key).  FILL=0 writes the counts, FILL=1 the entries at
// start + rank.  Sub-bins are numbered ((z*m + y)*m + x)*kSX + q: each (y, z)
// row of the PP grid is one contiguous, x-ordered run of the CSR.
template <int IZ, int FILL>
__global__ void __launch_bounds__(256) k_pp_bin(PPLevel L, const PPBox* __restrict__ boxes, int nbox,
                                                const XT<IZ>* __restrict__ xp, const int32_t* __restrict__ rhoc,
                                                const int64_t* __restrict__ off, int32_t* __restrict__ count,
                                                const int64_t* __restrict__ start, float4* __restrict__ ent) {
  extern __shared__ int lcnt[];
  const int nr = L.ntt + 2;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t wv = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (wv >= (int64_t)nbox * nr * nr * nr) return;
  const int ib = (int)(wv / (nr * nr * nr));
  const int qc = (int)(wv % (nr * nr * nr));
  const PPBox B = boxes[ib];
  const int c[3] = {B.ifs[0] * L.ntt + (qc % nr), B.ifs[1] * L.ntt + (qc / nr) % nr,
                    B.ifs[2] * L.ntt + qc / (nr * nr)};  // nc1-1 + offset (Fortran nc1 = ifs*ntt+1)
  const int64_t cell = rc_index(c[0], c[1], c[2], B.T, L.nte, L.ncb);
  const int cnt = rhoc[cell];
  const int64_t base = off[cell];
  const int m = L.npgrid + 2;
  const int sf = L.sf, nsb = sf * sf * sf * kSX;
  const int kbits = 32 - __clz(nsb - 1);  // bits of a sub-bin key
  int lo[3];  // PP index of the coarse cell's first PP cell per axis
  for (int d = 0; d < 3; ++d) lo[d] = sf * (c[d] - 1 - B.ifs[d] * L.ntt) + 1;
  int* lc = lcnt + wave * nsb;
  for (int t = lane; t < nsb; t += 64) lc[t] = 0;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int64_t cbase = (int64_t)ib * m * m * m;
  const unsigned long long lt = (1ull << lane) - 1ull;
  for (int l0 = 0; l0 < cnt; l0 += 64) {
    const int l = l0 + lane;
    float xf[3] = {0, 0, 0};
    int idx[3] = {0, 0, 0};
    int q = 0;
    bool in = false;
    if (l < cnt) in = pp_cell<IZ>(L, B, c[0], c[1], c[2], xp, base + l, xf, idx, q);
    const int key = in ? (((idx[2] - lo[2]) * sf + (idx[1] - lo[1])) * sf + (idx[0] - lo[0])) * kSX + q : 0;
    unsigned long long eq = __ballot(in);
    for (int bt = 0; bt < kbits; ++bt) {
      const unsigned long long mb = __ballot((key >> bt) & 1);
      eq &= ((key >> bt) & 1) ? mb : ~mb;
    }
    const int prev = in ? lc[key] : 0;
    if (in) {
      const int rank = (int)__popcll(eq & lt);
      if (FILL) {
        const int64_t pc = cbase + ((int64_t)idx[2] * m + idx[1]) * m + idx[0];
        ent[start[pc * kSX + q] + prev + rank] =
            make_float4(xf[0], xf[1], xf[2], __uint_as_float((unsigned)(base + l)));
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (in && (eq & lt) == 0) lc[key] = prev + (int)__popcll(eq);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  if (!FILL) {
    for (int t = lane; t < nsb; t += 64) {
      const int n = lc[t];
      if (n == 0) continue;
      const int pc = t / kSX;
      const int ix = lo[0] + pc % sf, iy = lo[1] + (pc / sf) % sf, iz = lo[2] + pc / (sf * sf);
      count[(cbase + ((int64_t)iz * m + iy) * m + ix) * kSX + t % kSX] = n;
    }
  }
}

// Physical PP cells per brick per dim: PB = 4 or 8, chosen with the load
// depth by the engine's autotuner (a target's sums do not depend on the
// brick it is in, so neither choice changes a result).
// Lanes per target in the FP32 force kernel: a target's candidates are dealt
// to kPPGroup lanes in turn and their partial sums combined at the end (26.1
// ms at 512^3 against 27.3 with one lane per target, 27.2 with four, 31.5
// with eight: pairs of lanes read consecutive entries and wait less on the
// longest column of the wavefront).
constexpr int kPPGroup = 2;
// Candidate loads in flight per lane: U in {4, 6, 8}, chosen per PP level at
// run time by the engine's autotuner (every U sums the pairs in the same
// order, so the choice never changes a result).


// F(r, r_pp) = c Q(ep), ep = 2r/r_pp < 2, with Q(ep) = pi(ep) for ep <= 1 and
// 12/ep^2 + po(ep) above (the two polynomials of F_ra, kick.f90:360-374).
// The FP32 force takes Q from a cubic Hermite table over kQSeg segments of
// [0, 2] kept in LDS: ep = 1 is a segment edge and Q is C^1 there, so every
// segment interpolates one smooth branch; the table's error is below 2e-5 on
// |Q| <= 104 (h^4 max|Q^(4)| / 384 with h = 1/32), i.e. below 1e-6 of
// 1/r_pp^2 in the pair force, at the level of FP32 rounding.
constexpr int kQSeg = 64;

__device__ inline void q_branch(double e, bool outer, double& q, double& dq) {
  if (outer) {
    q = 12.0 / (e * e) + (((((7.0 * e - 48.0) * e + 70.0) * e + 224.0) * e - 840.0) * e + 896.0) * e - 224.0;
    dq = -24.0 / (e * e * e) + ((((42.0 * e - 240.0) * e + 280.0) * e + 672.0) * e - 1680.0) * e + 896.0;
  } else {
    q = e * (224.0 + e * e * (-224.0 + e * (70.0 + e * (48.0 - 21.0 * e))));
    dq = 224.0 + e * e * (-672.0 + e * (280.0 + e * (240.0 - 126.0 * e)));
  }
}

// Segment s of the table: Q(s h + f h) = ((c.w f + c.z) f + c.y) f + c.x.
__device__ inline void fill_qtab(float4* tab) {
  for (int sg = threadIdx.x; sg < kQSeg; sg += blockDim.x) {
    const double h = 2.0 / kQSeg;
    const bool outer = sg >= kQSeg / 2;
    double y0, d0, y1, d1;
    q_branch(sg * h, outer, y0, d0);
    q_branch((sg + 1) * h, outer, y1, d1);
    d0 *= h;
    d1 *= h;
    tab[sg] = make_float4((float)y0, (float)d0, (float)(3.0 * (y1 - y0) - 2.0 * d0 - d1),
                          (float)(2.0 * (y0 - y1) + d0 + d1));
  }
}

// FP32 pair law from the table: F(r, app) - F(r, r_pp) = 1/r^2 - c Q(ep) for
// r > app; the rare r <= app pair takes F(r, app) in full.
__device__ inline float pair_law_tab(const PPLevel& L, const float4* qtab, float r2, float& rinv_out) {
  const float rinv = __builtin_amdgcn_rsqf(r2);
  const float r = r2 * rinv;
  const float t = r * (L.two_inv_range * (0.5f * kQSeg));
  const int i = min((int)t, kQSeg - 1);
  const float f = t - (float)i;
  const float4 c = qtab[i];
  const float q = fmaf(fmaf(fmaf(c.w, f, c.z), f, c.y), f, c.x);
  float s = fmaf(-L.c_range, q, rinv * rinv);
  if (r2 <= L.app2) {  // ep(r, app) <= 2
    const float ea = (2.0f * r) * L.inv_app;
    const float a2 = ea * ea;
    const float fa =
        ea > 1.0f
            ? L.c_app * (12.0f * __builtin_amdgcn_rcpf(a2) +
                         fmaf(ea, fmaf(ea, fmaf(ea, fmaf(ea, fmaf(ea, fmaf(ea, 7.0f, -48.0f), 70.0f), 224.0f),
                                                -840.0f),
                                       896.0f),
                              -224.0f))
            : L.c_app * ea * fmaf(a2, fmaf(ea, fmaf(ea, fmaf(ea, -21.0f, 48.0f), 70.0f), -224.0f), 224.0f);
    s = fmaf(-L.c_range, q, fa);
  }
  rinv_out = rinv;
  return s;
}

// Pair force of candidate o on target me (kick.f90:523-530, 554-560):
// f1 += (x2-x1)/r * (F(r,app) - F(r,r_pp)).
// FP32 form: r^-1 from the hardware reciprocal square root, F(r, app) = 1/r^2
// (r > app) and F(r, r_pp) from the LDS table (pair_law_tab); the error of a
// pair term stays at the level of FP32 rounding.  FP64 form: the reference's
// operations in FP64.
template <typename R>
__device__ inline void pair_force(const PPLevel& L, const float4* qtab, const float4 me, const float4 o, R f1[3]);

template <>
__device__ inline void pair_force<float>(const PPLevel& L, const float4* qtab, const float4 me, const float4 o,
                                         float f1[3]) {
  const float rv0 = o.x - me.x, rv1 = o.y - me.y, rv2 = o.z - me.z;
  const float r2 = (rv0 * rv0 + rv1 * rv1) + rv2 * rv2;
  float rinv;
  const float s = pair_law_tab(L, qtab, r2, rinv);
  const float sc = rinv * s;
  f1[0] = fmaf(rv0, sc, f1[0]);
  f1[1] = fmaf(rv1, sc, f1[1]);
  f1[2] = fmaf(rv2, sc, f1[2]);
}

template <>
__device__ inline void pair_force<double>(const PPLevel& L, const float4*, const float4 me, const float4 o,
                                          double f1[3]) {
  const double rv0 = (double)o.x - me.x, rv1 = (double)o.y - me.y, rv2 = (double)o.z - me.z;
  const double r2 = (rv0 * rv0 + rv1 * rv1) + rv2 * rv2;
  const double r = sqrt(r2);
  const double fr = f_ra<double>(r, (double)L.app, f_ra_coef<double>((double)L.app)) -
                    f_ra<double>(r, (double)L.pp_range, f_ra_coef<double>((double)L.pp_range));
  f1[0] = f1[0] + (rv0 / r) * fr;
  f1[1] = f1[1] + (rv1 / r) * fr;
  f1[2] = f1[2] + (rv2 / r) * fr;
}

// Persistent: each workgroup claims bricks of PB^3 physical PP cells from one
// of kPPClaimCounters counters (workgroup i uses counter i % 8, which hands
// out the bricks of sub-tiles i % 8, i % 8 + 8, ..., one sub-tile at a time,
// so that a sub-tile's bricks run on one XCD and share its L2; with fewer
// than 64 sub-tiles in a launch, bricks 8c + i % 8, so that every counter has
// work) until they run out.  The host orders the boxes
// densest sub-tile first (kick.f90:132-138 sorts by npc_max for
// schedule(dynamic,1), kick.f90:285), so every counter drains the heaviest
// bricks first and the launch does not end on a few heavy workgroups.  In a
// brick each group of G lanes owns a target particle and sums the pair forces
// of the 27 neighbouring PP cells in the fixed (dz, dy, x) order, lane j of
// the group taking candidates j, j + G, ... of each run; the G partial sums
// are added by a fixed butterfly, so the result is deterministic.  PP cells
// are split into kSX sub-bins along x and
// the CSR is x-ordered along each (y, z) row, so the part of a (dy, dz)
// column that can hold partners, x within x_t +- sqrt(r_pp^2 - e^2) for the
// column's nearest y/z distance e from the target, is one contiguous run
// found from two sub-bin starts; columns with e > r_pp are skipped.  Both windows carry a margin far
// above FP32 rounding, so every pair the reference tests in range is tested
// here with the reference's FP32 operations and the interacting set is the
// reference's.  The sub-bin starts of the (PB+2)^3 neighbourhood are staged
// in LDS; candidates are read U at a time from L1/L2 (the entry array is
// padded by kPPPad > (U-1) G entries so the reads need no clamp).
// Measured at 512^3 (texture address and data units ~90-100% busy, VALU ~65%
// busy in the chosen form), slower variants: a two-pass form
// (in-range tests of 32 candidates into a bit mask, then the pair law over
// the set bits only, each partner read again) measured 37.9 against 27.2 ms
// at 512^3: the second pass's reads come from L2, one per pair.  Staging
// each brick's neighbourhood (up to 2560 entries) in LDS and reading the
// candidates from there measured 50.3 ms: 50 KB of LDS per workgroup leaves
// three wavefronts per SIMD, and the lanes' scattered 16-byte LDS reads
// conflict.  8-byte entries (positions as exact integers in units of 2^-15
// fine cells, differences formed exactly through 2^23 + n in FP32) halved the
// load instructions but added ~30% VALU work: 30.3 ms.  Deferring the pair
// law until some lane holds two pairs (so it runs on most lanes at once)
// cut the VALU count but its branches and spills gave 32.4 ms.  Splitting PP
// cells also along y (fewer candidates, twice the columns) gave 30.2 ms.
// FP32 form: eight 256-thread workgroups per CU (the launch bound's second
// value), i.e. at most 64 VGPRs and eight wavefronts per SIMD to cover the
// candidate loads' latency (28.5 vs 30.2 ms at 512^3 with 71 VGPRs and seven
// wavefronts, a few values spilled); the FP64 form keeps its registers.  The
// 8^3-brick variants' staged starts (35.5 KB of LDS per workgroup) leave room
// for four workgroups only, so there the target is unmet (the compiler says
// so) and they take 88 VGPRs without spills; they are autotune candidates
// that lose at 512^3.
template <typename R>
constexpr int pp_min_blocks_per_cu() {
  return sizeof(R) == 4 ? 8 : 1;
}
constexpr int kPPClaimCounters = 8;

template <typename R, int U, int G, int PB>
__global__ void __launch_bounds__(256, pp_min_blocks_per_cu<R>()) k_pp_force(PPLevel L, const PPBox* __restrict__ boxes,
                                                  int64_t nbricks, const int64_t* __restrict__ start,
                                                  const float4* __restrict__ ent, float* __restrict__ vp,
                                                  PPOut* __restrict__ out, unsigned* __restrict__ claim,
                                                  bool xcd_local) {
  constexpr int NB = PB + 2, RW = kSX * NB + 1;  // staged sub-bin starts per (y, z) row
  __shared__ int pre[PB * PB * PB + 1];
  __shared__ int rs[NB * NB * RW];  // entry indices (a batch holds < 2^31 entries)
  __shared__ float red[4];
  __shared__ unsigned long long cnt_red[2];
  __shared__ float4 qtab[kQSeg];
  __shared__ int64_t s_brick;
  if (sizeof(R) == 4) fill_qtab(qtab);
  const double vsc = L.vdev ? *L.vdev : L.v_scale;
  const int nbr = (L.npgrid + PB - 1) / PB;
  const int ctr = blockIdx.x % kPPClaimCounters;
  if (threadIdx.x == 0) {
    cnt_red[0] = 0;
    cnt_red[1] = 0;
  }
  float f2 = 0, vm[3] = {0, 0, 0};
  unsigned long long npp = 0, npg = 0;
  for (;;) {
    __syncthreads();  // the previous brick's tables are no longer read
    if (threadIdx.x == 0) {
      const int64_t k = (int64_t)atomicAdd(&claim[ctr], 1u), nbps = (int64_t)nbr * nbr * nbr;
      if (xcd_local) {
        // counter ctr hands out the bricks of sub-tiles ctr, ctr + 8, ...
        // (in the host's densest-first order), one sub-tile's bricks in
        // turn: its workgroups (blockIdx % 8 == ctr) share an XCD, whose L2
        // then holds the neighbourhoods that a sub-tile's bricks read in
        // common
        const int64_t box = (k / nbps) * kPPClaimCounters + ctr;
        s_brick = box < nbricks / nbps ? box * nbps + k % nbps : nbricks;
      } else {
        s_brick = k * kPPClaimCounters + ctr;  // few sub-tiles: bricks 8k + ctr, every counter busy
      }
    }
    __syncthreads();
    int64_t b = s_brick;
    if (b >= nbricks) break;
    const int bx = (int)(b % nbr);
    b /= nbr;
    const int by = (int)(b % nbr);
    b /= nbr;
    const int bz = (int)(b % nbr);
    const int ib = (int)(b / nbr);
    const int m = L.npgrid + 2;
    const int64_t mS = (int64_t)m * kSX;
    const int64_t cbase = (int64_t)ib * m * m * m;
    const int c0[3] = {1 + bx * PB, 1 + by * PB, 1 + bz * PB};
    // rows (y, z) = c0-1 .. c0+PB, sub-bins kSX*(c0x-1) .. kSX*(c0x+PB+1); rows
    // outside the PP grid are empty, sub-bins past the row end read the row end
    for (int q = threadIdx.x; q < NB * NB * RW; q += blockDim.x) {
      const int xs = q % RW, y = c0[1] - 1 + (q / RW) % NB, z = c0[2] - 1 + q / (RW * NB);
      int st = 0;
      if (y <= L.npgrid + 1 && z <= L.npgrid + 1) {
        const int64_t row = (cbase + ((int64_t)z * m + y) * m) * kSX;
        st = (int)start[row + min((int64_t)kSX * (c0[0] - 1) + xs, mS)];
      }
      rs[q] = st;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int s = 0;
      for (int q = 0; q < PB * PB * PB; ++q) {
        pre[q] = s;
        const int lx = q % PB, ly = (q / PB) % PB, lz = q / (PB * PB);
        if (c0[0] + lx <= L.npgrid && c0[1] + ly <= L.npgrid && c0[2] + lz <= L.npgrid) {
          const int row = ((lz + 1) * NB + ly + 1) * RW;
          s += (int)(rs[row + kSX * (lx + 2)] - rs[row + kSX * (lx + 1)]);
        }
      }
      pre[PB * PB * PB] = s;
    }
    __syncthreads();
    const PPBox B = boxes[ib];
    const float xoff[3] = {(float)(B.ifs[0] * L.ntt), (float)(B.ifs[1] * L.ntt), (float)(B.ifs[2] * L.ntt)};
    const int ntg = pre[PB * PB * PB];
    const int gj = (int)threadIdx.x % G;  // lane within the target's group
    for (int t = (int)threadIdx.x / G; t < ntg; t += (int)blockDim.x / G) {
      int lo = 0, hi = PB * PB * PB - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= t) lo = mid;
        else hi = mid - 1;
      }
      const int lx = lo % PB, ly = (lo / PB) % PB, lz = lo / (PB * PB);
      const int rowt = ((lz + 1) * NB + ly + 1) * RW;
      const float4 me = ent[rs[rowt + kSX * (lx + 1)] + (t - pre[lo])];
      // PP coordinates of the target, the binning's xv = sf (xvec - ifs ntt)
      // (xvec = xf / rcs exactly for a power-of-two rcs)
      const float xme[3] = {me.x, me.y, me.z};
      float fr[3];
      for (int d = 0; d < 3; ++d) {
        const float xv = (float)L.sf * (xme[d] * L.inv_rcs - xoff[d]);
        fr[d] = xv - (float)(c0[d] - 1 + (d == 0 ? lx : (d == 1 ? ly : lz)));  // offset in its PP cell
      }
      // the target's x in local sub-bins (cells lx .. lx+2 of the staged rows)
      const float u = (fr[0] + (float)(lx + 1)) * (float)kSX;
      R f1[3] = {0, 0, 0};
      // pairs in range of this target, and those with a ghost partner (32-bit)
      int npair = 0, nghost = 0;
      const bool gx0 = c0[0] + lx - 1 == 0, gx1 = c0[0] + lx + 1 == L.npgrid + 1;
      for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy) {
          const float ey = dy < 0 ? fr[1] : (dy > 0 ? 1.0f - fr[1] : 0.0f);
          const float ez = dz < 0 ? fr[2] : (dz > 0 ? 1.0f - fr[2] : 0.0f);
          const float e2 = ey * ey + ez * ez;
          if (e2 >= L.rcut2) continue;
          // x window of the column: a partner at y/z distance >= (ey, ez)
          // is in range only within x +- sqrt(r_pp^2 - ey^2 - ez^2)
          const float hw = (float)kSX * sqrtf(L.rcut2 - e2);
          const int wlo = max((int)floorf(u - hw), kSX * lx);
          const int whi = min((int)floorf(u + hw), kSX * (lx + 3) - 1);
          const int ny = c0[1] + ly + dy, nz = c0[2] + lz + dz;
          const bool gyz = ny == 0 || nz == 0 || ny == L.npgrid + 1 || nz == L.npgrid + 1;
          const int row = rowt + (dz * NB + dy) * RW;
          const int beg = rs[row + wlo];
          const int nn = (int)(rs[row + whi + 1] - beg);
          // candidates [ph0, ph1) of the window are physical: the ghost cells
          // x-1 / x+1 lie before / after them, a ghost row holds none
          int ph0 = gx0 ? max((int)(rs[row + kSX * (lx + 1)] - beg), 0) : 0;
          int ph1 = gx1 ? min((int)(rs[row + kSX * (lx + 2)] - beg), nn) : nn;
          if (gyz) ph0 = ph1 = 0;
          const unsigned phn = (unsigned)(ph1 - ph0);
          const float4* run = ent + beg;
          for (int p = gj; p < nn; p += U * G) {
            float4 o[U];
  #pragma unroll
            for (int k = 0; k < U; ++k) o[k] = run[p + k * G];
  #pragma unroll
            for (int k = 0; k < U; ++k) {
              const float rv0 = o[k].x - me.x, rv1 = o[k].y - me.y, rv2 = o[k].z - me.z;
              const float r2 = (rv0 * rv0 + rv1 * rv1) + rv2 * rv2;  // exact-input FP32 test (kick.f90:525, 556)
              if (p + k * G < nn && 0.0f < r2 && r2 < L.pp_range2) {
                pair_force<R>(L, qtab, me, o[k], f1);
                ++npair;
                nghost += (unsigned)(p + k * G - ph0) >= phn;
              }
            }
          }
        }
      npg += (unsigned)nghost;
      npp += (unsigned)(npair - nghost);
      // the group's partial sums, combined by a butterfly: every lane ends
      // with the same value (each step adds the same two operands)
      for (int o = G / 2; o > 0; o >>= 1)
        for (int d = 0; d < 3; ++d) f1[d] += __shfl_xor(f1[d], o, 64);
      if (gj == 0) {
        // update v (kick.f90:576-582): g = fl32(f mass_scale) as the
        // reference forms it, times the FP64 coefficient, added in FP64
        const unsigned ip = __float_as_uint(me.w);
        float g[3];
        for (int d = 0; d < 3; ++d) g[d] = (float)f1[d] * L.mass_scale;
        f2 = fmaxf(f2, (g[0] * g[0] + g[1] * g[1]) + g[2] * g[2]);
        for (int d = 0; d < 3; ++d) {
          const float v = (float)((double)vp[3 * (int64_t)ip + d] + (double)g[d] * vsc);
          vp[3 * (int64_t)ip + d] = v;
          vm[d] = fmaxf(vm[d], fabsf(v));
        }
      }
    }
  }  // claimed bricks
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
  const int64_t ncell = (int64_t)nbox * m * m * m * kSX;
  zero_async(S.count, sizeof(int32_t) * ncell, s);
  const int nr = L.ntt + 2;
  const int64_t nthr = (int64_t)nbox * nr * nr * nr * 64;
  const size_t lds = sizeof(int) * 4 * L.sf * L.sf * L.sf * kSX;
  if (izipx == 2)
    hipLaunchKernelGGL((k_pp_bin<2, 0>), dim3((unsigned)ceil_div(nthr, 256)), dim3(256), lds, s, L, boxes, nbox,
                       (const int16_t*)xp, rhoc, off, S.count, nullptr, nullptr);
  else
    hipLaunchKernelGGL((k_pp_bin<1, 0>), dim3((unsigned)ceil_div(nthr, 256)), dim3(256), lds, s, L, boxes, nbox,
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
  const size_t lds = sizeof(int) * 4 * L.sf * L.sf * L.sf * kSX;
  if (izipx == 2)
    hipLaunchKernelGGL((k_pp_bin<2, 1>), dim3((unsigned)ceil_div(nthr, 256)), dim3(256), lds, s, L, boxes, nbox,
                       (const int16_t*)xp, rhoc, off, S.count, S.start, S.entries);
  else
    hipLaunchKernelGGL((k_pp_bin<1, 1>), dim3((unsigned)ceil_div(nthr, 256)), dim3(256), lds, s, L, boxes, nbox,
                       (const int8_t*)xp, rhoc, off, S.count, S.start, S.entries);
  CUBE_HIP(hipGetLastError());
}

template <int U, int G, int PB>
static void launch_pp_f32(dim3 grid, hipStream_t s, const PPLevel& L, const PPBox* boxes, int64_t nbricks,
                          PPScratch& S, float* vp, PPOut* out) {
  static_assert((U - 1) * G < kPPPad, "reads past a run stay inside the entry padding");
  // whole sub-tiles per counter only when there are enough of them to keep
  // the eight counters evenly busy
  const bool xcd_local = nbricks / ((int64_t)((L.npgrid + PB - 1) / PB) * ((L.npgrid + PB - 1) / PB) *
                                    ((L.npgrid + PB - 1) / PB)) >= 64;
  hipLaunchKernelGGL((k_pp_force<float, U, G, PB>), grid, dim3(256), 0, s, L, boxes, nbricks, S.start, S.entries,
                     vp, out, S.claim, xcd_local);
}

template <int G, int PB>
static void launch_pp_f32_u(int depth, dim3 grid, hipStream_t s, const PPLevel& L, const PPBox* boxes,
                            int64_t nbricks, PPScratch& S, float* vp, PPOut* out) {
  switch (depth) {
    case 4: launch_pp_f32<4, G, PB>(grid, s, L, boxes, nbricks, S, vp, out); break;
    case 8: launch_pp_f32<8, G, PB>(grid, s, L, boxes, nbricks, S, vp, out); break;
    default: launch_pp_f32<6, G, PB>(grid, s, L, boxes, nbricks, S, vp, out); break;
  }
}

void pp_force(const PPLevel& L, int precision, const PPBox* boxes, int nbox, float* vp, PPScratch& S, PPOut* out,
              hipStream_t s, int unroll) {
  const int pb = precision == 32 && unroll >= kPPBrick8 ? 8 : 4;
  const int depth = unroll % kPPBrick8;
  const int64_t nbr = (L.npgrid + pb - 1) / pb;
  const int64_t nbricks = nbr * nbr * nbr * nbox;
  static int n_cu = 0;
  if (n_cu == 0) {
    int dev = 0;
    CUBE_HIP(hipGetDevice(&dev));
    CUBE_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
  }
  zero_async(S.claim, sizeof(unsigned) * kPPClaimCounters, s);
  // as many workgroups as fit on the chip at once (eight per CU in the FP32
  // form, the launch bound's occupancy; four in the FP64 form, whose 117
  // VGPRs allow four wavefronts per SIMD), each then claims bricks until
  // none remain
  const int per_cu = precision == 64 ? 4 : pp_min_blocks_per_cu<float>();
  const dim3 grid((unsigned)std::min<int64_t>(nbricks, (int64_t)n_cu * per_cu));
  if (precision == 64)
    hipLaunchKernelGGL((k_pp_force<double, 6, 1, 4>), grid, dim3(256), 0, s, L, boxes, nbricks, S.start,
                       S.entries, vp, out, S.claim, nbox >= 64);
  else if (pb == 8)
    launch_pp_f32_u<kPPGroup, 8>(depth, grid, s, L, boxes, nbricks, S, vp, out);
  else
    launch_pp_f32_u<kPPGroup, 4>(depth, grid, s, L, boxes, nbricks, S, vp, out);
  CUBE_HIP(hipGetLastError());
}

}  // namespace cube
