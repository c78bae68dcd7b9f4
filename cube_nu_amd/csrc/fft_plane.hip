// fft_plane.hip -- size-specialised 3-pass FFT convolution for cubes with
// n <= 160: (A) per z-plane, r2c along x and forward c2c along y in LDS;
// (B) per y-slice, forward c2c along z, times G, inverse along z in LDS;
// (C) per z-plane, inverse along y and c2r along x, with the normalisation.
// The radix plan, line counts and per-thread butterfly counts are
// compile-time constants; stages run in place in LDS (each thread loads its
// butterflies into registers, barrier, writes them back).
#include "fft.h"
#include "timing.h"

namespace cube {
namespace {

struct Fac {
  int n;
  int r[16];
};
constexpr Fac factor(int n) {
  Fac f{};
  int m = n;
  // the largest in-register radices first: fewer LDS round trips per line
  while (m % 16 == 0) { f.r[f.n++] = 16; m /= 16; }
  while (m % 8 == 0) { f.r[f.n++] = 8; m /= 8; }
  while (m % 4 == 0) { f.r[f.n++] = 4; m /= 4; }
  while (m % 2 == 0) { f.r[f.n++] = 2; m /= 2; }
  while (m % 3 == 0) { f.r[f.n++] = 3; m /= 3; }
  while (m % 5 == 0) { f.r[f.n++] = 5; m /= 5; }
  return f;  // m == 1 for every specialised size
}

template <typename C>
struct CT;
template <>
struct CT<float2> {
  using R = float;
  __device__ static float2 mk(float a, float b) { return make_float2(a, b); }
};
template <>
struct CT<double2> {
  using R = double;
  __device__ static double2 mk(double a, double b) { return make_double2(a, b); }
};
template <typename C>
__device__ inline C add(C a, C b) { return CT<C>::mk(a.x + b.x, a.y + b.y); }
template <typename C>
__device__ inline C sub(C a, C b) { return CT<C>::mk(a.x - b.x, a.y - b.y); }
template <typename C>
__device__ inline C mul(C a, C b) { return CT<C>::mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
template <typename C>
__device__ inline C scl(C a, typename CT<C>::R s) { return CT<C>::mk(a.x * s, a.y * s); }
template <typename C>
__device__ inline C miu(C a, typename CT<C>::R u) { return CT<C>::mk(-u * a.y, u * a.x); }  // i*u*a
template <typename C>
__device__ inline C cj(C a) { return CT<C>::mk(a.x, -a.y); }

template <int SIGN, int R, typename C>
__device__ inline void bfly(C* v);

// 4-point DFT of (a, b, c, d) in place, e^{SIGN 2 pi i jk/4}
template <int SIGN, typename C>
__device__ inline void dft4(C& a, C& b, C& c, C& d) {
  using RR = typename CT<C>::R;
  const RR s = (RR)SIGN;
  const C p = add(a, c), q = sub(a, c), r = add(b, d), t = miu(sub(b, d), s);
  a = add(p, r);
  c = sub(p, r);
  b = add(q, t);
  d = sub(q, t);
}

template <int SIGN, int R, typename C>
__device__ inline void bfly(C* v) {
  using RR = typename CT<C>::R;
  const RR s = (RR)SIGN;
  if constexpr (R == 16) {
    // four 4-point DFTs over n = r + 4m, twiddles W16^(r k1), then 4-point
    // DFTs over r: X[k1 + 4 k2]
    C u[16];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      C a = v[r], b = v[r + 4], c = v[r + 8], d = v[r + 12];
      dft4<SIGN>(a, b, c, d);
      u[4 * r + 0] = a;
      u[4 * r + 1] = b;
      u[4 * r + 2] = c;
      u[4 * r + 3] = d;
    }
    constexpr double kC[10] = {1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173, 0.0,
                               -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613, -1.0,
                               -0.92387953251128675613};
    constexpr double kS[10] = {0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613, 1.0,
                               0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173, 0.0,
                               -0.38268343236508977173};
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      C t[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int j = r * k1;  // 0..9
        const C w = CT<C>::mk((RR)kC[j], s * (RR)kS[j]);
        t[r] = (j == 0) ? u[4 * r + k1] : mul(u[4 * r + k1], w);
      }
      dft4<SIGN>(t[0], t[1], t[2], t[3]);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
    }
  } else if constexpr (R == 8) {
    // even/odd 4-point DFTs, X[k] = E[k] + W8^k O[k], X[k+4] = E[k] - W8^k O[k]
    C e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6], o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<SIGN>(e0, e1, e2, e3);
    dft4<SIGN>(o0, o1, o2, o3);
    const RR h = (RR)0.70710678118654752440;
    const C w1 = CT<C>::mk(h, s * h), w3 = CT<C>::mk(-h, s * h);
    o1 = mul(o1, w1);
    o2 = miu(o2, s);
    o3 = mul(o3, w3);
    v[0] = add(e0, o0);
    v[4] = sub(e0, o0);
    v[1] = add(e1, o1);
    v[5] = sub(e1, o1);
    v[2] = add(e2, o2);
    v[6] = sub(e2, o2);
    v[3] = add(e3, o3);
    v[7] = sub(e3, o3);
  } else if constexpr (R == 2) {
    C a = v[0], b = v[1];
    v[0] = add(a, b);
    v[1] = sub(a, b);
  } else if constexpr (R == 4) {
    C a = add(v[0], v[2]), b = sub(v[0], v[2]);
    C c = add(v[1], v[3]), d = miu(sub(v[1], v[3]), s);
    v[0] = add(a, c);
    v[2] = sub(a, c);
    v[1] = add(b, d);
    v[3] = sub(b, d);
  } else if constexpr (R == 3) {
    const RR c = (RR)-0.5, u = s * (RR)0.86602540378443864676;
    C t1 = add(v[1], v[2]);
    C m = add(v[0], scl(t1, c));
    C r = miu(sub(v[1], v[2]), u);
    v[0] = add(v[0], t1);
    v[1] = add(m, r);
    v[2] = sub(m, r);
  } else {
    static_assert(R == 5, "radix");
    const RR c1 = (RR)0.30901699437494742410, c2 = (RR)-0.80901699437494742410;
    const RR s1 = (RR)0.95105651629515357212, s2 = (RR)0.58778525229247312917;
    C a1 = add(v[1], v[4]), a2 = add(v[2], v[3]);
    C b1 = sub(v[1], v[4]), b2 = sub(v[2], v[3]);
    C p1 = add(v[0], add(scl(a1, c1), scl(a2, c2)));
    C p2 = add(v[0], add(scl(a1, c2), scl(a2, c1)));
    C q1 = miu(add(scl(b1, s1), scl(b2, s2)), s);
    C q2 = miu(sub(scl(b1, s2), scl(b2, s1)), s);
    v[0] = add(v[0], add(a1, a2));
    v[1] = add(p1, q1);
    v[4] = sub(p1, q1);
    v[2] = add(p2, q2);
    v[3] = sub(p2, q2);
  }
}

// In-place Stockham over NL lines of length N held in LDS; element e of line
// l lives at buf[l*LS + e*ES].  LFAST selects the thread mapping (lines
// fastest for strided lines, elements fastest for contiguous lines).  tw is
// the W_{N*TWS} table; twiddles are W_N^e = tw[e*TWS].
template <int SIGN, int N, int NL, int LS, int ES, bool LFAST, int NT, int S, int NS, int TWS, typename C>
__device__ inline void stages(C* buf, const C* __restrict__ tw) {
  constexpr Fac F = factor(N);
  if constexpr (S < F.n) {
    constexpr int R = F.r[S];
    constexpr int NB = N / R;
    constexpr int TOT = NL * NB;
    constexpr int MB = (TOT + NT - 1) / NT;
    constexpr int STEP = N / (NS * R);
    C v[MB][R];
#pragma unroll
    for (int b = 0; b < MB; ++b) {
      const int idx = threadIdx.x + b * NT;
      if (idx < TOT) {
        const int l = LFAST ? idx % NL : idx / NB;
        const int j = LFAST ? idx / NL : idx % NB;
        const int jm = j % NS;
        const C* sp = buf + l * LS + j * ES;
        v[b][0] = sp[0];
#pragma unroll
        for (int r = 1; r < R; ++r) {
          C w = tw[(r * jm * STEP) * TWS];
          if (SIGN > 0) w.y = -w.y;
          v[b][r] = mul(sp[r * NB * ES], w);
        }
        bfly<SIGN, R>(v[b]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MB; ++b) {
      const int idx = threadIdx.x + b * NT;
      if (idx < TOT) {
        const int l = LFAST ? idx % NL : idx / NB;
        const int j = LFAST ? idx / NL : idx % NB;
        const int jm = j % NS;
        C* dp = buf + l * LS + ((j / NS) * NS * R + jm) * ES;
#pragma unroll
        for (int k = 0; k < R; ++k) dp[k * NS * ES] = v[b][k];
      }
    }
    __syncthreads();
    stages<SIGN, N, NL, LS, ES, LFAST, NT, S + 1, NS * R, TWS>(buf, tw);
  }
}

template <typename R>
using CP = typename Cplx<R>::T;

// Pass A: one z-plane: rho (N x N reals) -> complex (N rows of N/2+1).
template <typename R, int N, int NT>
__global__ void __launch_bounds__(NT) k_plane_fwd(const R* __restrict__ rho, CP<R>* __restrict__ out,
                                                  const CP<R>* __restrict__ twg) {
  using C = CP<R>;
  constexpr int H = N / 2, P = H + 1;
  __shared__ C buf[N * P];
  __shared__ C tw[N];  // W_N^e, staged once: the stages read it from LDS
  const int64_t plane = blockIdx.x;  // box * N + z
  const R* src = rho + plane * N * N;
  for (int i = threadIdx.x; i < N * H; i += NT) {
    const int y = i / H, j = i % H;
    buf[y * P + j] = CT<C>::mk(src[y * N + 2 * j], src[y * N + 2 * j + 1]);
  }
  for (int i = threadIdx.x; i < N; i += NT) tw[i] = twg[i];
  __syncthreads();
  stages<-1, H, N, P, 1, false, NT, 0, 1, 2>(buf, tw);  // half-length rows
  // split: X[k] = (Z[k]+conj Z[H-k])/2 - i/2 W_N^k (Z[k]-conj Z[H-k]), k=0..H
  const R half = (R)0.5;
  constexpr int NPAIR = H / 2 + 1;
  C xa[(N * NPAIR + NT - 1) / NT], xb[(N * NPAIR + NT - 1) / NT];
#pragma unroll
  for (int b = 0; b < (N * NPAIR + NT - 1) / NT; ++b) {
    const int i = threadIdx.x + b * NT;
    if (i < N * NPAIR) {
      const int y = i / NPAIR, k = i % NPAIR;
      const int k2 = H - k;
      C zk = buf[y * P + (k == H ? 0 : k)], zk2 = buf[y * P + (k2 == H ? 0 : k2)];
      // X[k]
      {
        C zc = cj(zk2);
        C xe = scl(add(zk, zc), half), d = scl(sub(zk, zc), half);
        C xo = CT<C>::mk(d.y, -d.x);
        xa[b] = add(xe, mul(tw[k], xo));
      }
      // X[H-k]
      {
        C zc = cj(zk);
        C xe = scl(add(zk2, zc), half), d = scl(sub(zk2, zc), half);
        C xo = CT<C>::mk(d.y, -d.x);
        xb[b] = add(xe, mul(tw[k2], xo));
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < (N * NPAIR + NT - 1) / NT; ++b) {
    const int i = threadIdx.x + b * NT;
    if (i < N * NPAIR) {
      const int y = i / NPAIR, k = i % NPAIR;
      buf[y * P + k] = xa[b];
      buf[y * P + (H - k)] = xb[b];
    }
  }
  __syncthreads();
  stages<-1, N, P, 1, P, true, NT, 0, 1, 1>(buf, tw);  // columns along y
  C* dst = out + plane * N * P;
  for (int i = threadIdx.x; i < N * P; i += NT) dst[i] = buf[i];
}

// Pass B: YB consecutive y-slices of one box: lines along z, times G, back.
// Several slices per block make the reads and writes of each z row longer
// contiguous runs (YB*(N/2+1) complex values).
template <typename R, int N, int NT, int YB>
__global__ void __launch_bounds__(NT) k_zslice(CP<R>* __restrict__ data, const float* __restrict__ G,
                                               const CP<R>* __restrict__ twg) {
  using C = CP<R>;
  constexpr int P = N / 2 + 1, W = YB * P;
  __shared__ C buf[N * W];
  __shared__ C tw[N];
  const int64_t box = blockIdx.x / (N / YB);
  const int y0 = (blockIdx.x % (N / YB)) * YB;
  C* base = data + box * (int64_t)N * N * P + (int64_t)y0 * P;
  for (int i = threadIdx.x; i < N * W; i += NT) {
    const int z = i / W, x = i % W;
    buf[i] = base[(int64_t)z * N * P + x];
  }
  for (int i = threadIdx.x; i < N; i += NT) tw[i] = twg[i];
  __syncthreads();
  stages<-1, N, W, 1, W, true, NT, 0, 1, 1>(buf, tw);
  const float* gb = G + (int64_t)y0 * P;
  for (int i = threadIdx.x; i < N * W; i += NT) {
    const int z = i / W, x = i % W;
    const R g = (R)gb[(int64_t)z * N * P + x];
    buf[i] = CT<C>::mk(buf[i].x * g, buf[i].y * g);  // cxyz*tf (kick.f90:698)
  }
  __syncthreads();
  stages<+1, N, W, 1, W, true, NT, 0, 1, 1>(buf, tw);
  for (int i = threadIdx.x; i < N * W; i += NT) {
    const int z = i / W, x = i % W;
    base[(int64_t)z * N * P + x] = buf[i];
  }
}

// Pass C: one z-plane: inverse along y, c2r along x, normalisation.
template <typename R, int N, int NT>
__global__ void __launch_bounds__(NT) k_plane_inv(const CP<R>* __restrict__ in, R* __restrict__ phi,
                                                  const CP<R>* __restrict__ twg, R div3, R div1) {
  using C = CP<R>;
  constexpr int H = N / 2, P = H + 1;
  __shared__ C buf[N * P];
  __shared__ C tw[N];
  const int64_t plane = blockIdx.x;
  const C* src = in + plane * N * P;
  for (int i = threadIdx.x; i < N * P; i += NT) buf[i] = src[i];
  for (int i = threadIdx.x; i < N; i += NT) tw[i] = twg[i];
  __syncthreads();
  stages<+1, N, P, 1, P, true, NT, 0, 1, 1>(buf, tw);
  // Z[k] = E + iO with E = X[k] + conj X[H-k], O = (X[k] - conj X[H-k]) W_N^-k
  constexpr int NPAIR = H / 2 + 1;
  constexpr int MB = (N * NPAIR + NT - 1) / NT;
  C za[MB], zb[MB];
#pragma unroll
  for (int b = 0; b < MB; ++b) {
    const int i = threadIdx.x + b * NT;
    if (i < N * NPAIR) {
      const int y = i / NPAIR, k = i % NPAIR, k2 = H - k;
      const C xk = buf[y * P + k], xk2 = buf[y * P + k2];
      {
        C xc = cj(xk2), w = tw[k];
        w.y = -w.y;
        za[b] = add(add(xk, xc), miu(mul(sub(xk, xc), w), (typename CT<C>::R)1));
      }
      {
        C xc = cj(xk), w = tw[k2];
        w.y = -w.y;
        zb[b] = add(add(xk2, xc), miu(mul(sub(xk2, xc), w), (typename CT<C>::R)1));
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < MB; ++b) {
    const int i = threadIdx.x + b * NT;
    if (i < N * NPAIR) {
      const int y = i / NPAIR, k = i % NPAIR, k2 = H - k;
      if (k < H) buf[y * P + k] = za[b];
      if (k2 < H) buf[y * P + k2] = zb[b];
    }
  }
  __syncthreads();
  stages<+1, H, N, P, 1, false, NT, 0, 1, 2>(buf, tw);
  R* dst = phi + plane * N * N;
  for (int i = threadIdx.x; i < N * H; i += NT) {
    const int y = i / H, j = i % H;
    const C z = buf[y * P + j];
    R a = z.x, b = z.y;
    if (div3 != (R)0) {
      a = a / div3;
      b = b / div3;
    } else {
      a = ((a / div1) / div1) / div1;
      b = ((b / div1) / div1) / div1;
    }
    dst[y * N + 2 * j] = a;
    dst[y * N + 2 * j + 1] = b;
  }
}

template <typename R, int N, int NT, int NTZ, int YB>
void run_plane_shape(const FFTCube<R>& c, R* rho, CP<R>* work, const float* G, int64_t nbox, R div3, R div1,
                     hipStream_t s, KTimer* tm, const std::string& t) {
  const double Gb = (double)N * N * N * sizeof(R) * nbox;
  const double Kb = (double)N * N * (N / 2 + 1) * sizeof(CP<R>) * nbox;
  const double Gk = (double)N * N * (N / 2 + 1) * sizeof(float);
  {
    KScope k(tm, (t + "fft_xy_fwd").c_str(), Gb + Kb);
    hipLaunchKernelGGL((k_plane_fwd<R, N, NT>), dim3((unsigned)(nbox * N)), dim3(NT), 0, s, rho, work, c.tw);
    CUBE_HIP(hipGetLastError());
  }
  {
    KScope k(tm, (t + "fft_z_green").c_str(), 2 * Kb + Gk);
    hipLaunchKernelGGL((k_zslice<R, N, NTZ, YB>), dim3((unsigned)(nbox * (N / YB))), dim3(NTZ), 0, s, work, G, c.tw);
    CUBE_HIP(hipGetLastError());
  }
  {
    KScope k(tm, (t + "fft_xy_inv").c_str(), Gb + Kb);
    hipLaunchKernelGGL((k_plane_inv<R, N, NT>), dim3((unsigned)(nbox * N)), dim3(NT), 0, s, work, rho, c.tw, div3,
                       div1);
    CUBE_HIP(hipGetLastError());
  }
}

// Launch shapes: shape 0 = threads per plane / slice measured best at n = 80 and
// 160 on MI355X; shape 1 = the other thread count for the plane passes; shape
// 2 = one z slice per workgroup instead of two (or half the threads where one
// slice already fills the workgroup's LDS).  Every shape computes each
// butterfly with the same operations, so all give the same result; the FP64
// form keeps shape 0.
template <typename R, int N>
bool run_plane(const FFTCube<R>& c, R* rho, CP<R>* work, const float* G, int64_t nbox, R div3, R div1,
               hipStream_t s, KTimer* tm, const std::string& t, int shape) {
  constexpr int NT = N >= 128 ? 1024 : (N >= 80 ? 512 : 256);
  constexpr int NTZ = N >= 80 ? 1024 : 256;
  // two slices per block where their LDS fits in 64 KB
  constexpr int YB = (N % 2 == 0 && 2 * N * (N / 2 + 1) * (int)sizeof(CP<R>) <= 65536) ? 2 : 1;
  if constexpr (std::is_same<R, float>::value && N >= 64) {
    constexpr int NT1 = N >= 128 ? 512 : (N >= 80 ? 256 : 512);
    constexpr int NTZ2 = YB == 2 ? NTZ : (NTZ >= 512 ? NTZ / 2 : 512);
    if (shape == 1) {
      run_plane_shape<R, N, NT1, NTZ, YB>(c, rho, work, G, nbox, div3, div1, s, tm, t);
      return true;
    }
    if (shape == 2) {
      run_plane_shape<R, N, NT, NTZ2, 1>(c, rho, work, G, nbox, div3, div1, s, tm, t);
      return true;
    }
  }
  run_plane_shape<R, N, NT, NTZ, YB>(c, rho, work, G, nbox, div3, div1, s, tm, t);
  return true;
}

}  // namespace

// Returns false when n has no specialised kernel (the caller then uses the
// generic line passes).
template <typename R>
bool fft_convolve_plane(const FFTCube<R>& c, R* rho, typename Cplx<R>::T* work, const float* G, int64_t nbox,
                        R div3, R div1, hipStream_t s, KTimer* tm, const char* tag, int shape) {
  std::string t(tag);
  switch (c.n) {
    case 16: return run_plane<R, 16>(c, rho, work, G, nbox, div3, div1, s, tm, t, shape);
    case 24: return run_plane<R, 24>(c, rho, work, G, nbox, div3, div1, s, tm, t, shape);
    case 32: return run_plane<R, 32>(c, rho, work, G, nbox, div3, div1, s, tm, t, shape);
    case 48: return run_plane<R, 48>(c, rho, work, G, nbox, div3, div1, s, tm, t, shape);
    case 64: return run_plane<R, 64>(c, rho, work, G, nbox, div3, div1, s, tm, t, shape);
    case 72: return run_plane<R, 72>(c, rho, work, G, nbox, div3, div1, s, tm, t, shape);
    case 80: return run_plane<R, 80>(c, rho, work, G, nbox, div3, div1, s, tm, t, shape);
    case 96: return run_plane<R, 96>(c, rho, work, G, nbox, div3, div1, s, tm, t, shape);
    default: break;
  }
  // n^2 (n/2+1) planes of FP64 beyond 96 do not fit in LDS
  if constexpr (std::is_same<R, float>::value) {
    if (c.n == 128) return run_plane<R, 128>(c, rho, work, G, nbox, div3, div1, s, tm, t, shape);
    if (c.n == 144) return run_plane<R, 144>(c, rho, work, G, nbox, div3, div1, s, tm, t, shape);  // PM3 at 1024^3
    if (c.n == 160) return run_plane<R, 160>(c, rho, work, G, nbox, div3, div1, s, tm, t, shape);
  }
  return false;
}

template bool fft_convolve_plane<float>(const FFTCube<float>&, float*, float2*, const float*, int64_t, float, float,
                                        hipStream_t, KTimer*, const char*, int);
template bool fft_convolve_plane<double>(const FFTCube<double>&, double*, double2*, const float*, int64_t, double,
                                         double, hipStream_t, KTimer*, const char*, int);

}  // namespace cube
