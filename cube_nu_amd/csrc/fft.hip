// fft.hip -- LDS-staged mixed-radix Stockham FFTs for the PM meshes.
//
// A 3-D r2c/c2r convolution is five passes over the batch of cubes:
//   x: r2c rows (half-length complex FFT + split), y: forward columns,
//   z: forward columns, times G, inverse columns (one kernel, data stays in
//   LDS), y: inverse columns, x: c2r rows (+ normalisation).
// Every pass stages whole lines in LDS; column passes load TX adjacent
// x-columns per workgroup so global accesses are contiguous runs.
#include "fft.h"
#include "timing.h"

namespace cube {

FFTStages make_stages(int n) {
  FFTStages s{};
  s.n = n;
  s.nst = 0;
  int m = n;
  // the largest in-register radices first: fewer LDS round trips per line
  while (m % 16 == 0) { s.rad[s.nst++] = 16; m /= 16; }
  while (m % 8 == 0) { s.rad[s.nst++] = 8; m /= 8; }
  while (m % 4 == 0) { s.rad[s.nst++] = 4; m /= 4; }
  while (m % 2 == 0) { s.rad[s.nst++] = 2; m /= 2; }
  while (m % 3 == 0) { s.rad[s.nst++] = 3; m /= 3; }
  while (m % 5 == 0) { s.rad[s.nst++] = 5; m /= 5; }
  for (int p = 7; m > 1; p += 2)
    while (m % p == 0) {
      if (p > 16) throw Error(1, "FFT length " + std::to_string(n) + " has a prime factor > 16");
      s.rad[s.nst++] = p;
      m /= p;
    }
  if (s.nst > 16) throw Error(1, "FFT length too composite");
  return s;
}

namespace {

template <typename C>
struct CO;
template <>
struct CO<float2> {
  using R = float;
  __device__ static float2 mk(float a, float b) { return make_float2(a, b); }
};
template <>
struct CO<double2> {
  using R = double;
  __device__ static double2 mk(double a, double b) { return make_double2(a, b); }
};

template <typename C>
__device__ inline C cadd(C a, C b) { return CO<C>::mk(a.x + b.x, a.y + b.y); }
template <typename C>
__device__ inline C csub(C a, C b) { return CO<C>::mk(a.x - b.x, a.y - b.y); }
template <typename C>
__device__ inline C cmul(C a, C b) { return CO<C>::mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
template <typename C>
__device__ inline C cscale(C a, typename CO<C>::R s) { return CO<C>::mk(a.x * s, a.y * s); }
// multiply by i*u
template <typename C>
__device__ inline C cmul_iu(C a, typename CO<C>::R u) { return CO<C>::mk(-u * a.y, u * a.x); }
template <typename C>
__device__ inline C cconj(C a) { return CO<C>::mk(a.x, -a.y); }

// Twiddle W_n^e (conjugated for the inverse direction).
template <int SIGN, typename C>
__device__ inline C twid(const C* tw, int e) {
  C w = tw[e];
  if (SIGN > 0) w.y = -w.y;
  return w;
}

// In-register R-point DFT, y_k = sum_r v_r exp(SIGN 2 pi i r k / R).
template <int SIGN, typename C>
__device__ inline void dft(int R, C* v, const C* tw, int twn, int twstride) {
  using RR = typename CO<C>::R;
  const RR s = (RR)SIGN;
  if (R == 2) {
    C a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if (R == 4) {
    C a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
    C c = cadd(v[1], v[3]), d = cmul_iu(csub(v[1], v[3]), s);
    v[0] = cadd(a, c);
    v[2] = csub(a, c);
    v[1] = cadd(b, d);
    v[3] = csub(b, d);
  } else if (R == 3) {
    const RR c = (RR)-0.5, u = s * (RR)0.86602540378443864676;
    C t1 = cadd(v[1], v[2]);
    C m = cadd(v[0], cscale(t1, c));
    C r = cmul_iu(csub(v[1], v[2]), u);
    v[0] = cadd(v[0], t1);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
  } else if (R == 5) {
    const RR c1 = (RR)0.30901699437494742410, c2 = (RR)-0.80901699437494742410;
    const RR s1 = (RR)0.95105651629515357212, s2 = (RR)0.58778525229247312917;
    C a1 = cadd(v[1], v[4]), a2 = cadd(v[2], v[3]);
    C b1 = csub(v[1], v[4]), b2 = csub(v[2], v[3]);
    C p1 = cadd(v[0], cadd(cscale(a1, c1), cscale(a2, c2)));
    C p2 = cadd(v[0], cadd(cscale(a1, c2), cscale(a2, c1)));
    C q1 = cmul_iu(cadd(cscale(b1, s1), cscale(b2, s2)), s);
    C q2 = cmul_iu(csub(cscale(b1, s2), cscale(b2, s1)), s);
    v[0] = cadd(v[0], cadd(a1, a2));
    v[1] = cadd(p1, q1);
    v[4] = csub(p1, q1);
    v[2] = cadd(p2, q2);
    v[3] = csub(p2, q2);
  } else {
    C y[16];
    for (int k = 0; k < R; ++k) {
      C acc = v[0];
      for (int r = 1; r < R; ++r) {
        int e = ((r * k) % R) * (twn / R);
        acc = cadd(acc, cmul(v[r], twid<SIGN>(tw, e * twstride)));
      }
      y[k] = acc;
    }
    for (int k = 0; k < R; ++k) v[k] = y[k];
  }
}

// 4-point DFT of (a, b, c, d) in place, e^{SIGN 2 pi i jk/4}.
template <int SIGN, typename C>
__device__ inline void dft4(C& a, C& b, C& c, C& d) {
  using RR = typename CO<C>::R;
  const RR s = (RR)SIGN;
  const C p = cadd(a, c), q = csub(a, c), r = cadd(b, d), t = cmul_iu(csub(b, d), s);
  a = cadd(p, r);
  c = csub(p, r);
  b = cadd(q, t);
  d = csub(q, t);
}

// Compile-time radix R in {2, 3, 4, 5, 8, 16}: the same butterflies as dft()
// for 2..5, register-resident (a runtime radix indexes its array and puts it
// in scratch memory).
template <int SIGN, int R, typename C>
__device__ inline void dft_fixed(C* v, const C* tw, int twn, int twstride) {
  using RR = typename CO<C>::R;
  const RR s = (RR)SIGN;
  if constexpr (R == 16) {
    // 4-point DFTs over n = r + 4m, twiddles W16^(r k1), 4-point DFTs over r
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
        const int j = r * k1;
        t[r] = (j == 0) ? u[4 * r + k1] : cmul(u[4 * r + k1], CO<C>::mk((RR)kC[j], s * (RR)kS[j]));
      }
      dft4<SIGN>(t[0], t[1], t[2], t[3]);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
    }
  } else if constexpr (R == 8) {
    C e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6], o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<SIGN>(e0, e1, e2, e3);
    dft4<SIGN>(o0, o1, o2, o3);
    const RR h = (RR)0.70710678118654752440;
    o1 = cmul(o1, CO<C>::mk(h, s * h));
    o2 = cmul_iu(o2, s);
    o3 = cmul(o3, CO<C>::mk(-h, s * h));
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, o1);
    v[5] = csub(e1, o1);
    v[2] = cadd(e2, o2);
    v[6] = csub(e2, o2);
    v[3] = cadd(e3, o3);
    v[7] = csub(e3, o3);
  } else if constexpr (R <= 5) {
    dft<SIGN>(R, v, tw, twn, twstride);
  } else {  // primes 7, 11, 13: the plain DFT with table twiddles
    C y[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      C acc = v[0];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const int e = ((r * k) % R) * (twn / R);
        acc = cadd(acc, cmul(v[r], twid<SIGN>(tw, e * twstride)));
      }
      y[k] = acc;
    }
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = y[k];
  }
}

template <int SIGN, int R, typename C>
__device__ inline void stockham_stage(const C* src, C* dst, int nl, int LS, int n, int Ns, const C* tw,
                                      int twstride) {
  const int nb = n / R;
  const int total = nl * nb;
  const int step = n / (Ns * R);
  for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
    const int t = idx / nb, j = idx - t * nb;
    const int jm = j % Ns;
    C v[R];
    const C* sp = src + t * LS + j;
    v[0] = sp[0];
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = cmul(sp[r * nb], twid<SIGN>(tw, (r * jm * step) * twstride));
    dft_fixed<SIGN, R>(v, tw, n, twstride);
    C* dp = dst + t * LS + (j / Ns) * Ns * R + jm;
#pragma unroll
    for (int k = 0; k < R; ++k) dp[k * Ns] = v[k];
  }
}

// Stockham passes over `nl` lines of length n held in LDS (line stride LS).
// Returns the buffer holding the natural-order result.
template <int SIGN, typename C>
__device__ C* stockham(C* A, C* B, int nl, int LS, const FFTStages& st, const C* tw, int twstride) {
  const int n = st.n;
  C* src = A;
  C* dst = B;
  int Ns = 1;
  for (int s = 0; s < st.nst; ++s) {
    const int R = st.rad[s];
    switch (R) {
      case 2: stockham_stage<SIGN, 2>(src, dst, nl, LS, n, Ns, tw, twstride); break;
      case 3: stockham_stage<SIGN, 3>(src, dst, nl, LS, n, Ns, tw, twstride); break;
      case 4: stockham_stage<SIGN, 4>(src, dst, nl, LS, n, Ns, tw, twstride); break;
      case 5: stockham_stage<SIGN, 5>(src, dst, nl, LS, n, Ns, tw, twstride); break;
      case 8: stockham_stage<SIGN, 8>(src, dst, nl, LS, n, Ns, tw, twstride); break;
      case 16: stockham_stage<SIGN, 16>(src, dst, nl, LS, n, Ns, tw, twstride); break;
      case 7: stockham_stage<SIGN, 7>(src, dst, nl, LS, n, Ns, tw, twstride); break;
      case 11: stockham_stage<SIGN, 11>(src, dst, nl, LS, n, Ns, tw, twstride); break;
      default: stockham_stage<SIGN, 13>(src, dst, nl, LS, n, Ns, tw, twstride); break;  // make_stages: 13
    }
    __syncthreads();
    C* tmp = src;
    src = dst;
    dst = tmp;
    Ns *= R;
  }
  return src;
}

// x pass, forward: rows of n reals -> n/2+1 complex (FFTW r2c layout).
template <typename R>
__global__ void k_rows_r2c(const R* __restrict__ in, typename Cplx<R>::T* __restrict__ out, int64_t nrows,
                           int RW, FFTStages half, const typename Cplx<R>::T* __restrict__ tw) {
  using C = typename Cplx<R>::T;
  extern __shared__ unsigned char smem_raw[];
  C* A = reinterpret_cast<C*>(smem_raw);
  const int h = half.n, n = 2 * h, LS = h + 1;
  C* B = A + RW * LS;
  const int64_t row0 = (int64_t)blockIdx.x * RW;
  const int nl = (int)min((int64_t)RW, nrows - row0);
  for (int idx = threadIdx.x; idx < nl * h; idx += blockDim.x) {
    int t = idx / h, j = idx - t * h;
    const R* rp = in + (row0 + t) * n;
    A[t * LS + j] = CO<C>::mk(rp[2 * j], rp[2 * j + 1]);
  }
  __syncthreads();
  C* Z = stockham<-1>(A, B, nl, LS, half, tw, 2);
  const R half_r = (R)0.5;
  for (int idx = threadIdx.x; idx < nl * (h + 1); idx += blockDim.x) {
    int t = idx / (h + 1), k = idx - t * (h + 1);
    C zk = Z[t * LS + (k == h ? 0 : k)];
    C zc = cconj(Z[t * LS + (k == 0 ? 0 : h - k)]);
    C xe = cscale(cadd(zk, zc), half_r);
    C d = cscale(csub(zk, zc), half_r);  // = i * Xo
    // X = Xe + W_n^k * Xo with Xo = -i d
    C xo = CO<C>::mk(d.y, -d.x);
    C w = tw[k % n];
    out[(row0 + t) * (h + 1) + k] = cadd(xe, cmul(w, xo));
  }
}

// x pass, inverse: n/2+1 complex -> n reals (FFTW c2r, unnormalised), then
// the reference normalisation.
template <typename R>
__global__ void k_rows_c2r(const typename Cplx<R>::T* __restrict__ in, R* __restrict__ out, int64_t nrows, int RW,
                           FFTStages half, const typename Cplx<R>::T* __restrict__ tw, R div3, R div1) {
  using C = typename Cplx<R>::T;
  extern __shared__ unsigned char smem_raw[];
  C* A = reinterpret_cast<C*>(smem_raw);
  const int h = half.n, n = 2 * h, LS = h + 1;
  C* B = A + RW * LS;
  const int64_t row0 = (int64_t)blockIdx.x * RW;
  const int nl = (int)min((int64_t)RW, nrows - row0);
  for (int idx = threadIdx.x; idx < nl * h; idx += blockDim.x) {
    int t = idx / h, k = idx - t * h;
    const C* rp = in + (row0 + t) * (h + 1);
    C xk = rp[k];
    C xc = cconj(rp[h - k]);
    C e = cadd(xk, xc);
    C w = tw[k];
    w.y = -w.y;  // W_n^{-k}
    C o = cmul(csub(xk, xc), w);
    A[t * LS + k] = cadd(e, cmul_iu(o, (R)1));  // E + i O
  }
  __syncthreads();
  C* Z = stockham<+1>(A, B, nl, LS, half, tw, 2);
  for (int idx = threadIdx.x; idx < nl * h; idx += blockDim.x) {
    int t = idx / h, j = idx - t * h;
    C z = Z[t * LS + j];
    R a = z.x, b = z.y;
    if (div3 != (R)0) {
      a = a / div3;
      b = b / div3;
    } else {
      a = ((a / div1) / div1) / div1;
      b = ((b / div1) / div1) / div1;
    }
    R* op = out + (row0 + t) * n;
    op[2 * j] = a;
    op[2 * j + 1] = b;
  }
}

// Column passes along y (stride nx) or z (stride n*nx) on [box][z][y][nx].
// MODE 0: forward; 1: inverse; 2: forward, times G, inverse (z axis).
template <typename R, int MODE>
__global__ void k_cols(typename Cplx<R>::T* __restrict__ data, int nx, int64_t S, int64_t O, int64_t BS,
                       int ncx, int nother, int TX, FFTStages st, const typename Cplx<R>::T* __restrict__ tw,
                       const float* __restrict__ G) {
  using C = typename Cplx<R>::T;
  extern __shared__ unsigned char smem_raw[];
  C* A = reinterpret_cast<C*>(smem_raw);
  const int n = st.n, LS = n + 1;
  C* B = A + TX * LS;
  int64_t bid = blockIdx.x;
  const int cx = (int)(bid % ncx);
  bid /= ncx;
  const int o = (int)(bid % nother);
  const int64_t box = bid / nother;
  const int x0 = cx * TX;
  const int nl = min(TX, nx - x0);
  C* base = data + box * BS + (int64_t)o * O + x0;
  for (int idx = threadIdx.x; idx < n * nl; idx += blockDim.x) {
    int e = idx / nl, t = idx - e * nl;
    A[t * LS + e] = base[(int64_t)e * S + t];
  }
  __syncthreads();
  C* res;
  if (MODE == 1) {
    res = stockham<+1>(A, B, nl, LS, st, tw, 1);
  } else {
    res = stockham<-1>(A, B, nl, LS, st, tw, 1);
  }
  if (MODE == 2) {
    // cxyz = cxyz * tf (complex(4) times real(4), kick.f90:698, 704)
    const float* gb = G + (int64_t)o * O + x0;
    for (int idx = threadIdx.x; idx < n * nl; idx += blockDim.x) {
      int e = idx / nl, t = idx - e * nl;
      R g = (R)gb[(int64_t)e * S + t];
      C v = res[t * LS + e];
      res[t * LS + e] = CO<C>::mk(v.x * g, v.y * g);
    }
    __syncthreads();
    C* other = (res == A) ? B : A;
    res = stockham<+1>(res, other, nl, LS, st, tw, 1);
  }
  for (int idx = threadIdx.x; idx < n * nl; idx += blockDim.x) {
    int e = idx / nl, t = idx - e * nl;
    base[(int64_t)e * S + t] = res[t * LS + e];
  }
}

KTimer& null_timer() {
  static thread_local KTimer t;
  return t;
}

// CUBE_FFT_GENERIC=1 forces the generic five-pass line kernels (tests use it
// to check the specialised path against them).
bool use_plane_path() {
  const char* e = std::getenv("CUBE_FFT_GENERIC");
  return !(e && e[0] == '1');
}

template <typename R>
int rows_per_block(int h) {
  int bytes_per_row = 2 * (h + 1) * (int)sizeof(typename Cplx<R>::T);
  int rw = 32768 / bytes_per_row;
  return std::max(1, std::min(rw, 16));
}
template <typename R>
int cols_per_block(int n) {
  int bytes_per_line = 2 * (n + 1) * (int)sizeof(typename Cplx<R>::T);
  int tx = 40960 / bytes_per_line;
  int p = 1;
  while (p * 2 <= tx && p < 16) p *= 2;
  return p;
}

template <typename R, int MODE>
void launch_cols(const FFTCube<R>& c, typename Cplx<R>::T* data, int axis, int64_t nbox, const float* G,
                 hipStream_t s) {
  using C = typename Cplx<R>::T;
  const int n = c.n, nx = n / 2 + 1;
  const int TX = cols_per_block<R>(n);
  const int ncx = (int)ceil_div(nx, TX);
  int64_t S = axis == 1 ? nx : (int64_t)n * nx;
  int64_t O = axis == 1 ? (int64_t)n * nx : nx;
  int64_t BS = (int64_t)n * n * nx;
  int64_t nblk = (int64_t)ncx * n * nbox;
  size_t lds = 2 * (size_t)TX * (n + 1) * sizeof(C);
  hipLaunchKernelGGL((k_cols<R, MODE>), dim3((unsigned)nblk), dim3(256), lds, s, data, nx, S, O, BS, ncx, n, TX,
                     c.full, c.tw, G);
  CUBE_HIP(hipGetLastError());
}

}  // namespace

template <typename R>
void FFTCube<R>::init(int n_) {
  using C = typename Cplx<R>::T;
  n = n_;
  if (n % 2) throw Error(1, "FFT length must be even");
  full = make_stages(n);
  half = make_stages(n / 2);
  std::vector<C> h(n);
  for (int t = 0; t < n; ++t) {
    double a = -2.0 * kPi * (double)t / (double)n;
    h[t].x = (R)std::cos(a);
    h[t].y = (R)std::sin(a);
  }
  CUBE_HIP(hipMalloc(&tw, sizeof(C) * n));
  CUBE_HIP(hipMemcpy(tw, h.data(), sizeof(C) * n, hipMemcpyHostToDevice));
}

template <typename R>
void FFTCube<R>::release() {
  if (tw) (void)hipFree(tw);
  tw = nullptr;
}

template <typename R>
void fft_r2c3d(const FFTCube<R>& c, const R* rho, typename Cplx<R>::T* out, int64_t nbox, hipStream_t s) {
  using C = typename Cplx<R>::T;
  const int n = c.n, h = n / 2;
  int64_t nrows = nbox * n * n;
  int RW = rows_per_block<R>(h);
  size_t lds = 2 * (size_t)RW * (h + 1) * sizeof(C);
  hipLaunchKernelGGL(k_rows_r2c<R>, dim3((unsigned)ceil_div(nrows, RW)), dim3(256), lds, s, rho, out, nrows, RW,
                     c.half, c.tw);
  CUBE_HIP(hipGetLastError());
  launch_cols<R, 0>(c, out, 1, nbox, nullptr, s);
  launch_cols<R, 0>(c, out, 2, nbox, nullptr, s);
}

template <typename R>
void fft_c2r3d(const FFTCube<R>& c, typename Cplx<R>::T* in, R* out, int64_t nbox, hipStream_t s) {
  using C = typename Cplx<R>::T;
  const int n = c.n, h = n / 2;
  launch_cols<R, 1>(c, in, 2, nbox, nullptr, s);
  launch_cols<R, 1>(c, in, 1, nbox, nullptr, s);
  int64_t nrows = nbox * n * n;
  int RW = rows_per_block<R>(h);
  size_t lds = 2 * (size_t)RW * (h + 1) * sizeof(C);
  hipLaunchKernelGGL(k_rows_c2r<R>, dim3((unsigned)ceil_div(nrows, RW)), dim3(256), lds, s, in, out, nrows, RW,
                     c.half, c.tw, (R)1, (R)0);
  CUBE_HIP(hipGetLastError());
}

bool fft_plane_tunable(int n) {
  if (!use_plane_path()) return false;
  switch (n) {
    case 64: case 72: case 80: case 96: case 128: case 144: case 160: return true;
    default: return false;
  }
}

template <typename R>
void fft_convolve(const FFTCube<R>& c, R* rho, typename Cplx<R>::T* work, const float* G, int64_t nbox, R div3,
                  R div1, hipStream_t s, KTimer* tm, const char* tag, int shape) {
  using C = typename Cplx<R>::T;
  if (use_plane_path() && shape >= 0 &&
      fft_convolve_plane<R>(c, rho, work, G, nbox, div3, div1, s, tm ? tm : &null_timer(), tag, shape))
    return;
  const int n = c.n, h = n / 2;
  const int64_t nrows = nbox * n * n;
  const double Gb = (double)n * n * n * sizeof(R) * nbox;          // real cube bytes
  const double Kb = (double)n * n * (h + 1) * sizeof(C) * nbox;    // half-complex bytes
  const double Gk = (double)n * n * (h + 1) * sizeof(float);       // kernel bytes (shared)
  const int RW = rows_per_block<R>(h);
  const size_t lds = 2 * (size_t)RW * (h + 1) * sizeof(C);
  std::string t(tag);
  {
    KScope k(tm ? tm : &null_timer(), (t + "fft_x_r2c").c_str(), Gb + Kb);
    hipLaunchKernelGGL(k_rows_r2c<R>, dim3((unsigned)ceil_div(nrows, RW)), dim3(256), lds, s, rho, work, nrows, RW,
                       c.half, c.tw);
    CUBE_HIP(hipGetLastError());
  }
  {
    KScope k(tm ? tm : &null_timer(), (t + "fft_y_fwd").c_str(), 2 * Kb);
    launch_cols<R, 0>(c, work, 1, nbox, nullptr, s);
  }
  {
    KScope k(tm ? tm : &null_timer(), (t + "fft_z_green").c_str(), 2 * Kb + Gk);
    launch_cols<R, 2>(c, work, 2, nbox, G, s);
  }
  {
    KScope k(tm ? tm : &null_timer(), (t + "fft_y_inv").c_str(), 2 * Kb);
    launch_cols<R, 1>(c, work, 1, nbox, nullptr, s);
  }
  {
    KScope k(tm ? tm : &null_timer(), (t + "fft_x_c2r").c_str(), Gb + Kb);
    hipLaunchKernelGGL(k_rows_c2r<R>, dim3((unsigned)ceil_div(nrows, RW)), dim3(256), lds, s, work, rho, nrows, RW,
                       c.half, c.tw, div3, div1);
    CUBE_HIP(hipGetLastError());
  }
}

template struct FFTCube<float>;
template struct FFTCube<double>;
template void fft_convolve<float>(const FFTCube<float>&, float*, float2*, const float*, int64_t, float, float,
                                  hipStream_t, KTimer*, const char*, int);
template void fft_convolve<double>(const FFTCube<double>&, double*, double2*, const float*, int64_t, double, double,
                                   hipStream_t, KTimer*, const char*, int);
template void fft_r2c3d<float>(const FFTCube<float>&, const float*, float2*, int64_t, hipStream_t);
template void fft_r2c3d<double>(const FFTCube<double>&, const double*, double2*, int64_t, hipStream_t);
template void fft_c2r3d<float>(const FFTCube<float>&, float2*, float*, int64_t, hipStream_t);
template void fft_c2r3d<double>(const FFTCube<double>&, double2*, double*, int64_t, hipStream_t);

}  // namespace cube
