// cube_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of CUBE's P3M force-and-update step (cosmologychen/cube_nu),
// written to be the checker for the HIP implementation.  It is imported only
// by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by
// the product path.
//
// Arithmetic follows the reference's FP32 rounding order (compiled with
// -ffp-contract=off): decode and TSC positions in FP64, weights and
// accumulations in FP32, PP in FP32.  One deliberate difference: the kick
// and drift coefficients (a_mid*dt/6/pi, dt_mid = (dt_old+dt)/2 and the
// displacement dt_mid*v) are formed in FP64, where the reference rounds them
// to real(4) (particle_mesh.f90:205, kick.f90:574, 797, drift.f90:21, 57,
// 117), because the implementation under test keeps them FP64.
// FFTs are computed in FP64 (the reference's FFTW/MKL FP32 rounding is not
// reproducible; see the DFT-definition tests).  Loop orders follow the
// reference exactly, including the plane-colour passes that fix summation
// and drift orders.
//
// Index conventions: arrays are the reference's column-major arrays passed as
// flat C pointers; particle indices are 0-based here (Fortran ip-1).
//
// Every function is checked against an independent statement in
// tests/test_oracle.py (closed forms, numpy restatements, brute-force sums,
// invariants) before it is trusted as the checker of the GPU path.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef std::complex<double> cd;

extern "C" {

// Parameters mirrored from parameters.f90:8-132 (nn = 1).
struct OParams {
  int32_t ng, nnt, nns, ratio_cs, ncb, ngb, izipx;
  float app, apm1c, apm2, apm3f;
  int32_t pm3, pp, pp_corr;
  double f_nu, mass_nu;
  float tile_buffer, image_buffer;
  int32_t ires_fixed;  // 0: thresholds of kick.f90:116-130; 2..5: fixed (kick.f90:131, commented there)
};

struct OKickIn {
  double a_mid, dt;  // FP64 step coefficients (real(4) in the reference, variables.f90:35)
  float mass_p_cdm;
  int32_t phase_mask;
};

struct OKickOut {
  float dt_pm1, dt_pm2, dt_pm3, dt_pp, dt_vmax, vz_max;
  float f2max_pm1, f2max_pm2, f2max_pm3, f2max_pp;
  float vmax[3];
  int32_t napm[7];
  double testrho;
};

struct ODriftIn {
  double dt_old, dt;
  float vz_max;
};
struct ODriftOut {
  int64_t nplocal;
  float overhead_tile;
  int32_t nlayer;
};
}

namespace {

const double PI = 3.14159265358979323846;  // parameters.f90:94 pi=4*atan(1d0)
const int RATIO_SF[8] = {0, 1, 2, 4, 6, 8, 12, 16};  // parameters.f90:26 (1-based)

// Derived sizes, parameters.f90:57-82 and variables.f90:24-26.
struct Geo {
  int64_t ng, nnt, nns, rcs, ncb, ngb, nc, nt, ntt, ngp, ngt, nte, izipx, B;
  int64_t sf[8], nfp[8], nfb[8], nft[8];
  double xres;
  explicit Geo(const OParams& p) {
    ng = p.ng; nnt = p.nnt; nns = p.nns; rcs = p.ratio_cs; ncb = p.ncb; ngb = p.ngb;
    izipx = p.izipx; B = 8 * izipx;
    nc = ng / rcs; nt = nc / nnt; ntt = nt / nns; ngp = ng / nnt;
    ngt = ngp + 2 * ngb; nte = nt + 2 * ncb;
    for (int i = 1; i <= 7; ++i) {
      sf[i] = RATIO_SF[i];
      nfp[i] = (ngp / nns) * sf[i];
      nfb[i] = rcs * sf[i];
      nft[i] = nfp[i] + 2 * nfb[i];
    }
    xres = std::ldexp(1.0, -(int)B);  // parameters.f90:132
  }
  int64_t ntiles() const { return nnt * nnt * nnt; }
  int64_t ncell_tile() const { return nte * nte * nte; }
  // rhoc(i,j,k,itx,ity,itz) with Fortran indices (i in 1-ncb..nt+ncb, it in 1..nnt).
  int64_t rc(int64_t i, int64_t j, int64_t k, int64_t T) const {
    return ((T * nte + (k - 1 + ncb)) * nte + (j - 1 + ncb)) * nte + (i - 1 + ncb);
  }
  int64_t tile_lin(int64_t itx, int64_t ity, int64_t itz) const {
    return ((itz - 1) * nnt + (ity - 1)) * nnt + (itx - 1);
  }
};

// Offset (number of particles before) of every extended cell: the global
// exclusive scan of rhoc in memory order, which equals
// nzero = idx_b_r(j,k,tile) - sum(rhoc(i:,j,k,tile)) (particle_mesh.f90:32,
// spine_image variables.f90:245-270).
std::vector<int64_t> cell_offsets(const Geo& g, const int32_t* rhoc) {
  int64_t n = g.ntiles() * g.ncell_tile();
  std::vector<int64_t> off(n + 1);
  int64_t s = 0;
  for (int64_t c = 0; c < n; ++c) { off[c] = s; s += rhoc[c]; }
  off[n] = s;
  return off;
}

// Compressed coordinate decode, parameters.f90:91-92 and particle_mesh.f90:36:
// (int(xp+ishift,izipx)+rshift)*x_resolution = (u + 0.5) * 2^-B with u the
// two's-complement unsigned value of xp.
inline double decode(const void* xp, int64_t ip, int d, const Geo& g) {
  uint32_t u = g.izipx == 2 ? (uint32_t)(uint16_t)((const int16_t*)xp)[3 * ip + d]
                            : (uint32_t)(uint8_t)((const int8_t*)xp)[3 * ip + d];
  return ((double)u + 0.5) * g.xres;
}

// TSC weights, particle_mesh.f90:56-62: l3 = real(4)(xpos - floor(xpos)).
inline int64_t tsc(double xpos, float w[3]) {
  double fl = std::floor(xpos);
  float l3 = (float)(xpos - fl);
  float a = 1.0f - l3;
  w[0] = (a * a) / 2.0f;
  w[2] = (l3 * l3) / 2.0f;
  w[1] = (1.0f - w[0]) - w[2];
  return (int64_t)fl;  // Fortran idx(:,1) = floor(xpos)
}

// weight = [(alpha-1)/4, -alpha/2, alpha/2, (1-alpha)/4], alpha = 4./3 in FP32
// (parameters.f90:53-54).
struct FDW {
  float w[4];
  FDW() {
    float alpha = 4.0f / 3.0f;
    w[0] = (alpha - 1.0f) / 4.0f;
    w[1] = -alpha / 2.0f;
    w[2] = alpha / 2.0f;
    w[3] = (1.0f - alpha) / 4.0f;
  }
};
const FDW FD;

// G_grid = 1.0/6.0/pi stored as real(4) (variables.f90:28).
inline float g_grid() { return (float)((double)(1.0f / 6.0f) / PI); }

// Kick update, particle_mesh.f90:205 / kick.f90:797:
// vp = vp + dv*a_mid*dt/6/pi.  The coefficient a_mid*dt/6/pi is kept in FP64
// (the reference forms dv*a_mid*dt in real(4)); product and add in FP64.
inline double kick_coef(double a_mid, double dt) { return ((a_mid * dt) / 6.0) / PI; }
inline float kick_update(float vp, float dv, double kc) { return (float)((double)vp + (double)dv * kc); }

// ---------------------------------------------------------------------------
// FFT (FP64 Stockham, any radix; used for the r2c/c2r 3-D transforms).
struct Plan1D {
  int n;
  std::vector<int> rad;
  std::vector<cd> tw;  // e^{-2 pi i t / n}
  explicit Plan1D(int n_) : n(n_) {
    int m = n;
    for (int r : {4, 2, 3, 5}) while (m % r == 0) { rad.push_back(r); m /= r; }
    for (int r = 7; m > 1; r += 2) while (m % r == 0) { rad.push_back(r); m /= r; }
    tw.resize(n);
    for (int t = 0; t < n; ++t) tw[t] = std::polar(1.0, -2.0 * PI * t / n);
  }
};

// Out-of-place Stockham (Govindaraju et al. formulation), sign = -1 forward.
void fft1d(const Plan1D& p, cd* x, cd* work, int sign) {
  int n = p.n;
  cd* a = x;
  cd* b = work;
  int Ns = 1;
  cd v[16];
  for (int R : p.rad) {
    int nb = n / R;
    for (int j = 0; j < nb; ++j) {
      int jm = j % Ns;
      for (int r = 0; r < R; ++r) {
        cd t = a[j + r * nb];
        int e = (int)(((int64_t)r * jm * (n / (Ns * R))) % n);
        cd w = p.tw[e];
        if (sign > 0) w = std::conj(w);
        v[r] = t * w;
      }
      cd out[16];
      for (int k = 0; k < R; ++k) {
        cd s = 0;
        for (int r = 0; r < R; ++r) {
          int e = (int)(((int64_t)r * k * (n / R)) % n);
          cd w = p.tw[e];
          if (sign > 0) w = std::conj(w);
          s += v[r] * w;
        }
        out[k] = s;
      }
      int idx = (j / Ns) * Ns * R + jm;
      for (int k = 0; k < R; ++k) b[idx + k * Ns] = out[k];
    }
    std::swap(a, b);
    Ns *= R;
  }
  if (a != x) std::memcpy(x, a, sizeof(cd) * n);
}

// 3-D r2c: real [z][y][x] (n^3) -> complex [z][y][kx], kx = 0..n/2, e^{-i}
// unnormalised (FFTW convention used at kick.f90:697).
void r2c3d(int n, const float* in, cd* out) {
  int nh = n / 2 + 1;
  Plan1D p(n);
  std::vector<cd> line(n), work(n);
  for (int64_t zy = 0; zy < (int64_t)n * n; ++zy) {
    for (int x = 0; x < n; ++x) line[x] = cd(in[zy * n + x], 0.0);
    fft1d(p, line.data(), work.data(), -1);
    for (int kx = 0; kx < nh; ++kx) out[zy * nh + kx] = line[kx];
  }
  for (int z = 0; z < n; ++z)
    for (int kx = 0; kx < nh; ++kx) {
      for (int y = 0; y < n; ++y) line[y] = out[((int64_t)z * n + y) * nh + kx];
      fft1d(p, line.data(), work.data(), -1);
      for (int y = 0; y < n; ++y) out[((int64_t)z * n + y) * nh + kx] = line[y];
    }
  for (int y = 0; y < n; ++y)
    for (int kx = 0; kx < nh; ++kx) {
      for (int z = 0; z < n; ++z) line[z] = out[((int64_t)z * n + y) * nh + kx];
      fft1d(p, line.data(), work.data(), -1);
      for (int z = 0; z < n; ++z) out[((int64_t)z * n + y) * nh + kx] = line[z];
    }
}

// 3-D c2r, unnormalised (FFTW backward): the inverse of r2c3d times n^3.
void c2r3d(int n, cd* in, double* out) {
  int nh = n / 2 + 1;
  Plan1D p(n);
  std::vector<cd> line(n), work(n);
  for (int y = 0; y < n; ++y)
    for (int kx = 0; kx < nh; ++kx) {
      for (int z = 0; z < n; ++z) line[z] = in[((int64_t)z * n + y) * nh + kx];
      fft1d(p, line.data(), work.data(), +1);
      for (int z = 0; z < n; ++z) in[((int64_t)z * n + y) * nh + kx] = line[z];
    }
  for (int z = 0; z < n; ++z)
    for (int kx = 0; kx < nh; ++kx) {
      for (int y = 0; y < n; ++y) line[y] = in[((int64_t)z * n + y) * nh + kx];
      fft1d(p, line.data(), work.data(), +1);
      for (int y = 0; y < n; ++y) in[((int64_t)z * n + y) * nh + kx] = line[y];
    }
  for (int64_t zy = 0; zy < (int64_t)n * n; ++zy) {
    for (int kx = 0; kx < nh; ++kx) line[kx] = in[zy * nh + kx];
    for (int kx = nh; kx < n; ++kx) line[kx] = std::conj(in[zy * nh + (n - kx)]);
    // Hermitian completion: imaginary parts of self-conjugate bins are ignored
    // as a c2r transform does.
    line[0] = cd(line[0].real(), 0.0);
    if (n % 2 == 0) line[n / 2] = cd(line[n / 2].real(), 0.0);
    fft1d(p, line.data(), work.data(), +1);
    for (int x = 0; x < n; ++x) out[zy * n + x] = line[x].real();
  }
}

// Periodic Poisson solve of one mesh: r2c, multiply by the real tf, c2r.
// rho (FP32, n^3) -> phi (FP32, unnormalised, as FFTW leaves it).
void poisson(int n, const float* rho, const float* tf, float* phi) {
  int nh = n / 2 + 1;
  std::vector<cd> ck((size_t)nh * n * n);
  r2c3d(n, rho, ck.data());
  // cxyz = cxyz * tf : complex(4) times real(4) in FP32 (kick.f90:698)
  for (size_t i = 0; i < ck.size(); ++i) {
    float re = (float)ck[i].real(), im = (float)ck[i].imag();
    ck[i] = cd((double)(re * tf[i]), (double)(im * tf[i]));
  }
  std::vector<double> out((size_t)n * n * n);
  c2r3d(n, ck.data(), out.data());
  for (size_t i = 0; i < out.size(); ++i) phi[i] = (float)out[i];
}

// ---------------------------------------------------------------------------
// Green_3D (Green/Green.f90:1-80) at nn=1 (kshift terms vanish).
void green3d(int64_t n, float apm_soft, float apm_range, float gridsize, float* Gk) {
  int64_t nh = n / 2 + 1;
  std::vector<double> fftspace(n);
  float inv = 1.0f / (float)n;  // (1./nglobal) in default real
  for (int64_t k1 = 1; k1 <= n; ++k1)
    fftspace[k1 - 1] = (2.0 * PI * (double)inv) * (double)(((k1 + n / 2 - 1) % n) - n / 2);
  const double alpha = (double)(4.0f / 3.0f);
  const double one_m_alpha = (double)(1.0f - 4.0f / 3.0f);
  const double c215 = (double)(2.0f / 15.0f);
  const int p = 2, n_int = 2;
#pragma omp parallel for schedule(dynamic)
  for (int64_t k3 = 0; k3 < n; ++k3)
    for (int64_t k2 = 0; k2 < n; ++k2)
      for (int64_t k1 = 0; k1 < nh; ++k1) {
        double kv[3] = {fftspace[k1], fftspace[k2], fftspace[k3]};
        double D[3], ur[3] = {0, 0, 0};
        for (int d = 0; d < 3; ++d) D[d] = alpha * std::sin(kv[d]) + one_m_alpha * std::sin(2 * kv[d]) / 2;
        bool zero = (kv[0] == 0 && kv[1] == 0 && kv[2] == 0);
        for (int i3 = -n_int; i3 <= n_int; ++i3)
          for (int i2 = -n_int; i2 <= n_int; ++i2)
            for (int i1 = -n_int; i1 <= n_int; ++i1) {
              double kn[3] = {kv[0] + 2 * PI * i1, kv[1] + 2 * PI * i2, kv[2] + 2 * PI * i3};
              double kmag = std::sqrt(kn[0] * kn[0] + kn[1] * kn[1] + kn[2] * kn[2]);
              double U = 1;
              for (int d = 0; d < 3; ++d) U *= (kn[d] == 0) ? 1.0 : std::sin(kn[d] / 2) / (kn[d] / 2);
              U = std::pow(U, 2 * p + 2);
              double y = (double)apm_soft * kmag / 2;
              double Ss = (kmag == 0) ? 1.0 : 12 * std::pow(y, -4) * (2 - 2 * std::cos(y) - y * std::sin(y));
              double Sr;
              if (apm_range == 0) Sr = 0;
              else if (kmag == 0) Sr = 1;
              else {
                y = (double)apm_range * kmag / 2;
                Sr = 12 * std::pow(y, -4) * (2 - 2 * std::cos(y) - y * std::sin(y));
              }
              if (kmag == 0) continue;  // Rvec_n = 0/0 only at k=0, zeroed below
              for (int d = 0; d < 3; ++d) ur[d] += U * (kn[d] * (Ss * Ss - Sr * Sr) / (kmag * kmag));
            }
        double u2n = 1;
        for (int d = 0; d < 3; ++d) {
          double s = std::sin(kv[d] / 2);
          u2n *= 1 - s * s + c215 * (s * s) * (s * s);
        }
        double sDu = D[0] * ur[0] + D[1] * ur[1] + D[2] * ur[2];
        double sDD = D[0] * D[0] + D[1] * D[1] + D[2] * D[2];
        float gs2 = gridsize * gridsize;
        double G = 4 * PI * sDu / sDD / (u2n * u2n) / (double)gs2;
        if (zero) G = 0;
        Gk[(k3 * n + k2) * nh + k1] = (float)G;
      }
  // Gk(::nglobal/2,::nglobal/2,::nglobal/2)=0 (Green.f90:71)
  for (int64_t a : {(int64_t)0, n / 2})
    for (int64_t b : {(int64_t)0, n / 2})
      for (int64_t c : {(int64_t)0, n / 2}) Gk[(c * n + b) * nh + a] = 0;
}

// ---------------------------------------------------------------------------
// FD gradient of one mesh (particle_mesh.f90:224-247, kick.f90:839-862):
// force(d, i0) = sum(phi(i0 + [-2,-1,1,2]) * weight) for i0 in 0..nforce+1.
// phi index p maps to phiv[p - p0] along each axis (p0 = lower bound).
struct ForceField {
  int64_t nf;  // nforce
  std::vector<float> f;  // (3, 0:nf+1)^3
  float& at(int d, int64_t i, int64_t j, int64_t k) {
    int64_t m = nf + 2;
    return f[(((k * m + j) * m + i) * 3) + d];
  }
};

template <class PHI>
void fd_gradient(int64_t nforce, PHI phi, ForceField& F, float* f2max) {
  F.nf = nforce;
  int64_t m = nforce + 2;
  F.f.assign((size_t)3 * m * m * m, 0.0f);
  const int off[4] = {-2, -1, 1, 2};
#pragma omp parallel for
  for (int64_t k = 0; k < m; ++k)
    for (int64_t j = 0; j < m; ++j)
      for (int64_t i = 0; i < m; ++i) {
        float sx = 0, sy = 0, sz = 0;
        for (int q = 0; q < 4; ++q) {
          sx = sx + phi(i + off[q], j, k) * FD.w[q];
          sy = sy + phi(i, j + off[q], k) * FD.w[q];
          sz = sz + phi(i, j, k + off[q]) * FD.w[q];
        }
        F.at(0, i, j, k) = sx;
        F.at(1, i, j, k) = sy;
        F.at(2, i, j, k) = sz;
      }
  float mx = 0;
  for (int64_t k = 1; k <= nforce; ++k)
    for (int64_t j = 1; j <= nforce; ++j)
      for (int64_t i = 1; i <= nforce; ++i) {
        float a = F.at(0, i, j, k), b = F.at(1, i, j, k), c = F.at(2, i, j, k);
        float s = (a * a + b * b) + c * c;
        mx = std::max(mx, s);
      }
  *f2max = mx;
}

// TSC gather + kick for one particle (particle_mesh.f90:178-208).
inline void gather_kick(ForceField& F, double xpos[3], float* v, double kc,
                        float vmax[3]) {
  float w[3][3];
  int64_t i0[3];
  for (int d = 0; d < 3; ++d) i0[d] = tsc(xpos[d], w[d]);
  float dv[3] = {0, 0, 0};
  for (int i3 = 0; i3 < 3; ++i3)
    for (int i2 = 0; i2 < 3; ++i2)
      for (int i1 = 0; i1 < 3; ++i1)
        for (int d = 0; d < 3; ++d)
          dv[d] = dv[d] + ((F.at(d, i0[0] + i1, i0[1] + i2, i0[2] + i3) * w[0][i1]) * w[1][i2]) * w[2][i3];
  for (int d = 0; d < 3; ++d) {
    v[d] = kick_update(v[d], dv[d], kc);
    vmax[d] = std::max(vmax[d], std::fabs(v[d]));
  }
}

inline void atomic_max_f(float* a, float v) {
#pragma omp critical(amax)
  { if (v > *a) *a = v; }
}

// ---------------------------------------------------------------------------
// PM1 (particle_mesh.f90:1-222 with the pm settings of kick.f90:34-37).
void pm1(const Geo& g, const void* xp, float* vp, const int32_t* rhoc, const std::vector<int64_t>& off,
         const float* tf1, const OKickIn& kin, float* rho1_out, OKickOut& out) {
  const int64_t nc = g.nc, nt = g.nt, nnt = g.nnt;
  std::vector<float> rho1((size_t)nc * nc * nc, 0.0f);
  const double gridsize = (double)(float)g.rcs;
  double testrho = 0;
  // Deposit tile by tile onto rho_ex(1-nex:nt+nex), nex=2, cells 0..nt+1.
#pragma omp parallel for schedule(dynamic) reduction(+ : testrho)
  for (int64_t T = 0; T < g.ntiles(); ++T) {
    int64_t itx = T % nnt + 1, ity = (T / nnt) % nnt + 1, itz = T / (nnt * nnt) + 1;
    const int64_t nex = 2, L = nt + 2 * nex;
    std::vector<float> rex((size_t)L * L * L, 0.0f);
    auto R = [&](int64_t i, int64_t j, int64_t k) -> float& {
      return rex[((k + nex - 1) * L + (j + nex - 1)) * L + (i + nex - 1)];
    };
    for (int ilayer = 0; ilayer < 3; ++ilayer)
      for (int64_t k = 0 + ilayer; k <= nt + 1; k += 3)
        for (int64_t j = 0; j <= nt + 1; ++j)
          for (int64_t i = 0; i <= nt + 1; ++i) {
            int64_t c = g.rc(i, j, k, T);
            for (int64_t ip = off[c]; ip < off[c] + rhoc[c]; ++ip) {
              double xpos[3];
              int64_t ijk[3] = {i, j, k};
              float w[3][3];
              int64_t f[3];
              for (int d = 0; d < 3; ++d) {
                xpos[d] = (double)(ijk[d] - 1) + decode(xp, ip, d, g);
                xpos[d] = xpos[d] * (double)g.rcs / gridsize;
                f[d] = tsc(xpos[d], w[d]);  // idx(:,1)=floor(xpos)
              }
              for (int i3 = 0; i3 < 3; ++i3)
                for (int i2 = 0; i2 < 3; ++i2)
                  for (int i1 = 0; i1 < 3; ++i1) {
                    float& r = R(f[0] + i1, f[1] + i2, f[2] + i3);
                    r = r + (w[0][i1] * w[1][i2]) * w[2][i3];
                  }
            }
          }
    double tr = 0;
    for (int64_t k = 1; k <= nt; ++k)
      for (int64_t j = 1; j <= nt; ++j)
        for (int64_t i = 1; i <= nt; ++i) {
          float r = R(i, j, k);
          int64_t gi = (itx - 1) * nt + i - 1, gj = (ity - 1) * nt + j - 1, gk = (itz - 1) * nt + k - 1;
          rho1[(gk * nc + gj) * nc + gi] = r * kin.mass_p_cdm;
          tr += (double)r;
        }
    testrho += tr;
  }
  if (rho1_out) std::memcpy(rho1_out, rho1.data(), sizeof(float) * rho1.size());
  out.testrho = testrho;
  // Poisson solve; rho1 = rho1/nw/nw/nw in FP32 (pencil_fft.f90:77).
  std::vector<float> phi(rho1.size());
  poisson((int)nc, rho1.data(), tf1, phi.data());
  float fn = (float)nc;
  for (auto& v : phi) v = ((v / fn) / fn) / fn;
  // Periodic halo (variables.f90:120-129) + gradient.
  auto PH = [&](int64_t i, int64_t j, int64_t k) -> float {
    i = ((i - 1) % nc + nc) % nc; j = ((j - 1) % nc + nc) % nc; k = ((k - 1) % nc + nc) % nc;
    return phi[(k * nc + j) * nc + i];
  };
  ForceField F;
  float f2max;
  fd_gradient(nc, PH, F, &f2max);
  out.f2max_pm1 = f2max;
  // Interpolation + kick of physical particles (particle_mesh.f90:148-217).
  float vmax[3] = {0, 0, 0};
#pragma omp parallel for schedule(dynamic)
  for (int64_t T = 0; T < g.ntiles(); ++T) {
    int64_t it[3] = {T % nnt + 1, (T / nnt) % nnt + 1, T / (nnt * nnt) + 1};
    float vm[3] = {0, 0, 0};
    for (int64_t k = 1; k <= nt; ++k)
      for (int64_t j = 1; j <= nt; ++j)
        for (int64_t i = 1; i <= nt; ++i) {
          int64_t c = g.rc(i, j, k, T);
          int64_t ijk[3] = {i, j, k};
          for (int64_t ip = off[c]; ip < off[c] + rhoc[c]; ++ip) {
            double xpos[3];
            for (int d = 0; d < 3; ++d) {
              xpos[d] = (double)((it[d] - 1) * nt + (ijk[d] - 1)) + decode(xp, ip, d, g);
              xpos[d] = xpos[d] * (double)g.rcs / gridsize;
            }
            gather_kick(F, xpos, vp + 3 * ip, kick_coef(kin.a_mid, kin.dt), vm);
          }
        }
    for (int d = 0; d < 3; ++d) atomic_max_f(&vmax[d], vm[d]);
  }
  for (int d = 0; d < 3; ++d) out.vmax[d] = vmax[d];
  out.dt_pm1 = 0.5f * std::sqrt(1.0f / ((std::sqrt(f2max) * (float)kin.a_mid) * g_grid()));
}

// One tile or sub-tile of PM2/PM3 (pm_thread, kick.f90:593-837).
struct LevelBox {
  int64_t nwork, nb, nphy, nforce, nloop, nex;
  double gridsize;
  int64_t ushift[3];    // utile_shift * nfp(iapm)
  int64_t nc1[3], nc2[3];
  int64_t T;            // tile (0-based linear)
};

void pm_box(const Geo& g, const void* xp, float* vp, const int32_t* rhoc, const std::vector<int64_t>& off,
            const float* tf, const LevelBox& b, const OKickIn& kin, float* f2max_t, float vmax_t[3]) {
  // A box whose deposit window holds no particle has a zero density, hence a
  // zero potential and force (f2max_t = 0), and no physical particle to kick:
  // its result is known without the transforms.
  {
    int64_t npart = 0;
    for (int64_t k = b.nc1[2] - b.nloop; k <= b.nc2[2] + b.nloop && npart == 0; ++k)
      for (int64_t j = b.nc1[1] - b.nloop; j <= b.nc2[1] + b.nloop; ++j)
        for (int64_t i = b.nc1[0] - b.nloop; i <= b.nc2[0] + b.nloop; ++i) npart += rhoc[g.rc(i, j, k, b.T)];
    if (npart == 0) {
      *f2max_t = 0;
      for (int d = 0; d < 3; ++d) vmax_t[d] = 0;
      return;
    }
  }
  const int64_t L = b.nphy + 2 * b.nex;
  std::vector<float> rth((size_t)L * L * L, 0.0f);
  auto R = [&](int64_t i, int64_t j, int64_t k) -> float& {
    return rth[((k + b.nex - 1) * L + (j + b.nex - 1)) * L + (i + b.nex - 1)];
  };
  for (int ilayer = 0; ilayer < 3; ++ilayer)
    for (int64_t k = b.nc1[2] - b.nloop + ilayer; k <= b.nc2[2] + b.nloop; k += 3)
      for (int64_t j = b.nc1[1] - b.nloop; j <= b.nc2[1] + b.nloop; ++j)
        for (int64_t i = b.nc1[0] - b.nloop; i <= b.nc2[0] + b.nloop; ++i) {
          int64_t c = g.rc(i, j, k, b.T);
          int64_t ijk[3] = {i, j, k};
          for (int64_t ip = off[c]; ip < off[c] + rhoc[c]; ++ip) {
            float w[3][3];
            int64_t f[3];
            for (int d = 0; d < 3; ++d) {
              double xpos = (double)(ijk[d] - 1) + decode(xp, ip, d, g);
              xpos = xpos * (double)g.rcs / b.gridsize - (double)b.ushift[d];
              f[d] = tsc(xpos, w[d]);
            }
            for (int i3 = 0; i3 < 3; ++i3)
              for (int i2 = 0; i2 < 3; ++i2)
                for (int i1 = 0; i1 < 3; ++i1) {
                  float& r = R(f[0] + i1, f[1] + i2, f[2] + i3);
                  r = r + (w[0][i1] * w[1][i2]) * w[2][i3];
                }
          }
        }
  // density(1-nb:nphy+nb) = rho_th(m1:m2)*mass_p_cdm (kick.f90:686)
  const int64_t n = b.nwork;
  std::vector<float> dens((size_t)n * n * n), phi((size_t)n * n * n);
  for (int64_t k = 0; k < n; ++k)
    for (int64_t j = 0; j < n; ++j)
      for (int64_t i = 0; i < n; ++i)
        dens[(k * n + j) * n + i] = R(i + 1 - b.nb, j + 1 - b.nb, k + 1 - b.nb) * kin.mass_p_cdm;
  poisson((int)n, dens.data(), tf, phi.data());
  float fn = (float)n;
  float n3 = (fn * fn) * fn;  // real(pm%nwork)**3
  for (auto& v : phi) v = v / n3;
  auto PH = [&](int64_t i, int64_t j, int64_t k) -> float {
    return phi[((k - (1 - b.nb)) * n + (j - (1 - b.nb))) * n + (i - (1 - b.nb))];
  };
  ForceField F;
  fd_gradient(b.nforce, PH, F, f2max_t);
  for (int d = 0; d < 3; ++d) vmax_t[d] = 0;
  for (int64_t k = b.nc1[2]; k <= b.nc2[2]; ++k)
    for (int64_t j = b.nc1[1]; j <= b.nc2[1]; ++j)
      for (int64_t i = b.nc1[0]; i <= b.nc2[0]; ++i) {
        int64_t c = g.rc(i, j, k, b.T);
        int64_t ijk[3] = {i, j, k};
        for (int64_t ip = off[c]; ip < off[c] + rhoc[c]; ++ip) {
          double xpos[3];
          for (int d = 0; d < 3; ++d) {
            xpos[d] = (double)(ijk[d] - 1) + decode(xp, ip, d, g);
            xpos[d] = xpos[d] * (double)g.rcs / b.gridsize - (double)b.ushift[d];
          }
          gather_kick(F, xpos, vp + 3 * ip, kick_coef(kin.a_mid, kin.dt), vmax_t);
        }
      }
}

// PM2 settings for one tile, kick.f90:54-58.
LevelBox pm2_box(const Geo& g, int64_t T) {
  LevelBox b;
  b.nwork = g.ngt; b.nb = g.ngb; b.nphy = g.ngp; b.nforce = g.ngp; b.nloop = g.ncb; b.nex = g.ngb + 1;
  b.gridsize = 1.0;
  for (int d = 0; d < 3; ++d) { b.ushift[d] = 0; b.nc1[d] = 1; b.nc2[d] = g.nt; }
  b.T = T;
  return b;
}

// PM3 settings for one sub-tile, kick.f90:212-219.
LevelBox pm3_box(const Geo& g, int64_t T, const int64_t ifs[3], int iapm) {
  LevelBox b;
  b.nwork = g.nft[iapm]; b.nb = g.nfb[iapm]; b.nphy = g.nfp[iapm]; b.nforce = g.nfp[iapm];
  b.nloop = 1; b.nex = g.nfb[iapm] + 1;
  b.gridsize = (double)(1.0f / (float)g.sf[iapm]);
  for (int d = 0; d < 3; ++d) {
    b.ushift[d] = (ifs[d] - 1) * g.nfp[iapm];
    b.nc1[d] = (ifs[d] - 1) * g.ntt + 1;
    b.nc2[d] = ifs[d] * g.ntt;
  }
  b.T = T;
  return b;
}

// Sub-tile classification, kick.f90:108-133.
int classify(const Geo& g, const int32_t* rhoc, int64_t T, const int64_t ifs[3]) {
  int32_t mx = 0;
  for (int64_t k = (ifs[2] - 1) * g.ntt + 1; k <= ifs[2] * g.ntt; ++k)
    for (int64_t j = (ifs[1] - 1) * g.ntt + 1; j <= ifs[1] * g.ntt; ++j)
      for (int64_t i = (ifs[0] - 1) * g.ntt + 1; i <= ifs[0] * g.ntt; ++i)
        mx = std::max(mx, rhoc[g.rc(i, j, k, T)]);
  float npc = (float)mx / 64.0f;
  if (npc < 24) return 2;
  if (npc < 114) return 3;
  if (npc < 320) return 4;
  return 5;
}

// F_ra_inline, kick.f90:360-374 (powers expanded by repeated products).
inline float f_ra(float r, float apm) {
  float ep = (2.0f * r) / apm;
  if (apm == 0.0f || ep > 2.0f) return 1.0f / (r * r);
  float c = (1.0f / 35.0f) / (apm * apm);
  float e2 = ep * ep, e3 = e2 * ep, e4 = e2 * e2, e5 = e4 * ep, e6 = e3 * e3;
  if (ep > 1.0f)
    return c * ((((((12.0f * (1.0f / e2) - 224.0f) + 896.0f * ep) - 840.0f * e2) + 224.0f * e3) + 70.0f * e4 - 48.0f * e5) + 7.0f * e6);
  if (ep > 0.0f) return c * ((((224.0f * ep - 224.0f * e3) + 70.0f * e4) + 48.0f * e5) - 21.0f * e6);
  return 0.0f;
}

// PP_Force for one sub-tile (kick.f90:376-588).
void pp_box(const Geo& g, const OParams& P, const void* xp, float* vp, const int32_t* rhoc,
            const std::vector<int64_t>& off, int64_t T, const int64_t ifs[3], int iapm, const OKickIn& kin,
            float* f2max_t, float vmax_t[3], int64_t* npairs) {
  float pp_range = 3.5f / (float)g.sf[iapm];  // apm3 = 3.5/ratio_sf (parameters.f90:43)
  float pp_range2 = pp_range * pp_range;
  int64_t rcp = g.sf[iapm], npgrid = g.ntt * rcp;
  *f2max_t = 0;
  for (int d = 0; d < 3; ++d) vmax_t[d] = 0;
  int64_t nc1[3], nc2[3];
  for (int d = 0; d < 3; ++d) { nc1[d] = (ifs[d] - 1) * g.ntt + 1; nc2[d] = ifs[d] * g.ntt; }
  int64_t m = npgrid + 2;
  auto CI = [&](int64_t i, int64_t j, int64_t k) { return (k * m + j) * m + i; };
  std::vector<int> count((size_t)m * m * m, 0), start((size_t)m * m * m, 0);
  struct Tmp { float xf[3]; int64_t ip; int64_t cell; };
  std::vector<Tmp> tmp;
  // Pass 1 (kick.f90:411-459)
  for (int64_t k = nc1[2] - 1; k <= nc2[2] + 1; ++k)
    for (int64_t j = nc1[1] - 1; j <= nc2[1] + 1; ++j)
      for (int64_t i = nc1[0] - 1; i <= nc2[0] + 1; ++i) {
        int64_t c = g.rc(i, j, k, T);
        int64_t ijk[3] = {i, j, k};
        for (int64_t ip = off[c]; ip < off[c] + rhoc[c]; ++ip) {
          Tmp t;
          t.ip = ip;
          int64_t idx[3];
          bool in = true;
          for (int d = 0; d < 3; ++d) {
            float xvec = (float)((double)(ijk[d] - 1) + decode(xp, ip, d, g));
            t.xf[d] = xvec * (float)g.rcs;
            float xv = (float)rcp * (xvec - (float)((ifs[d] - 1) * g.ntt));
            idx[d] = (int64_t)std::floor(xv) + 1;
            if (idx[d] < 0 || idx[d] > npgrid + 1) in = false;
          }
          if (in) {
            t.cell = CI(idx[0], idx[1], idx[2]);
            count[t.cell]++;
          } else {
            t.cell = -1;
          }
          tmp.push_back(t);
        }
      }
  if (tmp.empty()) return;
  // Pass 2: physical cells first (kick.f90:468-499)
  int64_t pos = 0, nphy = 0;
  for (int64_t k = 1; k <= npgrid; ++k)
    for (int64_t j = 1; j <= npgrid; ++j)
      for (int64_t i = 1; i <= npgrid; ++i) { start[CI(i, j, k)] = (int)pos; pos += count[CI(i, j, k)]; }
  nphy = pos;
  for (int64_t k = 0; k <= npgrid + 1; ++k)
    for (int64_t j = 0; j <= npgrid + 1; ++j)
      for (int64_t i = 0; i <= npgrid + 1; ++i) {
        if (i >= 1 && i <= npgrid && j >= 1 && j <= npgrid && k >= 1 && k <= npgrid) continue;
        start[CI(i, j, k)] = (int)pos;
        pos += count[CI(i, j, k)];
      }
  // Pass 3 (kick.f90:502-511)
  std::vector<float> xf((size_t)3 * std::max<int64_t>(pos, 1)), af((size_t)3 * std::max<int64_t>(pos, 1), 0.0f);
  std::vector<int64_t> iploc(std::max<int64_t>(nphy, 1));
  std::vector<int> cur(start);
  for (auto& t : tmp) {
    if (t.cell < 0) continue;
    int p = cur[t.cell]++;
    for (int d = 0; d < 3; ++d) xf[3 * p + d] = t.xf[d];
    if (p < nphy) iploc[p] = t.ip;
  }
  int64_t pairs = 0;
  auto pair = [&](int p1, int p2, float f1[3], bool guard0) {
    float rv[3] = {xf[3 * p2] - xf[3 * p1], xf[3 * p2 + 1] - xf[3 * p1 + 1], xf[3 * p2 + 2] - xf[3 * p1 + 2]};
    float r2 = (rv[0] * rv[0] + rv[1] * rv[1]) + rv[2] * rv[2];
    bool ok = guard0 ? (0.0f < r2 && r2 < pp_range2) : (r2 < pp_range2);
    if (!ok) return;
    ++pairs;
    float rmag = std::sqrt(r2);
    float fr = f_ra(rmag, P.app) - f_ra(rmag, pp_range);
    for (int d = 0; d < 3; ++d) {
      float fp = (rv[d] / rmag) * fr;
      f1[d] = f1[d] + fp;
      af[3 * p2 + d] = af[3 * p2 + d] - fp;
    }
  };
  // self pp over physical cells (kick.f90:516-534)
  for (int64_t k = 1; k <= npgrid; ++k)
    for (int64_t j = 1; j <= npgrid; ++j)
      for (int64_t i = 1; i <= npgrid; ++i) {
        int64_t c = CI(i, j, k);
        if (count[c] == 0) continue;
        int s1 = start[c], e1 = s1 + count[c] - 1;
        for (int p1 = s1; p1 <= e1 - 1; ++p1) {
          float f1[3] = {af[3 * p1], af[3 * p1 + 1], af[3 * p1 + 2]};
          for (int p2 = p1 + 1; p2 <= e1; ++p2) pair(p1, p2, f1, true);
          for (int d = 0; d < 3; ++d) af[3 * p1 + d] = f1[d];
        }
      }
  // neighbour pp: half shell ijk (initialize.f90:369-390), kick.f90:537-566
  int nb[13][3];
  {
    int l = 0;
    for (int k = -1; k <= -1; ++k) for (int j = -1; j <= 1; ++j) for (int i = -1; i <= 1; ++i) { nb[l][0] = i; nb[l][1] = j; nb[l][2] = k; ++l; }
    for (int j = -1; j <= -1; ++j) for (int i = -1; i <= 1; ++i) { nb[l][0] = i; nb[l][1] = j; nb[l][2] = 0; ++l; }
    nb[l][0] = -1; nb[l][1] = 0; nb[l][2] = 0;
  }
  for (int64_t k = 0; k <= npgrid + 1; ++k)
    for (int64_t j = 0; j <= npgrid + 1; ++j)
      for (int64_t i = 0; i <= npgrid + 1; ++i) {
        int64_t c = CI(i, j, k);
        if (count[c] == 0) continue;
        int s1 = start[c], e1 = s1 + count[c] - 1;
        bool ghost = std::min({i, j, k}) == 0 || std::max({i, j, k}) == npgrid + 1;
        for (int q = 0; q < 13; ++q) {
          int64_t a = i + nb[q][0], bq = j + nb[q][1], cc = k + nb[q][2];
          if (std::min({a, bq, cc}) < 0 || std::max({a, bq, cc}) > npgrid + 1) continue;
          if (ghost && (std::min({a, bq, cc}) == 0 || std::max({a, bq, cc}) == npgrid + 1)) continue;
          int64_t c2 = CI(a, bq, cc);
          if (count[c2] == 0) continue;
          int s2 = start[c2], e2 = s2 + count[c2] - 1;
          for (int p1 = s1; p1 <= e1; ++p1) {
            float f1[3] = {af[3 * p1], af[3 * p1 + 1], af[3 * p1 + 2]};
            for (int p2 = s2; p2 <= e2; ++p2) pair(p1, p2, f1, false);
            for (int d = 0; d < 3; ++d) af[3 * p1 + d] = f1[d];
          }
        }
      }
  // update v (kick.f90:569-582)
  float mass_scale;
  if (P.pp_corr && P.mass_nu > 0)
    mass_scale = (float)((double)kin.mass_p_cdm * ((1.0 - P.f_nu) * (1.0 - P.f_nu)));
  else
    mass_scale = kin.mass_p_cdm;
  const double v_scale = kick_coef(kin.a_mid, kin.dt);  // FP64 (kick.f90:574 forms it in real(4))
  for (int64_t p = 0; p < nphy; ++p) {
    float f1[3];
    for (int d = 0; d < 3; ++d) f1[d] = af[3 * p + d] * mass_scale;
    *f2max_t = std::max(*f2max_t, (f1[0] * f1[0] + f1[1] * f1[1]) + f1[2] * f1[2]);
    int64_t ip = iploc[p];
    for (int d = 0; d < 3; ++d) {
      vp[3 * ip + d] = (float)((double)vp[3 * ip + d] + (double)f1[d] * v_scale);
      vmax_t[d] = std::max(vmax_t[d], std::fabs(vp[3 * ip + d]));
    }
  }
  *npairs = pairs;
}

inline void subtile_ids(const Geo& g, int64_t it, int64_t* T, int64_t ifs[3]) {
  // ixyz3 order (initialize.f90:352-367): ifx fastest, then ify, ifz, itx, ity, itz.
  int64_t nns = g.nns;
  ifs[0] = it % nns + 1; ifs[1] = (it / nns) % nns + 1; ifs[2] = (it / (nns * nns)) % nns + 1;
  *T = it / (nns * nns * nns);
}

}  // namespace

extern "C" {

int orc_green(int64_t n, float apm_soft, float apm_range, float gridsize, float* Gk) {
  green3d(n, apm_soft, apm_range, gridsize, Gk);
  return 0;
}

int orc_r2c3d(int n, const float* in, double* out_complex) {
  r2c3d(n, in, (cd*)out_complex);
  return 0;
}
int orc_c2r3d(int n, const double* in_complex, double* out) {
  int nh = n / 2 + 1;
  std::vector<cd> tmp((const cd*)in_complex, (const cd*)in_complex + (size_t)nh * n * n);
  c2r3d(n, tmp.data(), out);
  return 0;
}
int orc_fft1d(int n, double* data, int sign) {
  Plan1D p(n);
  std::vector<cd> w(n);
  fft1d(p, (cd*)data, w.data(), sign);
  return 0;
}

// Decode of n*3 compressed coordinates (parameters.f90:91-92,
// particle_mesh.f90:36) into cell fractions.
int orc_decode(int izipx, int64_t n, const void* xp, double* out) {
  OParams p{};
  p.ng = 4; p.nnt = 1; p.nns = 1; p.ratio_cs = 1; p.izipx = izipx;
  Geo g(p);
  for (int64_t ip = 0; ip < n; ++ip)
    for (int d = 0; d < 3; ++d) out[3 * ip + d] = decode(xp, ip, d, g);
  return 0;
}

// TSC weights and floor index of n positions (particle_mesh.f90:56-62).
int orc_tsc(int64_t n, const double* xpos, int64_t* idx, float* w) {
  for (int64_t i = 0; i < n; ++i) idx[i] = tsc(xpos[i], w + 3 * i);
  return 0;
}

// FD gradient of a periodic n^3 potential phi[k][j][i] (cells 1..n), the
// 4-point stencil of particle_mesh.f90:237-242 over the interior 1..n:
// force[k][j][i][d], plus f2max (particle_mesh.f90:144).
int orc_fd_gradient(int64_t n, const float* phi, float* force, float* f2max) {
  auto PH = [&](int64_t i, int64_t j, int64_t k) -> float {
    i = ((i - 1) % n + n) % n; j = ((j - 1) % n + n) % n; k = ((k - 1) % n + n) % n;
    return phi[(k * n + j) * n + i];
  };
  ForceField F;
  fd_gradient(n, PH, F, f2max);
  for (int64_t k = 1; k <= n; ++k)
    for (int64_t j = 1; j <= n; ++j)
      for (int64_t i = 1; i <= n; ++i)
        for (int d = 0; d < 3; ++d) force[(((k - 1) * n + (j - 1)) * n + (i - 1)) * 3 + d] = F.at(d, i, j, k);
  return 0;
}

// Periodic Poisson solve of one mesh as the PM levels do it (r2c, times the
// real tf in FP32, c2r, unnormalised).
int orc_poisson(int n, const float* rho, const float* tf, float* phi) {
  poisson(n, rho, tf, phi);
  return 0;
}

// Timing sample for the CPU baseline: with a stride s > 1 the kick's PM2, PM3
// and PP loops and the drift process only tiles T with T % s == 0 (PM1, a
// global mesh, runs whole); the outputs are then incomplete and serve only to
// time the work.  Seconds per phase of the last kick/drift, in the reference's
// tcat slots: [0] PM1 (11), [1] PM2 (12), [2] PM3 (13), [3] PP (14), [4] drift (2).
static int64_t g_sample_stride = 1;
static double g_phase_s[5] = {0, 0, 0, 0, 0};

int orc_set_sample_stride(int64_t s) {
  g_sample_stride = s < 1 ? 1 : s;
  return 0;
}

int orc_phase_seconds(double* t) {
  for (int i = 0; i < 5; ++i) t[i] = g_phase_s[i];
  return 0;
}

// Classification and napm (kick.f90:108-133).
int orc_classify(const OParams* P, const int32_t* rhoc, int32_t* ires, int32_t napm[7]) {
  Geo g(*P);
  int64_t ns3 = g.ntiles() * g.nns * g.nns * g.nns;
  for (int i = 0; i < 7; ++i) napm[i] = 0;
  for (int64_t it = 0; it < ns3; ++it) {
    int64_t T, ifs[3];
    subtile_ids(g, it, &T, ifs);
    // a fixed level replaces the thresholds (the commented ires(itile)=4, kick.f90:131)
    ires[it] = P->ires_fixed >= 2 ? P->ires_fixed : classify(g, rhoc, T, ifs);
    napm[ires[it] - 1]++;
  }
  return 0;
}

// Full kick (kick.f90:1-353) with per-phase control.  tf[0] = tf1, tf[1] = tf2,
// tf[2..5] = tf3_2..tf3_8.  ires_override (may be NULL) forces sub-tile levels.
int orc_kick(const OParams* P, const void* xp, float* vp, const int32_t* rhoc, const float* const* tf,
             const OKickIn* kin, OKickOut* out, float* rho1_out, int64_t* npairs_out) {
  Geo g(*P);
  std::memset(out, 0, sizeof(*out));
  std::vector<int64_t> off = cell_offsets(g, rhoc);
  int mask = kin->phase_mask ? kin->phase_mask : 15;
  float vmax[3] = {0, 0, 0};
  const int64_t stride = g_sample_stride;
  for (int i = 0; i < 4; ++i) g_phase_s[i] = 0;
  double t0 = omp_get_wtime();
  if (mask & 1) {
    OKickOut o1;
    std::memset(&o1, 0, sizeof(o1));
    pm1(g, xp, vp, rhoc, off, tf[0], *kin, rho1_out, o1);
    out->dt_pm1 = o1.dt_pm1; out->f2max_pm1 = o1.f2max_pm1; out->testrho = o1.testrho;
    for (int d = 0; d < 3; ++d) vmax[d] = o1.vmax[d];
  }
  g_phase_s[0] = omp_get_wtime() - t0;
  t0 = omp_get_wtime();
  float gg = g_grid();
  if (mask & 2) {
    float f2 = 0, vm[3] = {0, 0, 0};
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t T = 0; T < g.ntiles(); ++T) {
      if (T % stride) continue;
      float f2t, vt[3];
      LevelBox b = pm2_box(g, T);
      pm_box(g, xp, vp, rhoc, off, tf[1], b, *kin, &f2t, vt);
      atomic_max_f(&f2, f2t);
      for (int d = 0; d < 3; ++d) atomic_max_f(&vm[d], vt[d]);
    }
    out->f2max_pm2 = f2;
    out->dt_pm2 = 0.5f * std::sqrt(1.0f / ((std::sqrt(f2) * (float)kin->a_mid) * gg));
    for (int d = 0; d < 3; ++d) vmax[d] = vm[d];
  }
  g_phase_s[1] = omp_get_wtime() - t0;
  t0 = omp_get_wtime();
  int64_t ns3 = g.ntiles() * g.nns * g.nns * g.nns;
  std::vector<int32_t> ires(ns3);
  if (mask & 12) orc_classify(P, rhoc, ires.data(), out->napm);
  if ((mask & 4) && P->pm3) {
    float f2 = 0, vm[3] = {0, 0, 0};
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t it = 0; it < ns3; ++it) {
      int64_t T, ifs[3];
      subtile_ids(g, it, &T, ifs);
      if (T % stride) continue;
      float f2t, vt[3];
      LevelBox b = pm3_box(g, T, ifs, ires[it]);
      pm_box(g, xp, vp, rhoc, off, tf[ires[it]], b, *kin, &f2t, vt);
      atomic_max_f(&f2, f2t);
      for (int d = 0; d < 3; ++d) atomic_max_f(&vm[d], vt[d]);
    }
    out->f2max_pm3 = f2;
    out->dt_pm3 = 0.5f * std::sqrt(1.0f / ((std::sqrt(f2) * (float)kin->a_mid) * gg));
    for (int d = 0; d < 3; ++d) vmax[d] = vm[d];
  }
  g_phase_s[2] = omp_get_wtime() - t0;
  t0 = omp_get_wtime();
  int64_t totpairs = 0;
  if ((mask & 8) && P->pp) {
    float f2 = 0, vm[3] = {0, 0, 0};
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : totpairs)
    for (int64_t it = 0; it < ns3; ++it) {
      int64_t T, ifs[3];
      subtile_ids(g, it, &T, ifs);
      if (T % stride) continue;
      float f2t = 0, vt[3] = {0, 0, 0};
      int64_t np = 0;
      pp_box(g, *P, xp, vp, rhoc, off, T, ifs, ires[it], *kin, &f2t, vt, &np);
      totpairs += np;
      atomic_max_f(&f2, f2t);
      for (int d = 0; d < 3; ++d) atomic_max_f(&vm[d], vt[d]);
    }
    out->f2max_pp = f2;
    out->dt_pp = 0.05f * std::sqrt(1.0f / ((std::sqrt(f2) * (float)kin->a_mid) * gg));
    for (int d = 0; d < 3; ++d) vmax[d] = vm[d];
  }
  g_phase_s[3] = omp_get_wtime() - t0;
  if (npairs_out) *npairs_out = totpairs;
  for (int d = 0; d < 3; ++d) out->vmax[d] = vmax[d];
  float mv = std::max({vmax[0], vmax[1], vmax[2]});
  out->dt_vmax = (0.9f * 14.0f) / mv;  // kick.f90:314 vbuf*14/maxval(vmax)
  out->vz_max = vmax[2];
  return 0;
}

// Per-level work on a single box, for focused tests: level 2 = PM2 tile T,
// level 3 = PM3 sub-tile `it` at resolution iapm.
int orc_pm_box(const OParams* P, const void* xp, float* vp, const int32_t* rhoc, const float* tf, int level,
               int64_t index, int iapm, const OKickIn* kin, float* f2max, float vmax[3]) {
  Geo g(*P);
  std::vector<int64_t> off = cell_offsets(g, rhoc);
  LevelBox b;
  if (level == 2) b = pm2_box(g, index);
  else {
    int64_t T, ifs[3];
    subtile_ids(g, index, &T, ifs);
    b = pm3_box(g, T, ifs, iapm);
  }
  pm_box(g, xp, vp, rhoc, off, tf, b, *kin, f2max, vmax);
  return 0;
}

// PP for one sub-tile (focused tests).
int orc_pp_box(const OParams* P, const void* xp, float* vp, const int32_t* rhoc, int64_t it, int iapm,
               const OKickIn* kin, float* f2max, float vmax[3], int64_t* npairs) {
  Geo g(*P);
  std::vector<int64_t> off = cell_offsets(g, rhoc);
  int64_t T, ifs[3];
  subtile_ids(g, it, &T, ifs);
  pp_box(g, *P, xp, vp, rhoc, off, T, ifs, iapm, *kin, f2max, vmax, npairs);
  return 0;
}

int orc_f_ra(int64_t n, const float* r, float apm, float* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = f_ra(r[i], apm);
  return 0;
}

// spine_image (variables.f90:225-295); values are the reference's counts
// (idx_b_l = particles before the row, idx_b_r = particles through the row).
int orc_spine_image(const OParams* P, const int32_t* rhoc, int64_t* idx_b_l, int64_t* idx_b_r, int64_t* ppe_l0,
                    int64_t* ppe_lr, int64_t* ppe_rl, int64_t* ppe_r0, int64_t* ppl, int64_t* ppr) {
  Geo g(*P);
  const int64_t nt = g.nt, ncb = g.ncb, nte = g.nte;
  std::vector<int64_t> bl(g.ntiles() * nte * nte), br(g.ntiles() * nte * nte);
  int64_t nsum = 0;
  for (int64_t T = 0; T < g.ntiles(); ++T)
    for (int64_t k = 1 - ncb; k <= nt + ncb; ++k)
      for (int64_t j = 1 - ncb; j <= nt + ncb; ++j) {
        int64_t r = (T * nte + (k - 1 + ncb)) * nte + (j - 1 + ncb);
        bl[r] = nsum;
        for (int64_t i = 1 - ncb; i <= nt + ncb; ++i) nsum += rhoc[g.rc(i, j, k, T)];
        br[r] = nsum;
      }
  if (idx_b_l) std::memcpy(idx_b_l, bl.data(), sizeof(int64_t) * bl.size());
  if (idx_b_r) std::memcpy(idx_b_r, br.data(), sizeof(int64_t) * br.size());
  int64_t nsum_p = 0;
  for (int64_t T = 0; T < g.ntiles(); ++T)
    for (int64_t k = 1; k <= nt; ++k)
      for (int64_t j = 1; j <= nt; ++j) {
        int64_t q = (T * nt + (k - 1)) * nt + (j - 1);
        int64_t r = (T * nte + (k - 1 + ncb)) * nte + (j - 1 + ncb);
        int64_t np_phy = 0, lsum = 0, rsum = 0, rl = 0, lr = 0;
        for (int64_t i = 1; i <= nt; ++i) np_phy += rhoc[g.rc(i, j, k, T)];
        for (int64_t i = 1 - ncb; i <= 0; ++i) lsum += rhoc[g.rc(i, j, k, T)];
        for (int64_t i = nt + 1; i <= nt + ncb; ++i) rsum += rhoc[g.rc(i, j, k, T)];
        for (int64_t i = nt - ncb + 1; i <= nt; ++i) rl += rhoc[g.rc(i, j, k, T)];
        for (int64_t i = 1; i <= ncb; ++i) lr += rhoc[g.rc(i, j, k, T)];
        if (ppl) ppl[q] = nsum_p;
        nsum_p += np_phy;
        if (ppr) ppr[q] = nsum_p;
        if (ppe_r0) ppe_r0[q] = br[r] - rsum;
        if (ppe_rl) ppe_rl[q] = br[r] - rsum - rl;
        if (ppe_l0) ppe_l0[q] = bl[r] + lsum;
        if (ppe_lr) ppe_lr[q] = bl[r] + lsum + lr;
      }
  return 0;
}

// drift (drift.f90:1-233), serial in the reference's loop order.
// Inputs: extended layout.  Outputs: physical layout of nplocal particles,
// rhoc_out with zero buffers.  Returns a CUBE-style status code.
int orc_drift(const OParams* P, const void* xp, const float* vp, const int64_t* pid, const int32_t* rhoc,
              const ODriftIn* in, void* xp_out, float* vp_out, int64_t* pid_out, int32_t* rhoc_out,
              int64_t np_out_cap, ODriftOut* out) {
  Geo g(*P);
  const int64_t nt = g.nt, ncb = g.ncb, nnt = g.nnt;
  std::vector<int64_t> off = cell_offsets(g, rhoc);
  // drift.f90:21 and :26, with the coefficient dt_mid in FP64
  const double dt_mid = (in->dt_old + in->dt) / 2.0;
  const int nlayer = 2 * (int)std::ceil((dt_mid * (double)in->vz_max) / (double)g.rcs) + 2;
  out->nlayer = nlayer;
  const int64_t ne = nt + 4 * ncb;  // rhoce(1-2ncb:nt+2ncb)
  auto RE = [&](int64_t i, int64_t j, int64_t k) { return ((k - 1 + 2 * ncb) * ne + (j - 1 + 2 * ncb)) * ne + (i - 1 + 2 * ncb); };
  // np_tile_max = np_image/nnt**3*(nte*1./nt)**3*tile_buffer (variables.f90:26)
  int64_t np_image = (int64_t)g.ng * g.ng * g.ng;
  float ratio = (float)g.nte / (float)g.nt;
  int64_t np_tile_max = (int64_t)((float)(np_image / (nnt * nnt * nnt)) * (ratio * ratio * ratio) * P->tile_buffer);
  int64_t np_prev = 0;
  float overhead = 0;
  std::memset(rhoc_out, 0, sizeof(int32_t) * g.ntiles() * g.ncell_tile());
  std::vector<int32_t> rhoce(ne * ne * ne), rholocal(ne * ne * ne);
  std::vector<int64_t> xnew, pnew;
  std::vector<float> vnew;
  const double t0 = omp_get_wtime();
  for (int64_t T = 0; T < g.ntiles(); ++T) {
    if (T % g_sample_stride) continue;
    std::fill(rhoce.begin(), rhoce.end(), 0);
    std::fill(rholocal.begin(), rholocal.end(), 0);
    auto dest = [&](int64_t ip, const int64_t ijk[3], int64_t gi[3]) -> bool {
      bool ok = true;
      for (int d = 0; d < 3; ++d) {
        double dx = (dt_mid * (double)vp[3 * ip + d]) / (double)g.rcs;
        double xq = ((double)ijk[d] - 1.0) + decode(xp, ip, d, g);
        gi[d] = (int64_t)std::ceil(xq + dx);
        if (gi[d] < 1 - 2 * ncb || gi[d] > nt + 2 * ncb) ok = false;
      }
      return ok;
    };
    for (int il = 0; il < nlayer; ++il)
      for (int64_t k = 1 - ncb + il; k <= nt + ncb; k += nlayer)
        for (int64_t j = 1 - ncb; j <= nt + ncb; ++j)
          for (int64_t i = 1 - ncb; i <= nt + ncb; ++i) {
            int64_t c = g.rc(i, j, k, T), ijk[3] = {i, j, k}, gi[3];
            for (int64_t ip = off[c]; ip < off[c] + rhoc[c]; ++ip) {
              if (!dest(ip, ijk, gi)) return 5;
              rhoce[RE(gi[0], gi[1], gi[2])]++;
            }
          }
    // spine_tile (variables.f90:180-223)
    std::vector<int64_t> idx_ex_r(ne * ne);
    int64_t nsum = 0;
    for (int64_t k = 1 - 2 * ncb; k <= nt + 2 * ncb; ++k)
      for (int64_t j = 1 - 2 * ncb; j <= nt + 2 * ncb; ++j) {
        for (int64_t i = 1 - 2 * ncb; i <= nt + 2 * ncb; ++i) nsum += rhoce[RE(i, j, k)];
        idx_ex_r[(k - 1 + 2 * ncb) * ne + (j - 1 + 2 * ncb)] = nsum;
      }
    overhead = std::max(overhead, (float)nsum / (float)np_tile_max);
    if (nsum > np_tile_max) return 3;
    xnew.assign(3 * std::max<int64_t>(nsum, 1), 0);
    vnew.assign(3 * std::max<int64_t>(nsum, 1), 0.0f);
    pnew.assign(std::max<int64_t>(nsum, 1), 0);
    // cell start in the tile-extended arrays (0-based)
    std::vector<int64_t> cst(ne * ne * ne);
    {
      int64_t s = 0;
      for (size_t q = 0; q < cst.size(); ++q) { cst[q] = s; s += rhoce[q]; }
    }
    for (int il = 0; il < nlayer; ++il)
      for (int64_t k = 1 - ncb + il; k <= nt + ncb; k += nlayer)
        for (int64_t j = 1 - ncb; j <= nt + ncb; ++j)
          for (int64_t i = 1 - ncb; i <= nt + ncb; ++i) {
            int64_t c = g.rc(i, j, k, T), ijk[3] = {i, j, k}, gi[3];
            for (int64_t ip = off[c]; ip < off[c] + rhoc[c]; ++ip) {
              dest(ip, ijk, gi);
              int64_t e = RE(gi[0], gi[1], gi[2]);
              int64_t idx = cst[e] + rholocal[e];
              rholocal[e]++;
              for (int d = 0; d < 3; ++d) {
                const double wv = dt_mid * (double)vp[3 * ip + d];
                double sh = wv / (g.xres * (double)g.rcs);
                int64_t dn = (int64_t)std::round(sh);  // nint: half away from zero
                int64_t xv = (g.izipx == 2) ? (int64_t)((const int16_t*)xp)[3 * ip + d] : (int64_t)((const int8_t*)xp)[3 * ip + d];
                xnew[3 * idx + d] = xv + dn;
                vnew[3 * idx + d] = vp[3 * ip + d];
              }
              pnew[idx] = pid ? pid[ip] : 1;
              if (pid && pid[ip] == 0) return 7;
            }
          }
    for (size_t q = 0; q < rhoce.size(); ++q)
      if (rhoce[q] != rholocal[q]) return 6;
    // compaction (drift.f90:153-173): physical cells in (k, j, i) order.
    for (int64_t k = 1; k <= nt; ++k)
      for (int64_t j = 1; j <= nt; ++j)
        for (int64_t i = 1; i <= nt; ++i) {
          int64_t e = RE(i, j, k);
          for (int64_t q = 0; q < rhoce[e]; ++q) {
            int64_t src = cst[e] + q;
            if (np_prev >= np_out_cap) return 4;
            for (int d = 0; d < 3; ++d) {
              int64_t v = xnew[3 * src + d];
              if (g.izipx == 2) ((int16_t*)xp_out)[3 * np_prev + d] = (int16_t)(uint16_t)(v & 0xffff);
              else ((int8_t*)xp_out)[3 * np_prev + d] = (int8_t)(uint8_t)(v & 0xff);
              vp_out[3 * np_prev + d] = vnew[3 * src + d];
            }
            if (pid_out) pid_out[np_prev] = pnew[src];
            ++np_prev;
          }
          rhoc_out[g.rc(i, j, k, T)] = rhoce[e];
        }
  }
  g_phase_s[4] = omp_get_wtime() - t0;
  out->nplocal = np_prev;
  out->overhead_tile = overhead;
  if (pid_out)
    for (int64_t q = 0; q < np_prev; ++q)
      if (pid_out[q] < 1) return 7;
  return 0;
}

// buffer_grid + buffer_x + buffer_v at nn=1: periodic buffer copies.
int orc_buffer(const OParams* P, const void* xp, const float* vp, const int64_t* pid, const int32_t* rhoc_phys,
               void* xp_ext, float* vp_ext, int64_t* pid_ext, int32_t* rhoc_ext, int64_t np_cap, int64_t* np_ext) {
  Geo g(*P);
  const int64_t nt = g.nt, ncb = g.ncb, nnt = g.nnt, nc = g.nc;
  // physical offsets in tile order, cells (k,j,i)
  std::vector<int64_t> poff(g.ntiles() * nt * nt * nt);
  int64_t s = 0;
  for (int64_t T = 0; T < g.ntiles(); ++T)
    for (int64_t k = 1; k <= nt; ++k)
      for (int64_t j = 1; j <= nt; ++j)
        for (int64_t i = 1; i <= nt; ++i) {
          poff[((T * nt + k - 1) * nt + j - 1) * nt + i - 1] = s;
          s += rhoc_phys[g.rc(i, j, k, T)];
        }
  int64_t q = 0;
  int esz = g.izipx;
  for (int64_t T = 0; T < g.ntiles(); ++T) {
    int64_t it[3] = {T % nnt, (T / nnt) % nnt, T / (nnt * nnt)};
    for (int64_t k = 1 - ncb; k <= nt + ncb; ++k)
      for (int64_t j = 1 - ncb; j <= nt + ncb; ++j)
        for (int64_t i = 1 - ncb; i <= nt + ncb; ++i) {
          int64_t ijk[3] = {i, j, k}, st[3], sc[3];
          for (int d = 0; d < 3; ++d) {
            int64_t gc = it[d] * nt + ijk[d] - 1;
            gc = ((gc % nc) + nc) % nc;
            st[d] = gc / nt; sc[d] = gc % nt + 1;
          }
          int64_t Ts = (st[2] * nnt + st[1]) * nnt + st[0];
          int32_t cnt = rhoc_phys[g.rc(sc[0], sc[1], sc[2], Ts)];
          rhoc_ext[g.rc(i, j, k, T)] = cnt;
          int64_t src = poff[((Ts * nt + sc[2] - 1) * nt + sc[1] - 1) * nt + sc[0] - 1];
          if (q + cnt > np_cap) return 4;
          std::memcpy((char*)xp_ext + 3 * esz * q, (const char*)xp + 3 * esz * src, 3 * esz * cnt);
          std::memcpy(vp_ext + 3 * q, vp + 3 * src, sizeof(float) * 3 * cnt);
          if (pid && pid_ext) std::memcpy(pid_ext + q, pid + src, sizeof(int64_t) * cnt);
          q += cnt;
        }
  }
  *np_ext = q;
  return 0;
}

// buffer_v at nn=1 (buffer_particle.f90:11-27 -> buffer_xvp(.false., .true.),
// 29-375): in the extended layout, every buffer cell's velocities and IDs are
// copied from the physical cell it is the periodic image of (the same cell
// mapping as buffer_x); physical cells are untouched.  pid may be NULL.
int orc_buffer_v(const OParams* P, float* vp, int64_t* pid, const int32_t* rhoc_ext, int64_t np) {
  Geo g(*P);
  const int64_t nt = g.nt, ncb = g.ncb, nnt = g.nnt, nc = g.nc, nte = nt + 2 * ncb;
  // extended offsets in storage order
  const int64_t ncell = g.ntiles() * nte * nte * nte;
  std::vector<int64_t> eoff(ncell + 1, 0);
  for (int64_t c = 0; c < ncell; ++c) eoff[c + 1] = eoff[c] + rhoc_ext[c];
  if (eoff[ncell] > np) return 4;
  for (int64_t T = 0; T < g.ntiles(); ++T) {
    int64_t it[3] = {T % nnt, (T / nnt) % nnt, T / (nnt * nnt)};
    for (int64_t k = 1 - ncb; k <= nt + ncb; ++k)
      for (int64_t j = 1 - ncb; j <= nt + ncb; ++j)
        for (int64_t i = 1 - ncb; i <= nt + ncb; ++i) {
          if (i >= 1 && i <= nt && j >= 1 && j <= nt && k >= 1 && k <= nt) continue;
          int64_t ijk[3] = {i, j, k}, st[3], sc[3];
          for (int d = 0; d < 3; ++d) {
            int64_t gc = it[d] * nt + ijk[d] - 1;
            gc = ((gc % nc) + nc) % nc;
            st[d] = gc / nt; sc[d] = gc % nt + 1;
          }
          const int64_t Ts = (st[2] * nnt + st[1]) * nnt + st[0];
          const int64_t cs = g.rc(sc[0], sc[1], sc[2], Ts), cd = g.rc(i, j, k, T);
          if (rhoc_ext[cs] != rhoc_ext[cd]) return 8;  // not a buffered layout
          std::memcpy(vp + 3 * eoff[cd], vp + 3 * eoff[cs], sizeof(float) * 3 * rhoc_ext[cd]);
          if (pid) std::memcpy(pid + eoff[cd], pid + eoff[cs], sizeof(int64_t) * rhoc_ext[cd]);
        }
  }
  return 0;
}

// ---------------------------------------------------------------------------
// CIC power spectrum of the physical particles (utilities/cicpower.f90:67-179
// with utilities/powerspectrum.f90:140-259 auto_power and pk_correction),
// nn = 1.  Per-mode quantities in the reference's precision (FP32 positions,
// weights, kr, amp, C1k); the CIC mesh accumulates in FP32 in the reference's
// loop order; bin sums are FP64 (the reference's FP32 bin sums are
// order-dependent).  xi is (10, 0:nbin) column-major: xi[b*10 + c].
static double interp1_ref(const std::vector<double>& x, const std::vector<double>& y, int nbin, double xq) {
  // powerspectrum.f90:261-275 on xdata(1:nbin) = bins 0..nbin-1
  int i1 = 1, i2 = nbin;
  while (i2 - i1 > 1) {
    int im = (i1 + i2) / 2;
    if (xq > x[im - 1]) i1 = im;
    else i2 = im;
  }
  return y[i1 - 1] + (xq - x[i1 - 1]) / (x[i2 - 1] - x[i1 - 1]) * (y[i2 - 1] - y[i1 - 1]);
}

int orc_cicpower(const OParams* P, const void* xp, const int32_t* rhoc_phys, int32_t nw, float box, float* delta,
                 double* mean_out, float* xi_out, int32_t nbin) {
  Geo g(*P);
  const int64_t nt = g.nt, nnt = g.nnt, nc = g.nc;
  const int64_t m = nw + 2;
  std::vector<float> grid((size_t)m * m * m, 0.0f);
  auto G = [&](int64_t i, int64_t j, int64_t k) -> float& { return grid[((size_t)k * m + j) * m + i]; };  // 0..nw+1
  int64_t nlast = 0;
  for (int64_t itz = 1; itz <= nnt; ++itz)
    for (int64_t ity = 1; ity <= nnt; ++ity)
      for (int64_t itx = 1; itx <= nnt; ++itx) {
        const int64_t T = g.tile_lin(itx, ity, itz);
        const int64_t it[3] = {itx, ity, itz};
        for (int64_t k = 1; k <= nt; ++k)
          for (int64_t j = 1; j <= nt; ++j)
            for (int64_t i = 1; i <= nt; ++i) {
              const int64_t np = rhoc_phys[g.rc(i, j, k, T)];
              const int64_t ijk[3] = {i, j, k};
              for (int64_t l = 1; l <= np; ++l) {
                const int64_t ip = nlast + l - 1;
                float pos[3], dx1[3], dx2[3];
                int64_t i1[3], i2[3];
                for (int d = 0; d < 3; ++d) {
                  pos[d] = (float)((double)(nt * (it[d] - 1) + (ijk[d] - 1)) + decode(xp, ip, d, g));
                  pos[d] = (pos[d] * (float)nw) / (float)nc - 0.5f;  // cicpower.f90:79
                  i1[d] = (int64_t)std::floor(pos[d]) + 1;
                  i2[d] = i1[d] + 1;
                  dx1[d] = (float)i1[d] - pos[d];
                  dx2[d] = 1.0f - dx1[d];
                }
                G(i1[0], i1[1], i1[2]) += (dx1[0] * dx1[1]) * dx1[2];
                G(i2[0], i1[1], i1[2]) += (dx2[0] * dx1[1]) * dx1[2];
                G(i1[0], i2[1], i1[2]) += (dx1[0] * dx2[1]) * dx1[2];
                G(i1[0], i1[1], i2[2]) += (dx1[0] * dx1[1]) * dx2[2];
                G(i1[0], i2[1], i2[2]) += (dx1[0] * dx2[1]) * dx2[2];
                G(i2[0], i1[1], i2[2]) += (dx2[0] * dx1[1]) * dx2[2];
                G(i2[0], i2[1], i1[2]) += (dx2[0] * dx2[1]) * dx1[2];
                G(i2[0], i2[1], i2[2]) += (dx2[0] * dx2[1]) * dx2[2];
              }
              nlast += np;
            }
      }
  // periodic halo, x then y then z (cicpower.f90:108-156 at nn = 1)
  for (int64_t a = 0; a < m; ++a)
    for (int64_t b = 0; b < m; ++b) { G(1, a, b) += G(nw + 1, a, b); G(nw, a, b) += G(0, a, b); }
  for (int64_t a = 0; a < m; ++a)
    for (int64_t b = 0; b < m; ++b) { G(a, 1, b) += G(a, nw + 1, b); G(a, nw, b) += G(a, 0, b); }
  for (int64_t a = 0; a < m; ++a)
    for (int64_t b = 0; b < m; ++b) { G(a, b, 1) += G(a, b, nw + 1); G(a, b, nw) += G(a, b, 0); }
  const int64_t n3 = (int64_t)nw * nw * nw;
  std::vector<float> rc((size_t)n3);
  double rho8 = 0;
  for (int64_t k = 1; k <= nw; ++k)
    for (int64_t j = 1; j <= nw; ++j)
      for (int64_t i = 1; i <= nw; ++i) {
        float v = G(i, j, k);
        rc[((size_t)(k - 1) * nw + (j - 1)) * nw + (i - 1)] = v;
        rho8 += (double)v;
      }
  const double mean = rho8 / (double)nw / (double)nw / (double)nw;
  for (auto& v : rc) v = (float)((double)v / mean - 1.0);  // cicpower.f90:176-179
  if (delta) std::memcpy(delta, rc.data(), sizeof(float) * n3);
  if (mean_out) *mean_out = mean;
  // auto_power (powerspectrum.f90:140-200)
  const int64_t nh = nw / 2 + 1, nyq = nw / 2;
  std::vector<cd> ck((size_t)nh * nw * nw);
  r2c3d(nw, rc.data(), ck.data());
  std::vector<double> x1(nbin + 1, 0), x2(nbin + 1, 0), x3(nbin + 1, 0), x4(nbin + 1, 0);
  const float fnw = (float)nw;
  for (int64_t kg = 1; kg <= nw; ++kg)
    for (int64_t jg = 1; jg <= nw; ++jg)
      for (int64_t ig = 1; ig <= nyq + 1; ++ig) {
        const int64_t kk[3] = {((ig + nyq - 1) % nw) - nyq, ((jg + nyq - 1) % nw) - nyq, ((kg + nyq - 1) % nw) - nyq};
        if (ig == 1 && jg == 1 && kg == 1) continue;
        if ((ig == 1 || ig == nyq + 1) && jg > nyq + 1) continue;
        if ((ig == 1 || ig == nyq + 1) && (jg == 1 || jg == nyq + 1) && kg > nyq + 1) continue;
        const float kx[3] = {(float)kk[0], (float)kk[1], (float)kk[2]};
        const float kr = std::sqrt((kx[0] * kx[0] + kx[1] * kx[1]) + kx[2] * kx[2]);
        const int64_t ib = (int64_t)std::lround(kr);
        if (ib > nbin) continue;
        const cd c = ck[((size_t)(kg - 1) * nw + (jg - 1)) * nh + (ig - 1)];
        const float re = (((float)c.real() / fnw) / fnw) / fnw, im = (((float)c.imag() / fnw) / fnw) / fnw;
        const float amp = re * re + im * im;
        float c1[3];
        for (int d = 0; d < 3; ++d) {
          const double sn = std::sin(PI * (double)kx[d] / (double)nw);
          c1[d] = (float)(1.0 - (double)(2.0f / 3.0f) * (sn * sn));
        }
        int64_t npart = nlast;
        const float Dk = ((c1[0] * c1[1]) * c1[2]) / (float)npart;
        x1[ib] += 1;
        x2[ib] += kr;
        x3[ib] += amp;
        x4[ib] += (double)(amp - Dk);
      }
  std::vector<double> x5(nbin + 1);
  for (int b = 0; b <= nbin; ++b) {
    x2[b] /= x1[b]; x3[b] /= x1[b]; x4[b] /= x1[b]; x5[b] = x4[b];
  }
  // pk_correction (powerspectrum.f90:213-259), applied while alpha is a number
  for (int iter = 0; iter < 4; ++iter) {
    const double alpha = (std::log(interp1_ref(x2, x5, nbin, (double)(float)nyq)) -
                          std::log(interp1_ref(x2, x5, nbin, (double)(float)nyq / 2))) / std::log(2.0);
    if (std::isnan(alpha)) break;
    std::vector<double> c0(nbin + 1, 0), c2(nbin + 1, 0);
#pragma omp parallel
    {
      std::vector<double> l0(nbin + 1, 0), l2(nbin + 1, 0);
#pragma omp for schedule(dynamic)
      for (int64_t i = 1; i <= nyq + 1; ++i)
        for (int64_t j = 1; j <= i; ++j)
          for (int64_t k = 1; k <= j; ++k) {
            const double kv[3] = {(double)(i - 1), (double)(j - 1), (double)(k - 1)};
            const double kmag = std::sqrt(kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2]);
            const int64_t ib = (int64_t)std::lround(kmag);
            if (ib < 1 || ib > nbin) continue;
            double w = 0;
            for (int in = -3; in <= 3; ++in)
              for (int jn = -3; jn <= 3; ++jn)
                for (int kn = -3; kn <= 3; ++kn) {
                  const double kvn[3] = {kv[0] + in * (double)nw, kv[1] + jn * (double)nw, kv[2] + kn * (double)nw};
                  const double kmn = std::sqrt(kvn[0] * kvn[0] + kvn[1] * kvn[1] + kvn[2] * kvn[2]);
                  double pr = 1;
                  for (int d = 0; d < 3; ++d) {
                    const double ks = PI * kvn[d] / (double)nw;
                    pr *= (ks == 0) ? 1.0 : std::sin(ks) / ks;
                  }
                  const double p2 = pr * pr;
                  w += (p2 * p2) * std::pow(kmn, alpha);  // sinc^(2p), p = 2 (CIC)
                }
            l0[ib] += 1;
            l2[ib] += w / std::pow(kmag, alpha);
          }
#pragma omp critical(pkc)
      for (int b = 0; b <= nbin; ++b) { c0[b] += l0[b]; c2[b] += l2[b]; }
    }
    for (int b = 1; b <= nbin; ++b) x5[b] = x4[b] / (c2[b] / c0[b]);
  }
  const double kfac = 2 * PI / (double)box, b3 = (double)box * box * box;
  for (int b = 0; b <= nbin; ++b) {
    float* o = xi_out + (size_t)b * 10;
    for (int c = 0; c < 10; ++c) o[c] = 0;
    o[0] = (float)x1[b];
    o[1] = (float)(x2[b] * kfac);
    o[2] = (float)(x3[b] * b3);
    o[3] = (float)(x4[b] * b3);
    o[4] = (float)(x5[b] * b3);
  }
  return 0;
}

int orc_num_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
