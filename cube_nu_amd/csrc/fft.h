// fft.h -- batched 3-D real FFT convolution on cubes (LDS-staged Stockham).
//
// Replaces the FFTW plans of cubefft.f90:5-38 (tile / sub-tile meshes) and
// pencil_fft.f90:23-79, 327-338 (coarse mesh at nn=1).  Conventions follow
// FFTW: forward r2c e^{-i}, backward c2r e^{+i}, both unnormalised.
#pragma once
#include "common.h"

namespace cube {

struct FFTStages {
  int n;             // line length
  int nst;           // number of radix stages
  int rad[16];
};

FFTStages make_stages(int n);

// One cube size: twiddle tables for n (full, used by r2c post-processing and
// the y/z passes) and the stage factorisations of n and n/2.
template <typename R>
struct FFTCube {
  int n = 0;
  FFTStages full{}, half{};
  typename Cplx<R>::T* tw = nullptr;  // e^{-2 pi i t / n}, t = 0..n-1
  void init(int n_);
  void release();
};

// Convolve `nbox` real cubes in place with a real kernel G(kx,ky,kz):
//   phi = c2r( G * r2c(rho) ) / div
// rho/phi: [box][z][y][x] real (n^3 per box); work: complex scratch of
// nbox * n*n*(n/2+1) elements.  The division follows the reference: one FP
// division by div3 when div3 != 0, else three divisions by div1.
// shape: launch shape of the size-specialised passes (0, 1 or 2; the same
// result for each), picked by the caller's autotuner; -1 selects the generic
// line passes for every size.
template <typename R>
void fft_convolve(const FFTCube<R>& c, R* rho, typename Cplx<R>::T* work, const float* G, int64_t nbox, R div3,
                  R div1, hipStream_t s, KTimer* tm = nullptr, const char* tag = "", int shape = 0);

// Size-specialised 3-pass form (fft_plane.hip); returns false when n has no
// specialisation.
template <typename R>
bool fft_convolve_plane(const FFTCube<R>& c, R* rho, typename Cplx<R>::T* work, const float* G, int64_t nbox,
                        R div3, R div1, hipStream_t s, KTimer* tm, const char* tag, int shape = 0);
// True when n has a size-specialised form with several launch shapes (FP32).
bool fft_plane_tunable(int n);

// Forward-only and inverse-only pieces (used by the input maker and tests).
template <typename R>
void fft_r2c3d(const FFTCube<R>& c, const R* rho, typename Cplx<R>::T* out, int64_t nbox, hipStream_t s);
template <typename R>
void fft_c2r3d(const FFTCube<R>& c, typename Cplx<R>::T* in, R* out, int64_t nbox, hipStream_t s);

}  // namespace cube
