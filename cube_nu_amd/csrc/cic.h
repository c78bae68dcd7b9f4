// cic.h -- CIC density and power spectrum of the resident particles
// (utilities/cicpower.f90:67-179, utilities/powerspectrum.f90:140-259).
#pragma once
#include "common.h"

namespace cube {

// CIC P(k) of the physical particles on an nw^3 periodic mesh.
//   xi:    host, (10, 0:nbin) column-major (xi[b*10 + c]); columns 1..5 are
//          filled: mode count, mean k (h/Mpc), raw power, shot-noise
//          subtracted power, alias-corrected power ((Mpc/h)^3);
//   delta: host nw^3 density contrast [z][y][x], or null;
//   mean:  mean CIC mass per cell (FP64), or null.
// rhoc/off: counts and cell offsets of the current layout (physical cells are
// the tile-local cells 1..nt in either layout).
void cic_power(const Geo& g, int nw, float box, const void* xp, const int32_t* rhoc, const int64_t* off,
               int64_t npart, float* xi, int nbin, float* delta, double* mean, hipStream_t s);

}  // namespace cube
