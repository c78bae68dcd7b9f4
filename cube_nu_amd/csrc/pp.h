// pp.h -- short-range particle-particle force (PP_Force, kick.f90:376-588).
#pragma once
#include "common.h"

namespace cube {

// PP cells are split into kSX sub-bins along x in the binned CSR, so that the
// partners of a target along a (dy, dz) column are one short contiguous run
// (8 sub-bins: 25.4 ms at 512^3 against 26.1 with 4 and 25.8 with 16, the
// window's edges cutting closer to x +- r_pp against a larger staged table).
constexpr int kSX = 8;

struct PPLevel {
  int sf;          // ratio_sf(iapm) = rcp
  int npgrid;      // ntt * rcp
  int ntt, nte, ncb, rcs;
  float pp_range, pp_range2, app;
  float c_app, c_range;      // (1./35.)/apm**2 for app and pp_range
  float inv_app, inv_range;  // 1/app, 1/pp_range
  float two_inv_range;       // 2/pp_range
  float app2;                // app^2
  float mass_scale;
  double v_scale;  // a_mid*dt/6/pi in FP64 (kick.f90:574 forms it in real(4))
  const double* vdev = nullptr;  // v_scale in device memory, read in its place when set
  float inv_rcs;   // 1/rcs: fine -> coarse units (exact for a power-of-two rcs)
  float rcut2;     // (r_pp in PP-cell units + margin)^2: y/z column cut
};

struct PPBox {
  int T;
  int ifs[3];      // 0-based sub-tile coordinates within the tile
};

struct PPOut {
  float f2max;
  float vmax[3];
  unsigned long long pairs_pp[kSumSlots];  // directed physical-physical pair count
  unsigned long long pairs_pg[kSumSlots];  // physical-ghost pair count
};

// Scratch the caller provides for one level batch.
struct PPScratch {
  int32_t* count = nullptr;   // nbox * m^3 * kSX sub-bins
  int64_t* start = nullptr;   // nbox * m^3 * kSX + 1
  float4* entries = nullptr;  // capacity entries + kPPPad readable padding
  int64_t capacity = 0;
  unsigned* claim = nullptr;  // 8 work-queue counters of the force kernel
  void* scan_tmp = nullptr;
  size_t scan_tmp_bytes = 0;
};

// Bin the particles of each sub-tile (plus its ghost ring) into its PP grid.
// Returns the number of entries needed (host sync); fills when it fits.
int64_t pp_bin_count(int izipx, const PPLevel& L, const PPBox* boxes_dev, int nbox, const void* xp,
                     const int32_t* rhoc, const int64_t* cell_off, PPScratch& S, hipStream_t s);
void pp_bin_fill(int izipx, const PPLevel& L, const PPBox* boxes_dev, int nbox, const void* xp, const int32_t* rhoc,
                 const int64_t* cell_off, PPScratch& S, hipStream_t s);
// unroll: candidate loads in flight per lane in the FP32 kernel (4, 6 or 8),
// plus kPPBrick8 for bricks of 8^3 PP cells instead of 4^3; every value gives
// the same result.
constexpr int kPPBrick8 = 100;
void pp_force(const PPLevel& L, int precision, const PPBox* boxes_dev, int nbox, float* vp, PPScratch& S, PPOut* out,
              hipStream_t s, int unroll = 6);

// Entries past the last one that the force kernel may read (and ignore).
constexpr int kPPPad = 16;

}  // namespace cube
