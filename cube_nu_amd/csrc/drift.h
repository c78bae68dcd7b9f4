// drift.h -- drift with re-sort/re-encode (drift.f90:1-233) and the nn=1
// buffer layout (buffer_grid.f90:1-226, buffer_particle.f90:1-375).
#pragma once
#include "common.h"

namespace cube {

constexpr int kMaxShift = 16;  // largest cell shift the offset pass orders

struct DriftParams {
  double dt_mid;  // (dt_old + dt)/2 in FP64 (the reference's real(4) coefficient, kept FP64 here)
  int nlayer;
  int nt, nte, ncb, nnt, rcs, izipx;
  double inv_rcs;  // 1/rcs, exact (rcs is a power of two), so x * inv_rcs == x / rcs
  int M[3];       // maximum |cell shift| per axis
  int D[3];       // 2M+1
};

struct DriftScratch {
  int32_t* flags = nullptr;     // [0] error bits, [1..3] max shift per axis
  int32_t* rank = nullptr;      // per source particle
  int32_t* table = nullptr;     // per source cell per direction (count -> offset)
  int64_t table_cap = 0;        // int32 elements
  int32_t* newcount = nullptr;  // ntiles * nt^3 physical destination counts
  int64_t* base = nullptr;      // ntiles * nt^3 + 1 (scan of newcount)
  int64_t* tile_tot = nullptr;  // per tile physical total
  void* scan_tmp = nullptr;
  size_t scan_tmp_bytes = 0;
};

// Error bits written to flags[0].
enum { DRIFT_OUT_OF_RANGE = 1, DRIFT_PID_ZERO = 2, DRIFT_SHIFT_BOUND = 8, DRIFT_TILE_OVERFLOW = 16 };

// Tile overflow check (drift.f90:85-93) on the device: flags[4] = the largest
// extended tile count (off[(T+1) nct] - off[T nct]), DRIFT_TILE_OVERFLOW in
// flags[0] when one exceeds np_tile_max.
void drift_tile_check(const int64_t* off, int64_t ncell_tile, int ntiles, int64_t np_tile_max, int32_t* flags,
                      hipStream_t s);
// Pass 1: destination range check and maximum shift.
void drift_shift(const DriftParams& P, const void* xp, const float* vp, const int32_t* rhoc, const int64_t* off,
                 int64_t ncells, DriftScratch& S, hipStream_t s);
// Passes 2-5 for tiles [t0, t1): rank, offsets, scatter (base scan done by caller).
// count = true folds pass 1 into the rank pass for all tiles at once, with
// P.M an upper bound of the cell shift given by the caller (flags as pass 1,
// plus DRIFT_SHIFT_BOUND when a shift exceeds P.M).
void drift_rank(const DriftParams& P, const void* xp, const float* vp, const int32_t* rhoc, const int64_t* off,
                int t0, int t1, DriftScratch& S, hipStream_t s, bool count = false);
void drift_offsets(const DriftParams& P, int t0, int t1, DriftScratch& S, hipStream_t s);
void drift_scatter(const DriftParams& P, const void* xp, const float* vp, const int64_t* pid, const int32_t* rhoc,
                   const int64_t* off, int t0, int t1, DriftScratch& S, void* xp_out, float* vp_out,
                   int64_t* pid_out, hipStream_t s);
// rhoc with zero buffers and physical counts from newcount.
void drift_write_rhoc(const DriftParams& P, const DriftScratch& S, int32_t* rhoc_out, hipStream_t s);

// Buffer: physical layout (counts in rhoc_phys, extended array with zero
// buffers) -> extended layout.
void buffer_rhoc(const Geo& g, const int32_t* rhoc_phys, int32_t* rhoc_ext, int32_t* physcount, hipStream_t s);
void buffer_copy(const Geo& g, const int32_t* rhoc_ext, const int64_t* ext_off, const int64_t* phys_off,
                 const void* xp, const float* vp, const int64_t* pid, void* xp_out, float* vp_out, int64_t* pid_out,
                 hipStream_t s);
// buffer_v: in the extended layout (offsets ext_off, the scan of the extended
// counts), buffer velocities (and IDs, when pid is given) from the physical
// cells they are images of.
void buffer_v(const Geo& g, const int64_t* ext_off, float* vp, int64_t* pid, hipStream_t s);

}  // namespace cube
