// pm.h -- particle-mesh kernels shared by PM1 (coarse), PM2 (tile) and PM3
// (sub-tile) levels: TSC deposit, fused gradient + interpolation + kick,
// Green's function generation and sub-tile classification.
#pragma once
#include "common.h"

namespace cube {

// Per-level geometry (the type_pm record of variables.f90:94-113).
struct PMLevel {
  int n;             // mesh points per dim of one box array
  int64_t bstride;   // elements between consecutive boxes in the mesh array
  int m0;            // Fortran mesh index held at array index 0 (box-local + org)
  int periodic;      // 1 for the coarse mesh (buffer_potential wrap)
  int win_lo, win_hi;// deposit window (box-local Fortran mesh indices)
  int nloop;         // coarse ring read around the box's cells (pm%nloop)
  int w;             // fine mesh cells per coarse cell (integer, for bounds)
  int rcs;           // ratio_cs
  double gridsize;   // pm%gridsize as stored (real(4) promoted)
  double xmul;       // ratio_cs / gridsize (exact when gridsize is a power of two)
  int xdiv;          // 1: divide by gridsize (not a power of two), 0: multiply by xmul
  int nforce;        // pm%nforce: physical force cells 1..nforce
  int nte, ncb, nt;  // tile geometry for rhoc indexing
  // kick coefficient in device memory, read in place of the by-value one when
  // set (a replayed graph keeps its launch arguments; the coefficient changes
  // every step)
  const double* kdev = nullptr;
};

// One box: a tile (PM1, PM2) or a sub-tile (PM3).
struct PMBox {
  int T;             // 0-based tile index
  int nc1[3], nc2[3];// physical coarse cells (tile-local Fortran indices)
  int ushift[3];     // utile_shift * nfp(iapm)
  int org[3];        // global coarse offset (PM1), else 0
  int64_t moff;      // element offset of this box's mesh array
};

struct PMOut {  // device-side reductions of one level
  float f2max;
  float vmax[3];
  double f2max_d;
  double vmax_d[3];
  unsigned long long testrho_fx[kSumSlots];  // PM1 density sum in 2^-32 fixed point
  double testrho_d[kSumSlots];
};

// Tile/sub-tile (fine) level deposit into each box's mesh.
template <typename R>
void pm_deposit(int izipx, const PMLevel& L, const PMBox* boxes_dev, int nbox, const void* xp, const int32_t* rhoc,
                const int64_t* cell_off, R* mesh, float mass, hipStream_t s);

// Coarse (PM1) deposit over all tiles into the periodic nc^3 mesh, with the
// exact testrho sum; `stencil` holds pm1_stencil_bytes(L) of scratch.
template <typename R>
void pm1_deposit(int izipx, const PMLevel& L, const void* xp, const int32_t* rhoc, const int64_t* cell_off,
                 void* stencil, R* mesh, float mass, PMOut* out, hipStream_t s);
size_t pm1_stencil_bytes(const PMLevel& L);

// Fills xmul/xdiv from rcs and gridsize.
void pm_level_scale(PMLevel& L);

// shape: block shape of the cell-group form (0: 8x8x16 mesh cells with 256
// threads, 1: 8x16x16 with 512, 2: 16x16x8 with 512; FP32 only, FP64 uses 0);
// every shape gives the same result.
template <typename R>
void pm_interp(int izipx, const PMLevel& L, const PMBox* boxes_dev, int nbox, const void* xp, float* vp,
               const int32_t* rhoc, const int64_t* cell_off, const R* phi, double kc, PMOut* out,
               hipStream_t s, int shape = 0);
// True when the level's interpolation takes the cell-group form (w = 4 or 8,
// power-of-two scale), whose block shape is tunable.
bool pm_interp_cells_form(const PMLevel& L);

// Green_3D (Green/Green.f90:1-80) evaluated in FP64 on the device, stored as
// FP32 (n/2+1, n, n) and scaled by `factor` ((1-f_nu) or 1).
void green_device(int n, float apm_soft, float apm_range, float gridsize, double factor, float* out, hipStream_t s);

// Sub-tile classification (kick.f90:110-133): ires(4096) and napm counts.
// ires and the largest physical cell count (cmax) of every sub-tile;
// ires_fixed in 2..5 puts every sub-tile at that level instead.
void classify_subtiles(const Geo& g, int ires_fixed, const int32_t* rhoc, int32_t* ires, int32_t* napm,
                       int32_t* cmax, hipStream_t s);

// Binned deposit for sub-tile levels with w <= 8 mesh cells per coarse cell
// and a power-of-two ratio_cs/gridsize: one pass lists, for every 16^3 mesh
// block of every box, the particles whose TSC taps reach it (entries carry
// the box-local cell and the raw coordinates), so the deposit reads each
// block's contributors only instead of every coarse cell near the block.
struct DepBins {
  int32_t* count = nullptr;            // nbox * nblk^3 (capacities, then entries used)
  int64_t* start = nullptr;            // nbox * nblk^3 + 1 (scan of the capacities)
  unsigned long long* ents = nullptr;  // entries
  void* scan_tmp = nullptr;
  size_t scan_tmp_bytes = 0;
};
bool pm_deposit_binned_applies(const PMLevel& L);
int64_t pm_dep_blocks(const PMLevel& L);  // blocks per box
// List capacities (the particles of the coarse cells that can reach each
// block) and their scan; returns the total capacity (host sync).
int64_t pm_dep_bin_count(int izipx, const PMLevel& L, const PMBox* boxes_dev, int nbox, const void* xp,
                         const int32_t* rhoc, const int64_t* cell_off, DepBins& S, hipStream_t s);
void pm_dep_bin_fill(int izipx, const PMLevel& L, const PMBox* boxes_dev, int nbox, const void* xp,
                     const int32_t* rhoc, const int64_t* cell_off, DepBins& S, hipStream_t s);
// Order-independent checksum of `bytes` (a multiple of 4) at p; scratch8 is
// 8 bytes of device memory.  Synchronises the stream.
uint64_t device_checksum(const void* p, size_t bytes, void* scratch8, hipStream_t s);
// threads per 16^3 block: 256, 512 or 1024 (the same mesh for each)
template <typename R>
void pm_deposit_lists(int izipx, const PMLevel& L, const PMBox* boxes_dev, int nbox, const DepBins& S, R* mesh,
                      float mass, hipStream_t s, int threads = 512);

// A fixed pseudo-random pattern in [-1/2, 1/2) (the autotuner's FFT input).
void fill_pattern(float* p, int64_t n, hipStream_t s);

// a[i] += b[i] (the density of a second species added to a level's meshes).
template <typename R>
void mesh_add(R* a, const R* b, int64_t n, hipStream_t s);

// Exclusive scan of int32 counts into int64 offsets (n+1 outputs).
void scan_counts(const int32_t* in, int64_t* out, int64_t n, void* tmp, size_t tmp_bytes, hipStream_t s);
size_t scan_tmp_bytes(int64_t n);

}  // namespace cube
