// ic.h -- seeded Zel'dovich particle generator (the algorithm of
// utilities/ic.f90:231-796, with a counter-based RNG and global binning).
#pragma once
#include "../../include/cube_p3m.h"
#include "common.h"

namespace cube {

struct ICResult {
  int64_t nplocal;
  float vz_max;
};

// Writes the physical layout (tile order, cells (k,j,i), ascending pid inside
// a cell) into xp/vp/pid (capacity >= ng^3) and rhoc (extended layout, zero
// buffers).
ICResult make_ic(const Geo& g, const cube_ic_params& p, void* xp, float* vp, int64_t* pid, int32_t* rhoc,
                 hipStream_t s);

// Adds Fermi-Dirac thermal velocities to n particles: u * v_scale in a
// uniformly random direction, u = pc/(kT) drawn from u^2/(e^u + 1), both keyed
// by (seed, pid) so that a particle's buffer copies draw the same velocity.
// Returns the largest |v_z| after the addition.
float add_thermal_velocities(float* vp, const int64_t* pid, int64_t n, double v_scale, uint64_t seed,
                             hipStream_t s);

// Host helpers exposed for tests of the cosmology inputs.
double bbks_sigma8_norm(const cube_ic_params& p);
double dgrow(double a, double om, double ol);

}  // namespace cube
