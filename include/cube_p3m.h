/*
 * cube_p3m.h -- C-ABI of the MI355X-native CUBE P3M force-and-update step.
 *
 * The Fortran host (one image, nn=1) keeps its module arrays and calls these
 * entry points through ISO_C_BINDING (cube_nu_amd/fortran/cube_p3m_iface.f90).
 * Every pointer is a plain host pointer into the Fortran arrays unless the
 * entry point name ends in _resident (those act on device copies owned by the
 * context).  Array layouts are the reference's column-major layouts:
 *
 *   xp  (3, np)  integer(izipx)   -> int8_t / int16_t, xp[3*ip + d]
 *   vp  (3, np)  real(4)          -> float,           vp[3*ip + d]
 *   pid (np)     integer(8)       -> int64_t
 *   rhoc(1-ncb:nt+ncb, 1-ncb:nt+ncb, 1-ncb:nt+ncb, nnt, nnt, nnt) integer(4)
 *   idx_b_l/idx_b_r (1-ncb:nt+ncb, 1-ncb:nt+ncb, nnt, nnt, nnt) integer(8)
 *   ppl0/pplr/pprl/ppr0/ppl/ppr (nt, nt, nnt, nnt, nnt) integer(8)
 *
 * Particle indices stay 1-based in the spine arrays (variables.f90:91-92),
 * exactly as spine_image writes them (variables.f90:225-295).
 *
 * All entry points return 0 on success or one of the CUBE_ERR_* codes; the
 * Fortran shims turn a non-zero code into `error stop` with the reference's
 * message.  No exception crosses the ABI.  A context owns one HIP stream and
 * is not re-entrant.
 */
#ifndef CUBE_P3M_H
#define CUBE_P3M_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (reference `error stop` sites in brackets) ---------- */
enum cube_status {
  CUBE_OK = 0,
  CUBE_ERR_ARG = 1,            /* invalid configuration or argument          */
  CUBE_ERR_HIP = 2,            /* HIP runtime failure                        */
  CUBE_ERR_TILE_OVERFLOW = 3,  /* np_tile_max exceeded   [drift.f90:87-93]   */
  CUBE_ERR_IMAGE_OVERFLOW = 4, /* np_image_max exceeded  [buffer_grid.f90:161-171] */
  CUBE_ERR_OUT_OF_RANGE = 5,   /* particle drifts beyond 2*ncb [drift.f90:64-73] */
  CUBE_ERR_COUNT_MISMATCH = 6, /* rholocal /= rhoce      [drift.f90:139-143] */
  CUBE_ERR_PID = 7,            /* pid <= 0               [drift.f90:124-129, 188-193] */
  CUBE_ERR_SPINE = 8,          /* idx_b_r inconsistent with rhoc             */
  CUBE_ERR_NO_DEVICE = 9,      /* no HIP device visible                      */
  CUBE_ERR_STATE = 10          /* resident state not initialised             */
};

/* ---- configuration: mirrors parameters.f90:8-132 at nn=1 --------------- */
typedef struct cube_p3m_config {
  int32_t ng;          /* particles (fine cells) per dim      parameters.f90:24 */
  int32_t nnt;         /* tiles per dim                       parameters.f90:28 */
  int32_t nns;         /* sub-tiles per tile per dim          parameters.f90:29 */
  int32_t ratio_cs;    /* fine cells per coarse cell          parameters.f90:25 */
  int32_t ncb;         /* coarse buffer depth                 parameters.f90:81 */
  int32_t ngb;         /* PM2 fine buffer depth               parameters.f90:61 */
  int32_t izipx;       /* bytes per coordinate: 1 or 2        parameters.f90:9-13 */
  int32_t precision;   /* 32: FP32 form; 64: plain FP64 form                    */
  float box;           /* Mpc/h                               parameters.f90:23 */
  float apm1c;         /* PM1 softening (coarse cells)        parameters.f90:38 */
  float apm2;          /* PM2 softening (fine cells)          parameters.f90:40 */
  float apm3f;         /* PM3 softening (fine-fine cells)     parameters.f90:42 */
  float app;           /* PP softening                        parameters.f90:46 */
  int32_t pm3;         /* kick.f90:9  */
  int32_t pp;          /* kick.f90:10 */
  int32_t pp_corr;     /* kick.f90:11 */
  double f_nu;         /* neutrino fraction                   parameters.f90:164 */
  double mass_nu;      /* sum of neutrino masses (eV)         parameters.f90:104 */
  float image_buffer;  /* parameters.f90:85 */
  float tile_buffer;   /* parameters.f90:86 */
  int64_t workspace_bytes; /* device budget for batched tile meshes; 0 = auto */
  int32_t ires_fixed;  /* 0: sub-tile PM3/PP resolution from the occupancy
                          thresholds (kick.f90:116-130); 2..5: every sub-tile
                          at that ires (the commented ires(itile)=4, kick.f90:131) */
} cube_p3m_config;

/* Kick inputs: the module scalars kick.f90 reads (variables.f90:35, sim). */
typedef struct cube_kick_in {
  double a_mid;        /* FP64: the kick coefficient a_mid*dt/6/pi is formed in FP64 */
  double dt;
  float mass_p_cdm;
  int32_t phase_mask;  /* bit0 PM1, bit1 PM2, bit2 PM3, bit3 PP; 0 = all */
} cube_kick_in;

/* Kick outputs: what kick.f90 writes into `sim`/module state (kick.f90:39,
 * 69, 239, 298, 314-315, 109-133; particle_mesh.f90:106). */
typedef struct cube_kick_out {
  float dt_pm1, dt_pm2, dt_pm3, dt_pp, dt_vmax, vz_max;
  float f2max_pm1, f2max_pm2, f2max_pm3, f2max_pp;
  float vmax[3];       /* vmax after the last phase run */
  int32_t napm[7];     /* sub-tiles per resolution (kick.f90:109-133) */
  double testrho;      /* FP64 sum of the PM1 density (particle_mesh.f90:106) */
} cube_kick_out;

typedef struct cube_drift_in {
  double dt_old;       /* variables.f90:35 (real(4) there; FP64 here: dt_mid and  */
  double dt;           /* the drift displacement dt_mid*v are formed in FP64)     */
  float vz_max;        /* sim%vz_max from the previous kick (kick.f90:315) */
} cube_drift_in;

typedef struct cube_drift_out {
  int64_t nplocal;     /* sim%nplocal (drift.f90:182) */
  float overhead_tile; /* drift.f90:85 */
  int32_t nlayer;      /* colour count used for the within-cell order (drift.f90:26) */
} cube_drift_out;

/* Seeded Zel'dovich input maker (follows utilities/ic.f90:231-796 with a
 * counter-based RNG and global binning). */
typedef struct cube_ic_params {
  uint64_t seed;
  double z;            /* redshift of the displacement amplitude */
  double h, omega_m, omega_b, n_s, sigma8;
  int32_t nic;         /* refinement of the potential grid (ic.f90 nic=2) */
  float box;
} cube_ic_params;

typedef struct cube_p3m_ctx cube_p3m_ctx;

/* Context: replaces the plan creation at initialize.f90:336-348 and the Green
 * function set-up at initialize.f90:207-334 (tf* = Gk* or (1-f_nu)*Gk*). */
int cube_p3m_create(const cube_p3m_config* cfg, cube_p3m_ctx** ctx);
/* Replaces finalize.f90:11-12 (plan destruction). */
int cube_p3m_destroy(cube_p3m_ctx* ctx);
const char* cube_p3m_last_error(const cube_p3m_ctx* ctx);

/* Overwrite one transfer array: level 1 = tf1, 2 = tf2, 3..6 = tf3_2,
 * tf3_4, tf3_6, tf3_8 (variables.f90:21-22), as get_tf_cb2matter does every
 * step (power_nu.f90:424-431).  n = number of floats (n/2+1)*n*n. */
int cube_p3m_set_tf(cube_p3m_ctx* ctx, int32_t level, const float* tf, int64_t n);
int cube_p3m_get_tf(cube_p3m_ctx* ctx, int32_t level, float* tf, int64_t n);

/* ---- swap-in entry points (host arrays, one call = one reference call) -- */

/* particle_mesh(rho1) for PM1 (particle_mesh.f90:1-222): writes the coarse
 * density rho1(nc,nc,nc) (may be NULL) and kicks vp of physical particles. */
int cube_p3m_particle_mesh(cube_p3m_ctx* ctx, const void* xp, float* vp,
                           const int32_t* rhoc, const int64_t* idx_b_r,
                           int64_t np, const cube_kick_in* in,
                           float* density, cube_kick_out* out);

/* kick (kick.f90:1-353): PM1 + PM2 + PM3 + PP velocity update. */
int cube_p3m_kick(cube_p3m_ctx* ctx, const void* xp, float* vp,
                  const int32_t* rhoc, const int64_t* idx_b_r, int64_t np,
                  const cube_kick_in* in, cube_kick_out* out);

/* drift (drift.f90:1-233): moves, re-sorts and re-encodes.  On return xp/vp/
 * pid hold the nplocal physical particles in tile order, rhoc holds physical
 * counts with zero buffers, and the spine arrays (each may be NULL) are
 * rebuilt as spine_image does.  np = length of the input arrays. */
int cube_p3m_drift(cube_p3m_ctx* ctx, void* xp, float* vp, int64_t* pid,
                   int32_t* rhoc, int64_t np,
                   int64_t* idx_b_l, int64_t* idx_b_r,
                   int64_t* ppl0, int64_t* pplr, int64_t* pprl, int64_t* ppr0,
                   int64_t* ppl, int64_t* ppr,
                   const cube_drift_in* in, cube_drift_out* out);

/* buffer_grid + buffer_x + buffer_v at nn=1 (buffer_grid.f90:1-226,
 * buffer_particle.f90:1-375): fills the ncb-deep periodic buffer of rhoc and
 * re-lays xp/vp/pid with the buffer cells' copies.  Input: nplocal physical
 * particles in tile order; output arrays have capacity np_cap. */
int cube_p3m_buffer(cube_p3m_ctx* ctx, void* xp, float* vp, int64_t* pid,
                    int32_t* rhoc, int64_t nplocal, int64_t np_cap,
                    int64_t* idx_b_l, int64_t* idx_b_r, int64_t* np_ext);

/* buffer_v at nn=1 (buffer_particle.f90:11-27; called after kick,
 * main.f90:38): in the extended layout of rhoc (np = length of the arrays),
 * the velocities (and the IDs, when pid is not NULL) of every buffer cell are
 * overwritten in place by those of the physical cell the buffer cell is the
 * periodic image of, so that the next drift moves every copy of a particle
 * alike.  Leaves no resident state behind. */
int cube_p3m_buffer_v(cube_p3m_ctx* ctx, float* vp, int64_t* pid,
                      const int32_t* rhoc, int64_t np);

/* ---- resident mode: state kept on the device between calls ------------- */

/* Upload the extended (buffered) layout. */
int cube_p3m_state_upload(cube_p3m_ctx* ctx, const void* xp, const float* vp,
                          const int64_t* pid, const int32_t* rhoc, int64_t np);
/* Download the current layout (extended after buffer/IC, physical after
 * drift).  *np receives the particle count of that layout. */
int cube_p3m_state_download(cube_p3m_ctx* ctx, void* xp, float* vp,
                            int64_t* pid, int32_t* rhoc, int64_t np_cap,
                            int64_t* np);
int cube_p3m_kick_resident(cube_p3m_ctx* ctx, const cube_kick_in* in,
                           cube_kick_out* out);
int cube_p3m_drift_resident(cube_p3m_ctx* ctx, const cube_drift_in* in,
                            cube_drift_out* out);
int cube_p3m_buffer_resident(cube_p3m_ctx* ctx, int64_t* np_ext);
/* buffer_v on the resident extended layout (velocities only: the buffer IDs
 * written by the buffer step are already the physical cells' IDs). */
int cube_p3m_buffer_v_resident(cube_p3m_ctx* ctx);
/* One full step in the main.f90:33-38 order after timestep:
 * drift -> buffer_grid/buffer_x -> kick -> buffer_v. */
int cube_p3m_step_resident(cube_p3m_ctx* ctx, const cube_drift_in* din,
                           const cube_kick_in* kin, cube_drift_out* dout,
                           cube_kick_out* kout);
/* Generate seeded Zel'dovich particles straight into the resident state
 * (extended layout, buffers filled).  Returns the velocity bound vz_max the
 * first drift needs (ic.f90:584-586). */
int cube_p3m_make_ic(cube_p3m_ctx* ctx, const cube_ic_params* ic,
                     float* vz_max, int64_t* nplocal);

/* ---- analysis on the resident state ------------------------------------- */

/* CIC power spectrum of the physical particles on an nw^3 periodic mesh:
 * utilities/cicpower.f90:67-245 (deposit, delta = rho/mean - 1) with
 * utilities/powerspectrum.f90:140-259 (auto_power, CIC window, shot noise,
 * pk_correction).  xi is the reference's xi(10, 0:nbin), column-major,
 * nbin = nint(nw/2*sqrt(3.)) (powerspectrum.f90:14); columns 1-5 are filled:
 * count, k (h/Mpc), raw power, shot-noise subtracted, alias corrected
 * ((Mpc/h)^3, box from the configuration).  delta (nw^3, x fastest) and mean
 * (FP64 mean CIC mass per cell) may be NULL. */
int cube_p3m_cicpower(cube_p3m_ctx* ctx, int32_t nw, float* xi, int32_t nbin,
                      float* delta, double* mean);

/* Wait until the context's work has finished (the kick's second stream is
 * joined into the context stream, so this waits for it too). */
int cube_p3m_synchronize(cube_p3m_ctx* ctx);

/* ---- neutrino particles ------------------------------------------------- */
/* Kick with a second, PM-only species (neutrino particles) resident in
 * another context of the same geometry, coordinates and precision (the
 * buffered layout, after its own buffer step).  Both species deposit into
 * the PM1 and PM2 meshes -- this context's particles with in->mass_p_cdm,
 * the other's with nu_mass -- and both are kicked by PM1 and PM2; PM3 and PP
 * act on this context's particles only.  *out_nu receives the second
 * species' PM1/PM2 maxima, Delta-t criteria, vz_max (for its drift) and PM1
 * density sum (testrho).  No reference counterpart: the reference represents
 * massive neutrinos by scaling the CDM transfer functions with (1 - f_nu)
 * (initialize.f90:316-334, neutrinos/power_nu.f90:346-498); with neutrino
 * particles the contexts keep mass_nu = 0 (unscaled tf) and the masses carry
 * f_nu. */
int cube_p3m_kick_resident_nu(cube_p3m_ctx* ctx, cube_p3m_ctx* nu,
                              const cube_kick_in* in, float nu_mass,
                              cube_kick_out* out, cube_kick_out* out_nu);
/* Add Fermi-Dirac thermal velocities to the resident particles (neutrino
 * particles; no reference counterpart): each particle, keyed by (seed,
 * particle ID) so that buffer copies agree, receives u * v_scale (code
 * velocity units) in a uniformly random direction, u = pc/(k T_nu) drawn from
 * u^2/(e^u + 1).  Use a seed other than the input maker's.  *vz_max receives
 * the largest |v_z| afterwards (the next drift's bound).  Needs IDs. */
int cube_p3m_add_thermal_velocities(cube_p3m_ctx* ctx, double v_scale,
                                    uint64_t seed, float* vz_max);
/* ---- graph capture and measurement -------------------------------------- */
/* Graph mode for full resident kicks (and the kick + buffer_v of
 * cube_p3m_step_resident): the kick's compute passes -- PM1 + PM2, then
 * PM3 + PP (+ buffer_v), whose lists are built on a second stream meanwhile
 * -- are captured into two HIP graphs and replayed while their launch
 * signatures (box counts per level, tuned shapes, buffer addresses) are
 * unchanged; the kick coefficient is read from device memory.  Results are
 * bit-identical to graph mode off.  Kicks with a level that needs more than
 * one workspace batch run uncaptured.  The timing report counts captures and
 * graph-run kicks (graph_captures, graph_launches).  No reference
 * counterpart (the reference's step loop is main.f90:30-59). */
int cube_p3m_graph_enable(cube_p3m_ctx* ctx, int32_t enable);
/* Enable HIP-event timing of every kernel launch on the context stream. */
int cube_p3m_timing_enable(cube_p3m_ctx* ctx, int32_t enable);
int cube_p3m_timing_reset(cube_p3m_ctx* ctx);
/* Text report, one line per kernel: "name calls total_ms bytes flops". */
int cube_p3m_timing_report(cube_p3m_ctx* ctx, char* buf, int64_t len);
/* Library build identification string. */
const char* cube_p3m_version(void);

#ifdef __cplusplus
}
#endif
#endif /* CUBE_P3M_H */
