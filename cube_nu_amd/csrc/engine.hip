// engine.hip -- context, level pipelines and the C-ABI of cube_p3m.h.
//
// One context = one image (nn=1) on the current HIP device, one stream.
// The kick runs PM1 (coarse periodic mesh), PM2 (64 tile meshes, batched),
// PM3 (sub-tile meshes grouped by resolution, batched) and PP, in the order
// of kick.f90:32-311; the drift and the nn=1 buffer layout complete a step.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>

#include "../../include/cube_p3m.h"
#include "cic.h"
#include "common.h"
#include "drift.h"
#include "fft.h"
#include "ic.h"
#include "pm.h"
#include "pp.h"
#include "timing.h"

using namespace cube;

namespace {

// Device buffer that only grows.  A growth allocates 1/16 more than asked:
// the particle sets, the PP entries and the deposit lists creep up by a few
// elements from step to step, and an exact-size regrowth would free and
// reallocate gigabytes (about a millisecond of an idle GPU) each time.
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  T* ensure(size_t m) {
    if (m > n) {
      release();
      const size_t cap = m + m / 16;
      hipError_t e = hipMalloc(&p, cap * sizeof(T));
      if (e != hipSuccess) {
        p = nullptr;
        n = 0;
        throw Error(CUBE_ERR_HIP, "hipMalloc of " + std::to_string(cap * sizeof(T)) + " bytes failed: " +
                                      hipGetErrorString(e));
      }
      n = cap;
    }
    return p;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { release(); }
};

// G_grid = 1.0/6.0/pi (variables.f90:28) as stored in real(4).
inline float g_grid() { return (float)((double)(1.0f / 6.0f) / kPi); }
// dt = 0.5*sqrt(1./(sqrt(f2max)*a_mid*G_grid)) (kick.f90:39, 69, 239), or
// with 0.05 for PP (kick.f90:298), evaluated in real(4).
inline float dt_from_f2(float c, float f2, float a_mid) {
  return c * std::sqrt(1.0f / ((std::sqrt(f2) * a_mid) * g_grid()));
}
// Kick coefficient a_mid*dt/6/pi (particle_mesh.f90:205, kick.f90:574, 797)
// in FP64: the reference rounds a_mid*dt to real(4); here the kick and drift
// coefficients stay FP64 on the GPU.
inline double kick_coef(const cube_kick_in& in) { return ((in.a_mid * in.dt) / 6.0) / kPi; }

inline uint64_t fbits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

__global__ void k_set_f64(double* p, double v) { *p = v; }

// Candidates of each tuned choice, in order of preference.
const std::vector<int> kTunePPDepth = {6, 4, 8, kPPBrick8 + 6, kPPBrick8 + 4, kPPBrick8 + 8};
const std::vector<int> kTuneDepListThreads = {512, 256, 1024};
const std::vector<int> kTuneFFTShape = {0, 1, 2};
const std::vector<int> kTuneInterpShape = {0, 1, 2};

// Autotuner picks persisted across runs: when the environment names a file in
// CUBE_AUTOTUNE_FILE, a context reads the picks recorded there (lines
// "key pick") and appends every pick it makes, so later runs skip the trials.
// A line whose key names no tuned choice, or whose pick is not one of that
// choice's candidates (a file from another build, or edited), is ignored.
std::mutex g_tune_file_mu;

bool tune_pick_valid(const std::string& key, int pick) {
  static const std::pair<const char*, const std::vector<int>*> choices[] = {
      {"pp_depth_l", &kTunePPDepth},
      {"dep_list_threads_l", &kTuneDepListThreads},
      {"fft_shape_l", &kTuneFFTShape},
      {"interp_shape_l", &kTuneInterpShape}};
  for (const auto& c : choices)
    if (key.compare(0, std::strlen(c.first), c.first) == 0)
      return std::find(c.second->begin(), c.second->end(), pick) != c.second->end();
  return false;
}

void tune_file_read(std::map<std::string, int>& picks) {
  const char* f = std::getenv("CUBE_AUTOTUNE_FILE");
  if (!f) return;
  std::lock_guard<std::mutex> lk(g_tune_file_mu);
  if (FILE* fp = std::fopen(f, "r")) {
    char k[256];
    int v;
    while (std::fscanf(fp, "%255s %d", k, &v) == 2)
      if (tune_pick_valid(k, v)) picks[k] = v;
    std::fclose(fp);
  }
}

void tune_file_append(const std::string& key, int pick) {
  const char* f = std::getenv("CUBE_AUTOTUNE_FILE");
  if (!f) return;
  std::lock_guard<std::mutex> lk(g_tune_file_mu);
  if (FILE* fp = std::fopen(f, "a")) {
    std::fprintf(fp, "%s %d\n", key.c_str(), pick);
    std::fclose(fp);
  }
}

}  // namespace

struct cube_p3m_ctx {
  cube_p3m_config cfg{};
  Geo g{};
  int prec = 32;
  int device = 0;
  hipStream_t s = nullptr;
  std::string err;
  KTimer tm;

  // transfer functions: [1]=tf1 (nc), [2]=tf2 (ngt), [3..6]=tf3_2..tf3_8
  DevBuf<float> tf[7];
  int tfn[7] = {0, 0, 0, 0, 0, 0, 0};
  FFTCube<float> fcf[7];
  FFTCube<double> fcd[7];

  // resident particle state (double-buffered)
  DevBuf<char> xp[2];
  DevBuf<float> vp[2];
  DevBuf<int64_t> pid[2];
  int cur = 0;

  int layout = 0;  // 0 none, 1 extended (buffered), 2 physical (after drift)
  // max |v| per axis after the last full resident kick (kick.f90:299-313): a
  // bound of every resident velocity until the next upload
  float last_vmax[3] = {0, 0, 0};
  bool vmax_valid = false;
  int64_t np = 0;
  bool has_pid = false;
  DevBuf<int32_t> rhoc[2];
  int rcur = 0;
  DevBuf<int64_t> off;
  bool off_valid = false;

  // scratch
  DevBuf<char> scan_tmp;
  DevBuf<char> work;
  DevBuf<char> stencil;  // coarse-level deposit: 27 tap sums per coarse cell
  // per-slot box lists and binned lists (slot 0 PM1, 1 PM2, 2..5 PM3 at
  // iapm 2..5; PP by iapm 2..5): a captured kick reads each level's lists
  // from where its sizing pass left them
  DevBuf<PMBox> lboxes[6];
  DevBuf<int32_t> ldepcount[6];
  DevBuf<int64_t> ldepstart[6];
  DevBuf<unsigned long long> ldepents[6];
  DevBuf<PPBox> lppboxes[6];
  DevBuf<int32_t> lppcount[6];
  DevBuf<int64_t> lppstart[6];
  DevBuf<float4> lppent[6];
  DevBuf<PMOut> pmout;
  DevBuf<PPOut> ppout;
  DevBuf<int32_t> ires, napm, cmax;
  DevBuf<unsigned> ppclaim;
  DevBuf<int32_t> dflags, drank, dtable, dnewcount;
  DevBuf<int64_t> dbase;
  DevBuf<int32_t> physcount;
  DevBuf<int64_t> physoff;
  int64_t last_pairs = 0;
  // autotuner: candidate times (ms) of every choice this context tuned, and
  // the scratch the PP trials run on
  struct TuneRecord {
    std::vector<int> cand;
    std::vector<float> ms;
    int pick = 0;
  };
  std::map<std::string, TuneRecord> tuned;
  std::map<std::string, int> tune_picks;  // picks of this context (and of CUBE_AUTOTUNE_FILE)
  DevBuf<float> vtune;
  DevBuf<PPOut> otune;
  DevBuf<unsigned long long> tune_sum;
  template <class Run>
  int autotune(const std::string& key, const std::vector<int>& cand, Run&& run, const void* result,
               size_t result_bytes);

  // Kick graph.  With graph_on, a full resident kick runs in two passes: the
  // sizing pass (classification, list counts and fills, box uploads, the
  // autotuner) runs on the stream as usual and reads its sizes back; the
  // compute pass (deposits, transforms, interpolation and kick, PP force,
  // buffer_v) is captured into a graph once and replayed while the launch
  // signature -- every level's box count and every buffer address the
  // compute pass reads -- stays the same.  The kick coefficient, the one
  // argument that changes every step, is read from device memory (kcoef).
  bool graph_on = false;
  // Overlapped kick (full resident kicks after a context's first, graph mode
  // off): the PM3 and PP sizing runs on a second stream while PM1 and PM2
  // compute on the first; the first stream waits for the sizing before PM3.
  bool overlap_on = true;
  hipStream_t s2 = nullptr;
  hipEvent_t ev_sized = nullptr;
  int64_t kicks_done = 0;
  bool classified = false;  // kick_pass: the classes are already on the host
  bool skip_zero = false;   // kick_pass: the reductions are already cleared
  bool no_trials = false;   // autotune: keep the first candidate, untimed
  bool trial_skipped = false;  // set when no_trials kept a candidate untried
  int kph = 0;           // 0 both passes in one, 1 sizing only, 2 compute only
  bool graph_ok = true;  // cleared by a sizing pass that needs more than one batch per level
  std::vector<uint64_t> gsig_acc, gsig[2];
  hipGraphExec_t gexec[2] = {nullptr, nullptr};  // PM1+PM2, then PM3+PP(+buffer_v)
  int64_t graph_captures = 0, graph_launches = 0;
  DevBuf<double> kcoef;
  int64_t wsb = -1;  // workspace budget held for the duration of a kick
  // A PM-only second species (neutrino particles) coupled for one kick: its
  // particles deposit into the PM1 and PM2 meshes with mass nu_mass and take
  // the PM1 and PM2 kicks; PM3 and PP act on this context's particles only.
  cube_p3m_ctx* nu = nullptr;
  float nu_mass = 0;
  DevBuf<char> work2;  // the second species' level meshes before they are added
  DevBuf<char> ftune;  // the FFT shape trials' meshes (released after the trials)
  void sig_add(std::initializer_list<uint64_t> v) { gsig_acc.insert(gsig_acc.end(), v.begin(), v.end()); }
  void graph_drop() {
    for (int i = 0; i < 2; ++i) {
      if (gexec[i]) (void)hipGraphExecDestroy(gexec[i]);
      gexec[i] = nullptr;
      gsig[i].clear();
    }
  }

  ~cube_p3m_ctx() {
    graph_drop();
    if (ev_sized) (void)hipEventDestroy(ev_sized);
    if (s2) (void)hipStreamDestroy(s2);
    for (int i = 0; i < 7; ++i) {
      fcf[i].release();
      fcd[i].release();
    }
    if (s) (void)hipStreamDestroy(s);
  }

  int esz() const { return (int)g.izipx; }
  void* xpc() { return xp[cur].p; }
  float* vpc() { return vp[cur].p; }
  int64_t* pidc() { return has_pid ? pid[cur].p : nullptr; }
  int32_t* rhocc() { return rhoc[rcur].p; }

  void* scan_scratch(int64_t n) {
    size_t b = scan_tmp_bytes(n);
    return scan_tmp.ensure(b);
  }
  void scan(const int32_t* in, int64_t* out, int64_t n) {
    void* t = scan_scratch(n);
    scan_counts(in, out, n, t, scan_tmp.n, s);
  }
  // Grow one particle buffer set; growing discards its contents, so callers
  // only grow the set they are about to overwrite.
  int64_t capb[2] = {0, 0};
  void ensure_set(int b, int64_t c) {
    if (c > capb[b]) {
      const int64_t cc = c + c / 16;  // room for the next steps' growth
      xp[b].ensure((size_t)cc * 3 * esz());
      vp[b].ensure((size_t)cc * 3);
      pid[b].ensure((size_t)cc);
      capb[b] = cc;
    }
    rhoc[0].ensure(g.ncells_ext);
    rhoc[1].ensure(g.ncells_ext);
    off.ensure(g.ncells_ext + 1);
  }
  void ensure_offsets() {
    if (off_valid) return;
    KScope k(&tm, "spine_scan", (double)g.ncells_ext * 12);
    scan(rhocc(), off.p, g.ncells_ext);
    off_valid = true;
  }
  int64_t ws_budget() const {
    if (cfg.workspace_bytes > 0) return cfg.workspace_bytes;
    if (wsb >= 0) return wsb;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return (int64_t)8 << 30;
    return std::max<int64_t>((int64_t)256 << 20, std::min<int64_t>((int64_t)24 << 30, (int64_t)(fr * 0.35)));
  }

  void init();
  void set_tf_default();
  template <typename R>
  void kick(const cube_kick_in& in, cube_kick_out& out, float* density_host, bool with_buffer_v = false);
  template <typename R>
  void kick_pass(const cube_kick_in& in, cube_kick_out& out, float* density_host, int mask, bool with_buffer_v,
                 std::vector<int32_t>& ires_h, std::vector<int32_t>& cmax_h);
  template <typename R>
  void kick_overlapped(const cube_kick_in& in, cube_kick_out& out, bool with_buffer_v, std::vector<int32_t>& ires_h,
                       std::vector<int32_t>& cmax_h);
  template <typename R>
  bool classify_and_reserve(cube_kick_out& out, std::vector<int32_t>& ires_h, std::vector<int32_t>& cmax_h);
  template <typename R>
  void kick_graph(const cube_kick_in& in, cube_kick_out& out, bool with_buffer_v, std::vector<int32_t>& ires_h,
                  std::vector<int32_t>& cmax_h);
  template <typename R>
  void run_level(int level, int slot, const std::vector<PMBox>& bx, const float* G, const FFTCube<R>& fc,
                 const PMLevel& L, const cube_kick_in& in, PMOut* o, const char* tag, float* density_host = nullptr);
  void pp_phase(const std::vector<int32_t>& ires_h, const std::vector<int32_t>& cmax_h, const cube_kick_in& in,
                PPOut* o);
  int tune_pp(int iapm, const PPLevel& L, const PPBox* boxes, int nb, PPScratch& S);
  template <typename R>
  int interp_shape(int level, int n, const PMLevel& L, const PMBox* boxes, int nb, const cube_kick_in& in, const R* phi,
                   bool trials);
  DevBuf<PMOut> pmtune;
  std::string tune_key(const char* what, int level, int64_t size) const {
    return std::string(what) + "_l" + std::to_string(level) + "_n" + std::to_string(size) + "_fp" +
           std::to_string(prec);
  }
  void drift(const cube_drift_in& in, cube_drift_out& out);
  void buffer(int64_t* np_ext);
  void buffer_vel();
};

// ---------------------------------------------------------------------------
void cube_p3m_ctx::init() {
  const cube_p3m_config& c = cfg;
  if (c.ng <= 0 || c.nnt <= 0 || c.nns <= 0 || c.ratio_cs <= 0 || c.ncb <= 0 || c.ngb < 0)
    throw Error(CUBE_ERR_ARG, "non-positive size in configuration");
  if (c.izipx != 1 && c.izipx != 2) throw Error(CUBE_ERR_ARG, "izipx must be 1 or 2");
  if (c.ires_fixed != 0 && (c.ires_fixed < 2 || c.ires_fixed > 5))
    throw Error(CUBE_ERR_ARG, "ires_fixed must be 0 (adaptive) or 2..5");
  if (c.precision != 32 && c.precision != 64) throw Error(CUBE_ERR_ARG, "precision must be 32 or 64");
  if (c.ng % (c.ratio_cs * c.nnt * c.nns) != 0) throw Error(CUBE_ERR_ARG, "ng must be a multiple of ratio_cs*nnt*nns");
  prec = c.precision;
  g.init(c.ng, c.nnt, c.nns, c.ratio_cs, c.ncb, c.ngb, c.izipx, c.image_buffer, c.tile_buffer);
  if (g.ngb < 4) throw Error(CUBE_ERR_ARG, "ngb must be >= 4 (gradient stencil)");
  // the drift divides by ratio_cs as a product with its reciprocal and the PP
  // binning takes xvec = xf / ratio_cs: both exact only for a power of two
  if (c.ratio_cs < 1 || (c.ratio_cs & (c.ratio_cs - 1)) != 0)
    throw Error(CUBE_ERR_ARG, "ratio_cs must be a power of two");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw Error(CUBE_ERR_NO_DEVICE, "no HIP device visible");
  CUBE_HIP(hipGetDevice(&device));
  CUBE_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  CUBE_HIP(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
  CUBE_HIP(hipEventCreateWithFlags(&ev_sized, hipEventDisableTiming));
  if (const char* e = std::getenv("CUBE_KICK_OVERLAP")) overlap_on = !(e[0] == '0');
  tm.stream = s;
  const int sizes[7] = {0, (int)g.nc, (int)g.ngt, (int)g.nft[2], (int)g.nft[3], (int)g.nft[4], (int)g.nft[5]};
  for (int l = 1; l <= 6; ++l) {
    tfn[l] = sizes[l];
    if (prec == 32) fcf[l].init(sizes[l]);
    else fcd[l].init(sizes[l]);
    tf[l].ensure((size_t)(sizes[l] / 2 + 1) * sizes[l] * sizes[l]);
  }
  tune_file_read(tune_picks);
  pmout.ensure(4);
  ppout.ensure(1);
  napm.ensure(8);
  ires.ensure(g.ns3);
  cmax.ensure(g.ns3);
  dflags.ensure(8);
  set_tf_default();
}

// tf* = Gk* or (1-f_nu)*Gk* (initialize.f90:248-334).
void cube_p3m_ctx::set_tf_default() {
  const double fac = cfg.mass_nu > 0 ? (1.0 - cfg.f_nu) : 1.0;
  KScope k(&tm, "green", 0);
  green_device(tfn[1], cfg.apm1c, 0.0f, (float)cfg.ratio_cs, fac, tf[1].p, s);  // Green.f90 call at initialize.f90:248
  green_device(tfn[2], cfg.apm2, cfg.apm1c * (float)cfg.ratio_cs, 1.0f, fac, tf[2].p, s);  // :254
  for (int iapm = 2; iapm <= 5; ++iapm) {
    float sf = (float)kRatioSF[iapm];
    green_device(tfn[iapm + 1], cfg.apm3f, cfg.apm2 * sf, 1.0f / sf, fac, tf[iapm + 1].p, s);  // :261-280
  }
  CUBE_HIP(hipStreamSynchronize(s));
}

template <typename R>
void cube_p3m_ctx::run_level(int level, int slot, const std::vector<PMBox>& bx_in, const float* G,
                             const FFTCube<R>& fc, const PMLevel& L0, const cube_kick_in& in, PMOut* o,
                             const char* tag, float* density_host) {
  using C = typename Cplx<R>::T;
  if (bx_in.empty()) return;
  const bool SZ = kph != 2, CP = kph != 1;  // sizing and compute parts of this call
  const int64_t n = L0.n;
  const int64_t nreal = n * n * n, ncplx = n * n * (n / 2 + 1);
  const int64_t per = nreal * (int64_t)sizeof(R) + ncplx * (int64_t)sizeof(C);
  std::string t(tag);
  // The coarse level is one periodic mesh shared by all tiles; tile and
  // sub-tile levels have one mesh per box, batched to fit the workspace.
  const int64_t nmesh_total = (level == 1) ? 1 : (int64_t)bx_in.size();
  const int64_t batch = (level == 1) ? 1 : std::max<int64_t>(1, std::min<int64_t>(nmesh_total, ws_budget() / per));
  work.ensure((size_t)(batch * per));
  R* mesh = (R*)work.p;
  C* cplx = (C*)(work.p + batch * nreal * sizeof(R));
  const size_t bchunk = (level == 1) ? bx_in.size() : (size_t)batch;
  if (kph == 1 && bchunk < bx_in.size()) graph_ok = false;  // a captured kick holds one batch per level
  DevBuf<PMBox>& boxes = lboxes[slot];
  boxes.ensure(std::max<size_t>(boxes.n, bchunk));
  for (size_t b0 = 0; b0 < bx_in.size(); b0 += bchunk) {
    const int nb = (int)std::min<size_t>(bchunk, bx_in.size() - b0);
    PMLevel L = L0;
    L.bstride = nreal;
    pm_level_scale(L);
    if (kph == 2) L.kdev = kcoef.p;
    const int nmesh = (level == 1) ? 1 : nb;
    const double cells = (double)nreal * nmesh;
    if (SZ) {
      std::vector<PMBox> bx(bx_in.begin() + b0, bx_in.begin() + b0 + nb);
      if (level != 1)
        for (int i = 0; i < nb; ++i) bx[i].moff = (int64_t)i * nreal;
      CUBE_HIP(hipMemcpyAsync(boxes.p, bx.data(), sizeof(PMBox) * nb, hipMemcpyHostToDevice, s));
      sig_add({(uint64_t)level, (uint64_t)slot, (uint64_t)nb, (uint64_t)n, (uint64_t)(uintptr_t)boxes.p,
               (uint64_t)(uintptr_t)work.p});
    }
    {
      KScope k(&tm, (t + "deposit").c_str(), cells * sizeof(R));
      if (level == 1) {
        void* st = stencil.ensure(pm1_stencil_bytes(L));
        if (SZ) sig_add({(uint64_t)(uintptr_t)st});
        if (CP)
          pm1_deposit<R>((int)g.izipx, L, xpc(), rhocc(), off.p, st, mesh, in.mass_p_cdm, o, s);
      } else if (level >= 3 && pm_deposit_binned_applies(L)) {
        const int64_t nblk = nb * pm_dep_blocks(L);
        DepBins S;
        S.count = ldepcount[slot].ensure((size_t)nblk);
        S.start = ldepstart[slot].ensure((size_t)nblk + 1);
        S.scan_tmp = scan_scratch(nblk);
        S.scan_tmp_bytes = scan_tmp.n;
        if (SZ) {
          const int64_t total = pm_dep_bin_count((int)g.izipx, L, boxes.p, nb, xpc(), rhocc(), off.p, S, s);
          ldepents[slot].ensure((size_t)std::max<int64_t>(total, 1));
        }
        S.ents = ldepents[slot].p;
        if (SZ) pm_dep_bin_fill((int)g.izipx, L, boxes.p, nb, xpc(), rhocc(), off.p, S, s);
        // threads per 16^3 block: the FP32 deposit writes every cell of its
        // blocks from integer sums, so the trials leave the mesh as one run
        // would; the FP64 form sums FP64 values in LDS atomics, whose order
        // (and last bits) follows the scheduling, and is not tuned
        const int thr = std::is_same<R, float>::value
                            ? autotune(tune_key("dep_list_threads", level, n), kTuneDepListThreads,
                                       [&](int t) {
                                         pm_deposit_lists<R>((int)g.izipx, L, boxes.p, nb, S, mesh, in.mass_p_cdm,
                                                             s, t);
                                       },
                                       mesh, sizeof(R) * (size_t)(nreal * nb))
                            : 512;
        if (SZ) sig_add({(uint64_t)(uintptr_t)S.count, (uint64_t)(uintptr_t)S.start, (uint64_t)(uintptr_t)S.ents,
                         (uint64_t)thr});
        if (CP) pm_deposit_lists<R>((int)g.izipx, L, boxes.p, nb, S, mesh, in.mass_p_cdm, s, thr);
      } else if (CP)
        pm_deposit<R>((int)g.izipx, L, boxes.p, nb, xpc(), rhocc(), off.p, mesh, in.mass_p_cdm, s);
    }
    // launch shape of the tile / sub-tile FFT passes, tuned on a fixed
    // pattern in scratch meshes (every shape gives the same result)
    int fshape = 0;
    if (std::is_same<R, float>::value && level >= 2 && fft_plane_tunable((int)n)) {
      const std::string key = tune_key("fft_shape", level, n);
      const int64_t nt = std::min<int64_t>(nmesh, 256);
      const bool trials = !tune_picks.count(key);
      if (trials) ftune.ensure(sizeof(R) * (size_t)(nt * nreal));
      const R n3t = (R)((float)n * (float)n * (float)n);
      fshape = autotune(
          key, kTuneFFTShape,
          [&](int sh) {
            fill_pattern((float*)ftune.p, nt * nreal, s);
            fft_convolve<R>(fc, (R*)ftune.p, cplx, G, nt, n3t, (R)0, s, nullptr, "", sh);
          },
          ftune.p, sizeof(R) * (size_t)(nt * nreal));
      if (trials) ftune.release();
      if (SZ) sig_add({(uint64_t)fshape});
    }
    // the interpolation's block shape so far (its trials need the potential)
    if (SZ) sig_add({(uint64_t)interp_shape<R>(level, (int)n, L, boxes.p, nb, in, (const R*)nullptr, false)});
    if (!CP) continue;
    if (nu && level <= 2) {
      KScope k(&tm, (t + "deposit_nu").c_str(), cells * sizeof(R));
      R* mesh2 = (R*)work2.ensure(sizeof(R) * (size_t)(nreal * nmesh));
      if (level == 1)
        pm1_deposit<R>((int)g.izipx, L, nu->xpc(), nu->rhocc(), nu->off.p, stencil.p, mesh2, nu_mass,
                       nu->pmout.p + 0, s);
      else
        pm_deposit<R>((int)g.izipx, L, boxes.p, nb, nu->xpc(), nu->rhocc(), nu->off.p, mesh2, nu_mass, s);
      mesh_add<R>(mesh, mesh2, nreal * nmesh, s);
    }
    if (density_host) {
      std::vector<R> tmp(nreal);
      CUBE_HIP(hipMemcpyAsync(tmp.data(), mesh, sizeof(R) * nreal, hipMemcpyDeviceToHost, s));
      CUBE_HIP(hipStreamSynchronize(s));
      for (int64_t i = 0; i < nreal; ++i) density_host[i] = (float)tmp[i];
    }
    const R n3 = (R)((float)n * (float)n * (float)n);
    // PM1 (pencil_fft.f90:77) on the generic line passes: under the overlapped
    // kick a coarse-mesh plane pass (one 1024-thread, 67.5 KB workgroup per
    // z plane) waits for room on the CUs behind the second stream's list
    // building, 1.9 ms against 0.34 for the line passes
    if (level == 1) fft_convolve<R>(fc, mesh, cplx, G, 1, (R)0, (R)(float)n, s, &tm, tag, -1);
    else fft_convolve<R>(fc, mesh, cplx, G, nmesh, n3, (R)0, s, &tm, tag, fshape);        // kick.f90:846
    const int ishape = interp_shape<R>(level, (int)n, L, boxes.p, nb, in, mesh, kph == 0);
    {
      KScope k(&tm, (t + "interp_kick").c_str(), cells * sizeof(R));
      pm_interp<R>((int)g.izipx, L, boxes.p, nb, xpc(), vpc(), rhocc(), off.p, mesh, kick_coef(in), o, s, ishape);
    }
    if (nu && level <= 2) {
      KScope k(&tm, (t + "interp_kick_nu").c_str(), cells * sizeof(R));
      pm_interp<R>((int)g.izipx, L, boxes.p, nb, nu->xpc(), nu->vpc(), nu->rhocc(), nu->off.p, mesh, kick_coef(in),
                   nu->pmout.p + (level - 1), s);
    }
  }
}

void cube_p3m_ctx::pp_phase(const std::vector<int32_t>& ires_h, const std::vector<int32_t>& cmax_h,
                            const cube_kick_in& in, PPOut* o) {
  const bool SZ = kph != 2, CP = kph != 1;
  // mass_scale and v_scale as kick.f90:569-574
  float mass_scale;
  if (cfg.pp_corr && cfg.mass_nu > 0)
    mass_scale = (float)((double)in.mass_p_cdm * ((1.0 - cfg.f_nu) * (1.0 - cfg.f_nu)));
  else
    mass_scale = in.mass_p_cdm;
  const double v_scale = kick_coef(in);
  const int64_t nns = g.nns;
  // sub-tiles by descending occupancy (kick.f90:132-138 sorts them by npc_max
  // for schedule(dynamic,1)): the workgroups of the densest sub-tiles are
  // dispatched first, so the launch does not end on a few heavy ones; each
  // sub-tile's result does not depend on the order
  std::vector<int64_t> order(g.ns3);
  for (int64_t it = 0; it < g.ns3; ++it) order[it] = it;
  if (SZ)
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return cmax_h[a] > cmax_h[b]; });
  for (int iapm = 2; iapm <= 5; ++iapm) {
    std::vector<PPBox> bx;
    for (int64_t it : order)
      if (ires_h[it] == iapm) {
        PPBox b;
        b.ifs[0] = (int)(it % nns);
        b.ifs[1] = (int)((it / nns) % nns);
        b.ifs[2] = (int)((it / (nns * nns)) % nns);
        b.T = (int)(it / (nns * nns * nns));
        bx.push_back(b);
      }
    if (bx.empty()) continue;
    PPLevel L;
    L.sf = kRatioSF[iapm];
    L.npgrid = (int)g.ntt * L.sf;
    L.ntt = (int)g.ntt;
    L.nte = (int)g.nte;
    L.ncb = (int)g.ncb;
    L.rcs = (int)g.rcs;
    L.pp_range = 3.5f / (float)L.sf;  // apm3 (parameters.f90:43)
    L.pp_range2 = L.pp_range * L.pp_range;
    L.app = cfg.app;
    L.c_app = f_ra_coef<float>(L.app);
    L.c_range = f_ra_coef<float>(L.pp_range);
    L.inv_app = 1.0f / L.app;
    L.inv_range = 1.0f / L.pp_range;
    L.two_inv_range = 2.0f / L.pp_range;
    L.app2 = L.app * L.app;
    L.mass_scale = mass_scale;
    L.v_scale = v_scale;
    if (kph == 2) L.vdev = kcoef.p;
    L.inv_rcs = 1.0f / (float)L.rcs;
    {
      // r_pp in PP-cell units plus a margin of 1e-3 cell, far above the FP32
      // rounding of the coordinates, so the windows never cut a pair in range
      const double rc = (double)L.pp_range * L.sf / L.rcs + 1e-3;
      L.rcut2 = (float)(rc * rc);
    }
    const int64_t m = L.npgrid + 2, m3 = m * m * m * kSX;
    const int64_t budget = ws_budget();
    int64_t batch = std::max<int64_t>(1, std::min<int64_t>((int64_t)bx.size(), budget / 4 / (m3 * 12)));
    if (kph == 1 && batch < (int64_t)bx.size()) graph_ok = false;
    DevBuf<PPBox>& boxes = lppboxes[iapm];
    for (size_t b0 = 0; b0 < bx.size(); b0 += (size_t)batch) {
      const int nb = (int)std::min<int64_t>(batch, (int64_t)(bx.size() - b0));
      boxes.ensure(std::max<size_t>(boxes.n, (size_t)nb));
      PPScratch S;
      S.count = lppcount[iapm].ensure((size_t)(nb * m3));
      S.start = lppstart[iapm].ensure((size_t)(nb * m3 + 1));
      S.scan_tmp = scan_scratch(nb * m3);
      S.scan_tmp_bytes = scan_tmp.n;
      S.claim = ppclaim.ensure(8);  // the force kernel's work-queue counters
      int64_t total = 0;
      if (SZ) {
        CUBE_HIP(hipMemcpyAsync(boxes.p, bx.data() + b0, sizeof(PPBox) * nb, hipMemcpyHostToDevice, s));
        {
          KScope k(&tm, "pp_bin_count", 0);
          total = pp_bin_count((int)g.izipx, L, boxes.p, nb, xpc(), rhocc(), off.p, S, s);
        }
        // the force kernel queues 31-bit entry indices
        if (total > (int64_t)INT32_MAX - kPPPad)
          throw Error(CUBE_ERR_ARG,
                      "PP batch of " + std::to_string(total) + " entries exceeds 2^31; lower workspace_bytes");
        S.entries = lppent[iapm].ensure((size_t)(total + kPPPad));
        S.capacity = total;
        {
          KScope k(&tm, "pp_bin_fill", (double)total * 16);
          pp_bin_fill((int)g.izipx, L, boxes.p, nb, xpc(), rhocc(), off.p, S, s);
        }
      }
      S.entries = lppent[iapm].p;
      const int unroll = prec == 32 ? tune_pp(iapm, L, boxes.p, nb, S) : 6;
      if (SZ)
        sig_add({(uint64_t)iapm, (uint64_t)nb, (uint64_t)(uintptr_t)boxes.p, (uint64_t)(uintptr_t)S.start,
                 (uint64_t)(uintptr_t)S.entries, (uint64_t)(uintptr_t)S.claim, (uint64_t)unroll,
                 fbits(mass_scale)});
      if (CP) {
        KScope k(&tm, "pp_force", (double)total * 16);
        pp_force(L, prec, boxes.p, nb, vpc(), S, o, s, unroll);
      }
    }
  }
}

// Autotuner.  A choice whose candidates all give the same result (the same
// sums in the same order, or integer sums) is timed on its first use: every
// candidate runs twice, timed with HIP events on the library stream, and the
// first candidate (in the caller's order of preference) within 1% of the
// fastest is kept for the context (and in CUBE_AUTOTUNE_FILE when set).  The
// candidate times are kept for the timing report.  After each trial the
// result buffer is checksummed; candidates that disagree stop the run with an
// error, so a pick can never change a result.
template <class Run>
int cube_p3m_ctx::autotune(const std::string& key, const std::vector<int>& cand, Run&& run, const void* result,
                           size_t result_bytes) {
  auto known = tune_picks.find(key);
  if (known != tune_picks.end()) return known->second;
  // no trials while another stream computes (they would share its scratch and
  // see its velocities change): the first candidate now, the trials at the
  // next use on the first stream
  if (no_trials) {
    trial_skipped = true;
    return cand[0];
  }
  int pick = 0;
  struct Ev {  // destroyed on every exit, errors included
    hipEvent_t e = nullptr;
    ~Ev() {
      if (e) (void)hipEventDestroy(e);
    }
  } a, b;
  CUBE_HIP(hipEventCreate(&a.e));
  CUBE_HIP(hipEventCreate(&b.e));
  std::vector<float> best(cand.size(), 0.0f);
  uint64_t sum0 = 0;
  for (int rep = 0; rep < 2; ++rep)
    for (size_t ci = 0; ci < cand.size(); ++ci) {
      CUBE_HIP(hipEventRecord(a.e, s));
      run(cand[ci]);
      CUBE_HIP(hipEventRecord(b.e, s));
      CUBE_HIP(hipEventSynchronize(b.e));
      float ms = 0;
      CUBE_HIP(hipEventElapsedTime(&ms, a.e, b.e));
      if (best[ci] == 0 || ms < best[ci]) best[ci] = ms;
      const uint64_t sum = device_checksum(result, result_bytes, tune_sum.ensure(1), s);
      if (rep == 0 && ci == 0) sum0 = sum;
      else if (sum != sum0)
        throw Error(CUBE_ERR_HIP, "autotune " + key + ": candidate " + std::to_string(cand[ci]) +
                                      " gives a different result");
    }
  const float fastest = *std::min_element(best.begin(), best.end());
  pick = cand.back();
  for (size_t ci = 0; ci < cand.size(); ++ci)
    if (best[ci] <= 1.01f * fastest) {
      pick = cand[ci];
      break;
    }
  tuned[key] = TuneRecord{cand, best, pick};
  tune_picks[key] = pick;
  tune_file_append(key, pick);
  return pick;
}

// Block shape of a tile / sub-tile level's interpolation (cell-group form,
// FP32), tuned on the level's first one-pass compute: each shape runs on a
// scratch copy of the velocities (the level's real potential, kick and
// boxes), preference 0 (8x8x16 cells, 256 threads), 1 (8x16x16, 512), 2
// (16x16x8, 512).  Graph captures, the overlapped kick's passes and untried
// levels otherwise use the pick so far, or shape 0; every shape gives the same
// velocities, so the choice changes the time only.
template <typename R>
int cube_p3m_ctx::interp_shape(int level, int n, const PMLevel& L, const PMBox* boxes, int nb, const cube_kick_in& in,
                               const R* phi, bool trials) {
  if (!std::is_same<R, float>::value || level < 2 || !pm_interp_cells_form(L)) return 0;
  const std::string key = tune_key("interp_shape", level, n);
  const auto it = tune_picks.find(key);
  if (it != tune_picks.end()) return it->second;
  if (!trials || no_trials) return 0;
  vtune.ensure((size_t)std::max<int64_t>(np, 1) * 3);
  pmtune.ensure(1);
  return autotune(
      key, kTuneInterpShape,
      [&](int sh) {
        CUBE_HIP(hipMemcpyAsync(vtune.p, vpc(), sizeof(float) * 3 * (size_t)np, hipMemcpyDeviceToDevice, s));
        zero_async(pmtune.p, sizeof(PMOut), s);
        pm_interp<R>((int)g.izipx, L, boxes, nb, xpc(), vtune.p, rhocc(), off.p, phi,
                     kick_coef(in), pmtune.p, s, sh);
      },
      vtune.p, sizeof(float) * 3 * (size_t)np);
}

// PP load depth (candidates in flight per lane), tuned per PP level on the
// level's first batch.  The trials run on a scratch copy of the velocities and
// of the outputs.  The preference order 6, 4, 8 keeps the middle depth unless
// another is more than 1% faster (a deeper one holds more registers, which the
// kernel's 64-VGPR bound turns into spill traffic).  All depths sum the pairs
// in the same order, so the choice changes the time of the step only.
int cube_p3m_ctx::tune_pp(int iapm, const PPLevel& L, const PPBox* boxes, int nb, PPScratch& S) {
  vtune.ensure((size_t)std::max<int64_t>(np, 1) * 3);
  otune.ensure(1);
  return autotune(
      tune_key("pp_depth", iapm, L.npgrid), kTunePPDepth,
      [&](int u) {
        CUBE_HIP(hipMemcpyAsync(vtune.p, vpc(), sizeof(float) * 3 * (size_t)np, hipMemcpyDeviceToDevice, s));
        zero_async(otune.p, sizeof(PPOut), s);
        pp_force(L, 32, boxes, nb, vtune.p, S, otune.p, s, u);
      },
      vtune.p, sizeof(float) * 3 * (size_t)np);
}

template <typename R>
void cube_p3m_ctx::kick_pass(const cube_kick_in& in, cube_kick_out& out, float* density_host, int mask,
                             bool with_buffer_v, std::vector<int32_t>& ires_h, std::vector<int32_t>& cmax_h) {
  const bool SZ = kph != 2, CP = kph != 1;
  if (CP && !skip_zero) {
    zero_async(pmout.p, sizeof(PMOut) * 4, s);
    zero_async(ppout.p, sizeof(PPOut), s);
    if (nu) zero_async(nu->pmout.p, sizeof(PMOut) * 4, s);
  }
  auto fc = [&](int l) -> const FFTCube<R>& {
    if constexpr (std::is_same<R, float>::value) return fcf[l];
    else return fcd[l];
  };
  const int64_t nnt = g.nnt;
  // PM1 (kick.f90:34-38, particle_mesh.f90)
  if (mask & 1) {
    PMLevel L{};
    L.n = (int)g.nc; L.m0 = 1; L.periodic = 1; L.win_lo = 1; L.win_hi = (int)g.nt; L.nloop = 1; L.w = 1;
    L.rcs = (int)g.rcs; L.gridsize = (double)(float)g.rcs; L.nforce = (int)g.nt;
    L.nte = (int)g.nte; L.ncb = (int)g.ncb; L.nt = (int)g.nt;
    std::vector<PMBox> bx;
    for (int64_t T = 0; T < g.ntiles; ++T) {
      PMBox b{};
      b.T = (int)T;
      const int64_t it[3] = {T % nnt, (T / nnt) % nnt, T / (nnt * nnt)};
      for (int d = 0; d < 3; ++d) {
        b.nc1[d] = 1; b.nc2[d] = (int)g.nt; b.ushift[d] = 0; b.org[d] = (int)(it[d] * g.nt);
      }
      b.moff = 0;
      bx.push_back(b);
    }
    run_level<R>(1, 0, bx, tf[1].p, fc(1), L, in, pmout.p + 0, "pm1_", density_host);
  }
  // PM2 (kick.f90:47-69)
  if (mask & 2) {
    PMLevel L{};
    L.n = (int)g.ngt; L.m0 = 1 - (int)g.ngb; L.periodic = 0; L.win_lo = 1 - (int)g.ngb;
    L.win_hi = (int)(g.ngp + g.ngb); L.nloop = (int)g.ncb; L.w = (int)g.rcs; L.rcs = (int)g.rcs; L.gridsize = 1.0;
    L.nforce = (int)g.ngp; L.nte = (int)g.nte; L.ncb = (int)g.ncb; L.nt = (int)g.nt;
    std::vector<PMBox> bx;
    for (int64_t T = 0; T < g.ntiles; ++T) {
      PMBox b{};
      b.T = (int)T;
      for (int d = 0; d < 3; ++d) { b.nc1[d] = 1; b.nc2[d] = (int)g.nt; b.ushift[d] = 0; b.org[d] = 0; }
      bx.push_back(b);
    }
    run_level<R>(2, 1, bx, tf[2].p, fc(2), L, in, pmout.p + 1, "pm2_");
  }
  // classification (kick.f90:108-138); a compute-only pass reuses the sizing
  // pass's classes
  if ((mask & 12) && SZ && !classified) {
    classify_subtiles(g, cfg.ires_fixed, rhocc(), ires.p, napm.p, cmax.p, s);
    CUBE_HIP(hipMemcpyAsync(ires_h.data(), ires.p, sizeof(int32_t) * g.ns3, hipMemcpyDeviceToHost, s));
    CUBE_HIP(hipMemcpyAsync(cmax_h.data(), cmax.p, sizeof(int32_t) * g.ns3, hipMemcpyDeviceToHost, s));
    CUBE_HIP(hipMemcpyAsync(out.napm, napm.p, sizeof(int32_t) * 7, hipMemcpyDeviceToHost, s));
    CUBE_HIP(hipStreamSynchronize(s));
  }
  // PM3 (kick.f90:203-239)
  if ((mask & 4) && cfg.pm3) {
    const int64_t nns = g.nns;
    for (int iapm = 2; iapm <= 5; ++iapm) {
      PMLevel L{};
      L.n = (int)g.nft[iapm]; L.m0 = 1 - (int)g.nfb[iapm]; L.periodic = 0; L.win_lo = 1 - (int)g.nfb[iapm];
      L.win_hi = (int)(g.nfp[iapm] + g.nfb[iapm]); L.nloop = 1; L.w = (int)(g.rcs * g.sf[iapm]);
      L.rcs = (int)g.rcs; L.gridsize = (double)(1.0f / (float)g.sf[iapm]); L.nforce = (int)g.nfp[iapm];
      L.nte = (int)g.nte; L.ncb = (int)g.ncb; L.nt = (int)g.nt;
      std::vector<PMBox> bx;
      for (int64_t it = 0; it < g.ns3; ++it) {
        if (ires_h[it] != iapm) continue;
        PMBox b{};
        const int64_t ifs[3] = {it % nns, (it / nns) % nns, (it / (nns * nns)) % nns};
        b.T = (int)(it / (nns * nns * nns));
        for (int d = 0; d < 3; ++d) {
          b.nc1[d] = (int)(ifs[d] * g.ntt + 1);
          b.nc2[d] = (int)((ifs[d] + 1) * g.ntt);
          b.ushift[d] = (int)(ifs[d] * g.nfp[iapm]);
          b.org[d] = 0;
        }
        bx.push_back(b);
      }
      char tag[32];
      snprintf(tag, sizeof(tag), "pm3_%d_", iapm);
      run_level<R>(3, iapm, bx, tf[iapm + 1].p, fc(iapm + 1), L, in, pmout.p + 2, tag);
    }
  }
  // PP (kick.f90:280-298)
  if ((mask & 8) && cfg.pp) pp_phase(ires_h, cmax_h, in, ppout.p);
  // buffer_v after the kick (main.f90:38)
  if (with_buffer_v && CP) {
    KScope k(&tm, "buffer_v", (double)np * 12);
    buffer_v(g, off.p, vpc(), nullptr, s);
  }
}

// Sub-tile classes (one read-back) and the workspace of the largest level,
// grown before any level is queued; true when a PM3 or PP level needs several
// workspace batches (its lists then cannot be built ahead of its compute).
template <typename R>
bool cube_p3m_ctx::classify_and_reserve(cube_kick_out& out, std::vector<int32_t>& ires_h,
                                        std::vector<int32_t>& cmax_h) {
  using C = typename Cplx<R>::T;
  classify_subtiles(g, cfg.ires_fixed, rhocc(), ires.p, napm.p, cmax.p, s);
  CUBE_HIP(hipMemcpyAsync(ires_h.data(), ires.p, sizeof(int32_t) * g.ns3, hipMemcpyDeviceToHost, s));
  CUBE_HIP(hipMemcpyAsync(cmax_h.data(), cmax.p, sizeof(int32_t) * g.ns3, hipMemcpyDeviceToHost, s));
  CUBE_HIP(hipMemcpyAsync(out.napm, napm.p, sizeof(int32_t) * 7, hipMemcpyDeviceToHost, s));
  CUBE_HIP(hipStreamSynchronize(s));
  classified = true;
  bool multi = false;  // a PM3 or PP level in several workspace batches
  {
    int64_t need = 0;
    auto level_ws = [&](int64_t n, int64_t nmesh, bool sub) {
      if (nmesh <= 0) return;
      const int64_t per = n * n * n * (int64_t)sizeof(R) + n * n * (n / 2 + 1) * (int64_t)sizeof(C);
      const int64_t batch = std::max<int64_t>(1, std::min<int64_t>(nmesh, ws_budget() / per));
      need = std::max(need, batch * per);
      if (sub && batch < nmesh) multi = true;
    };
    level_ws(g.nc, 1, false);
    level_ws(g.ngt, g.ntiles, false);
    for (int iapm = 2; iapm <= 5; ++iapm) {
      const int64_t nb = (int64_t)std::count(ires_h.begin(), ires_h.end(), iapm);
      level_ws(g.nft[iapm], nb, true);
      // the PP batch rule of pp_phase
      const int64_t m = (int64_t)g.ntt * kRatioSF[iapm] + 2, m3 = m * m * m * kSX;
      if (nb > 0 && std::max<int64_t>(1, ws_budget() / 4 / (m3 * 12)) < nb) multi = true;
    }
    work.ensure((size_t)need);
  }
  return multi;
}

// Full kick with the PM3/PP sizing on the second stream, overlapped with the
// PM1 and PM2 levels on the first (only the host-visible order changes: the
// same kernels on the same data, so the result is the same).  The sizing
// reads positions and counts only and writes the levels' own list buffers;
// the workspace meshes are grown for every level before anything is queued,
// so no buffer the first stream uses is reallocated while it runs.
template <typename R>
void cube_p3m_ctx::kick_overlapped(const cube_kick_in& in, cube_kick_out& out, bool with_buffer_v,
                                   std::vector<int32_t>& ires_h, std::vector<int32_t>& cmax_h) {
  struct Reset {  // one-pass kicks on the first stream again on every exit, errors included
    cube_p3m_ctx* c;
    hipStream_t s0;
    bool timing;
    ~Reset() {
      c->s = s0;
      c->kph = 0;
      c->classified = false;
      c->skip_zero = false;
      c->no_trials = false;
      c->tm.enabled = timing;
    }
  } reset{this, s, tm.enabled};
  // 1. classes (one read-back) and the largest level workspace
  const bool multi = classify_and_reserve<R>(out, ires_h, cmax_h);
  if (multi) {  // lists per batch: PM1/PM2, then PM3/PP, in one pass each on the first stream
    kph = 0;
    kick_pass<R>(in, out, nullptr, 3, false, ires_h, cmax_h);
    skip_zero = true;
    kick_pass<R>(in, out, nullptr, 12, with_buffer_v, ires_h, cmax_h);
    return;
  }
  // 2. PM1 and PM2 on the first stream
  kph = 0;
  kick_pass<R>(in, out, nullptr, 3, false, ires_h, cmax_h);
  skip_zero = true;
  // 3. PM3 and PP sizing on the second stream (untimed: the timer records on
  //    the first stream)
  tm.enabled = false;
  const hipStream_t s0 = s;
  s = s2;
  kph = 1;
  graph_ok = true;
  gsig_acc.clear();
  no_trials = true;
  kick_pass<R>(in, out, nullptr, 12, false, ires_h, cmax_h);
  no_trials = false;
  CUBE_HIP(hipEventRecord(ev_sized, s2));
  s = s0;
  tm.enabled = reset.timing;
  CUBE_HIP(hipStreamWaitEvent(s, ev_sized, 0));
  // 4. PM3 and PP compute (a level needing several workspace batches sizes
  //    and computes again in one pass: its batches share one set of lists)
  kph = graph_ok ? 2 : 0;
  kick_pass<R>(in, out, nullptr, 12, with_buffer_v, ires_h, cmax_h);
}

// Graph mode: the overlapped kick's two compute parts as two graphs.  Part 0
// (PM1 + PM2) is sized and launched on the library stream; part 1 (PM3 + PP +
// buffer_v) is sized on the second stream meanwhile and launched after it.
// Each part is captured on its first use and replayed while its launch
// signature holds.  A part whose sizing left autotuner trials untried runs
// uncaptured (the trials run there, on the library stream) and is captured at
// a later kick.
template <typename R>
void cube_p3m_ctx::kick_graph(const cube_kick_in& in, cube_kick_out& out, bool with_buffer_v,
                              std::vector<int32_t>& ires_h, std::vector<int32_t>& cmax_h) {
  struct Reset {
    cube_p3m_ctx* c;
    hipStream_t s0;
    bool timing;
    ~Reset() {
      c->s = s0;
      c->kph = 0;
      c->classified = false;
      c->skip_zero = false;
      c->no_trials = false;
      c->tm.enabled = timing;
    }
  } reset{this, s, tm.enabled};
  const hipStream_t s0 = s;
  if (classify_and_reserve<R>(out, ires_h, cmax_h)) {
    // several workspace batches in a level: uncaptured, one pass per part
    graph_drop();
    kph = 0;
    kick_pass<R>(in, out, nullptr, 3, false, ires_h, cmax_h);
    skip_zero = true;
    kick_pass<R>(in, out, nullptr, 12, with_buffer_v, ires_h, cmax_h);
    return;
  }
  kcoef.ensure(1);
  hipLaunchKernelGGL(k_set_f64, dim3(1), dim3(1), 0, s, kcoef.p, kick_coef(in));
  CUBE_HIP(hipGetLastError());
  auto common = [&](int part) {
    sig_add({(uint64_t)part, (uint64_t)prec, fbits(in.mass_p_cdm), (uint64_t)(part == 1 && with_buffer_v),
             (uint64_t)(uintptr_t)xpc(), (uint64_t)(uintptr_t)vpc(), (uint64_t)(uintptr_t)rhocc(),
             (uint64_t)(uintptr_t)off.p, (uint64_t)(uintptr_t)pmout.p, (uint64_t)(uintptr_t)ppout.p,
             (uint64_t)(uintptr_t)kcoef.p});
  };
  // capture (when the signature changed) and launch one part's compute pass
  auto launch = [&](int part, int pmask, bool bufv) {
    kph = 2;
    if (!gexec[part] || gsig_acc != gsig[part]) {
      if (gexec[part]) (void)hipGraphExecDestroy(gexec[part]);
      gexec[part] = nullptr;
      gsig[part].clear();
      hipGraph_t gr = nullptr;
      CUBE_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
      try {
        kick_pass<R>(in, out, nullptr, pmask, bufv, ires_h, cmax_h);
      } catch (...) {
        hipGraph_t tmp = nullptr;
        (void)hipStreamEndCapture(s, &tmp);
        if (tmp) (void)hipGraphDestroy(tmp);
        throw;
      }
      CUBE_HIP(hipStreamEndCapture(s, &gr));
      const hipError_t e = hipGraphInstantiate(&gexec[part], gr, nullptr, nullptr, 0);
      (void)hipGraphDestroy(gr);
      if (e != hipSuccess) {
        gexec[part] = nullptr;
        throw Error(CUBE_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
      }
      gsig[part] = gsig_acc;
      ++graph_captures;
    }
    CUBE_HIP(hipGraphLaunch(gexec[part], s));
  };
  tm.enabled = false;
  // part 0: PM1 + PM2 (its sizing is box tables, stencil and, at first use,
  // the PM2 FFT shape trials)
  kph = 1;
  gsig_acc.clear();
  kick_pass<R>(in, out, nullptr, 3, false, ires_h, cmax_h);
  common(0);
  launch(0, 3, false);
  // part 1 sizing on the second stream, without trials
  s = s2;
  kph = 1;
  gsig_acc.clear();
  no_trials = true;
  trial_skipped = false;
  kick_pass<R>(in, out, nullptr, 12, false, ires_h, cmax_h);
  no_trials = false;
  CUBE_HIP(hipEventRecord(ev_sized, s2));
  s = s0;
  CUBE_HIP(hipStreamWaitEvent(s, ev_sized, 0));
  skip_zero = true;
  if (trial_skipped) {
    kph = 2;  // uncaptured: the trials run now, on the library stream
    kick_pass<R>(in, out, nullptr, 12, with_buffer_v, ires_h, cmax_h);
  } else {
    common(1);
    launch(1, 12, with_buffer_v);
    ++graph_launches;
  }
}

template <typename R>
void cube_p3m_ctx::kick(const cube_kick_in& in, cube_kick_out& out, float* density_host, bool with_buffer_v) {
  if (layout != 1) throw Error(CUBE_ERR_STATE, "kick needs the buffered (extended) particle layout");
  std::memset(&out, 0, sizeof(out));
  ensure_offsets();
  const int mask = in.phase_mask ? in.phase_mask : 15;
  std::vector<int32_t> ires_h(g.ns3, 2), cmax_h(g.ns3, 0);
  struct Hold {  // the workspace budget is read once per kick, so both passes batch alike
    int64_t& w;
    ~Hold() { w = -1; }
  } hold{wsb};
  wsb = -1;
  wsb = ws_budget();
  const bool use_graph = graph_on && mask == 15 && !density_host && !nu;
  const bool overlap =
      !use_graph && overlap_on && mask == 15 && !density_host && !nu && cfg.pm3 && cfg.pp && kicks_done > 0;
  ++kicks_done;
  if (overlap) {
    kick_overlapped<R>(in, out, with_buffer_v, ires_h, cmax_h);
  } else if (!use_graph) {
    kph = 0;
    kick_pass<R>(in, out, density_host, mask, with_buffer_v, ires_h, cmax_h);
  } else {
    kick_graph<R>(in, out, with_buffer_v, ires_h, cmax_h);
  }
  // collect reductions
  PMOut po[4];
  PPOut pq;
  CUBE_HIP(hipMemcpyAsync(po, pmout.p, sizeof(PMOut) * 4, hipMemcpyDeviceToHost, s));
  CUBE_HIP(hipMemcpyAsync(&pq, ppout.p, sizeof(PPOut), hipMemcpyDeviceToHost, s));
  CUBE_HIP(hipStreamSynchronize(s));
  auto f2 = [&](const PMOut& p) { return std::is_same<R, float>::value ? p.f2max : (float)p.f2max_d; };
  auto vm = [&](const PMOut& p, int d) { return std::is_same<R, float>::value ? p.vmax[d] : (float)p.vmax_d[d]; };
  float vmax[3] = {0, 0, 0};
  if (mask & 1) {
    out.f2max_pm1 = f2(po[0]);
    out.dt_pm1 = dt_from_f2(0.5f, out.f2max_pm1, (float)in.a_mid);
    unsigned long long tfx = 0;
    double td = 0;
    for (int q = 0; q < kSumSlots; ++q) {
      tfx += po[0].testrho_fx[q];
      td += po[0].testrho_d[q];
    }
    out.testrho = std::is_same<R, float>::value ? (double)tfx / 4294967296.0 : td;
    for (int d = 0; d < 3; ++d) vmax[d] = vm(po[0], d);
  }
  if (mask & 2) {
    out.f2max_pm2 = f2(po[1]);
    out.dt_pm2 = dt_from_f2(0.5f, out.f2max_pm2, (float)in.a_mid);
    for (int d = 0; d < 3; ++d) vmax[d] = vm(po[1], d);
  }
  if ((mask & 4) && cfg.pm3) {
    out.f2max_pm3 = f2(po[2]);
    out.dt_pm3 = dt_from_f2(0.5f, out.f2max_pm3, (float)in.a_mid);
    for (int d = 0; d < 3; ++d) vmax[d] = vm(po[2], d);
  }
  if ((mask & 8) && cfg.pp) {
    out.f2max_pp = pq.f2max;
    out.dt_pp = dt_from_f2(0.05f, pq.f2max, (float)in.a_mid);
    for (int d = 0; d < 3; ++d) vmax[d] = pq.vmax[d];
    unsigned long long npp = 0, npg = 0;
    for (int q = 0; q < kSumSlots; ++q) {
      npp += pq.pairs_pp[q];
      npg += pq.pairs_pg[q];
    }
    last_pairs = (int64_t)(npp / 2 + npg);
  }
  for (int d = 0; d < 3; ++d) out.vmax[d] = vmax[d];
  out.dt_vmax = (0.9f * 14.0f) / std::max({vmax[0], vmax[1], vmax[2]});  // kick.f90:314
  // every physical particle took the last phase's kick, whose vmax covers it
  vmax_valid = mask == 15;
  for (int d = 0; d < 3; ++d) last_vmax[d] = vmax[d];
  out.vz_max = vmax[2];                                                     // kick.f90:315
}

// drift (drift.f90) on the resident extended layout.
void cube_p3m_ctx::drift(const cube_drift_in& in, cube_drift_out& out) {
  if (layout != 1) throw Error(CUBE_ERR_STATE, "drift needs the buffered (extended) particle layout");
  ensure_offsets();
  DriftParams P{};
  P.dt_mid = (in.dt_old + in.dt) / 2.0;                                                     // drift.f90:21
  P.nlayer = 2 * (int)std::ceil((P.dt_mid * (double)in.vz_max) / (double)g.rcs) + 2;        // drift.f90:26
  P.nt = (int)g.nt; P.nte = (int)g.nte; P.ncb = (int)g.ncb; P.nnt = (int)g.nnt; P.rcs = (int)g.rcs;
  P.inv_rcs = 1.0 / (double)g.rcs;
  P.izipx = (int)g.izipx;
  out.nlayer = P.nlayer;
  const int64_t nt3 = g.nt * g.nt * g.nt, nphys_cells = g.ntiles * nt3;
  DriftScratch S;
  S.flags = dflags.p;
  S.newcount = dnewcount.ensure(nphys_cells);
  S.base = dbase.ensure(nphys_cells + 1);
  const int other = 1 - cur;
  int32_t fl[5];
  int64_t nplocal = 0;
  // tile overflow (drift.f90:85-93): every particle of the extended tile lands
  // in the tile's rhoce range, so its total is the extended tile count; the
  // check runs on the device and is read back with the pass's flags (the
  // passes before the read write scratch arrays only)
  auto start_pass = [&]() {
    zero_async(S.flags, sizeof(int32_t) * 8, s);
    zero_async(S.newcount, sizeof(int32_t) * nphys_cells, s);
    drift_tile_check(off.p, g.ncell_tile, (int)g.ntiles, g.np_tile_max, S.flags, s);
  };
  auto table_tiles = [&](int64_t Dt) {  // tiles per rank/offset batch within a 2 GiB table
    const int64_t per_tile = g.ncell_tile * Dt;
    const int64_t tbudget = std::max<int64_t>(per_tile, ((int64_t)2 << 30) / 4);
    return (int)std::max<int64_t>(1, std::min<int64_t>(g.ntiles, tbudget / per_tile));
  };
  auto count_scan = [&]() {  // flags and the scan of the destination counts
    CUBE_HIP(hipMemcpyAsync(fl, S.flags, sizeof(fl), hipMemcpyDeviceToHost, s));
    {
      KScope k(&tm, "drift_scan", (double)nphys_cells * 12);
      scan(S.newcount, S.base, nphys_cells);
    }
    CUBE_HIP(hipMemcpyAsync(&nplocal, S.base + nphys_cells, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    CUBE_HIP(hipStreamSynchronize(s));
    if (fl[0] & DRIFT_TILE_OVERFLOW)
      throw Error(CUBE_ERR_TILE_OVERFLOW, "too many particles in this tile+buffer: " + std::to_string(fl[4]) + " > " +
                                              std::to_string(g.np_tile_max) + "; please set tile_buffer larger");
    out.overhead_tile = (float)fl[4] / (float)g.np_tile_max;
    if (fl[0] & DRIFT_OUT_OF_RANGE)
      throw Error(CUBE_ERR_OUT_OF_RANGE, "error: particle out of range (drift beyond 2*ncb coarse cells)");
  };
  auto offsets_scatter = [&](int t0, int t1) {
    {
      KScope k(&tm, "drift_offsets", (double)(t1 - t0) * nt3 * P.D[0] * P.D[1] * P.D[2] * 8);
      drift_offsets(P, t0, t1, S, s);
    }
    {
      KScope k(&tm, "drift_scatter", (double)np * 2 * (3 * esz() + 20) * (t1 - t0) / g.ntiles);
      drift_scatter(P, xpc(), vpc(), pidc(), rhocc(), off.p, t0, t1, S, xp[other].p, vp[other].p,
                    has_pid ? pid[other].p : nullptr, s);
    }
  };
  S.rank = drank.ensure((size_t)std::max<int64_t>(np, 1));
  // After a full resident kick the cell shift per axis is bounded by
  // ceil(fl32(fl32(dt_mid vmax)/rcs)) (a particle in cell i has xq in (i-1, i)
  // and moves by fl32(fl32(dt_mid v)/rcs), monotone in |v|), so the range
  // check and the destination counts ride on the rank pass's one read of the
  // particles; a shift beyond the bound (stale velocities) falls back to the
  // separate pass.
  bool fused = vmax_valid;
  int Mb[3] = {0, 0, 0};
  for (int d = 0; d < 3 && fused; ++d) {
    const double dm = std::fabs((P.dt_mid * (double)last_vmax[d]) / (double)g.rcs);
    if (!(dm <= (double)kMaxShift)) fused = false;
    else Mb[d] = (int)std::ceil(dm);
  }
  if (fused) fused = table_tiles((int64_t)(2 * Mb[0] + 1) * (2 * Mb[1] + 1) * (2 * Mb[2] + 1)) == (int)g.ntiles;
  if (fused) {
    for (int d = 0; d < 3; ++d) {
      P.M[d] = Mb[d];
      P.D[d] = 2 * Mb[d] + 1;
    }
    const int64_t Dt = (int64_t)P.D[0] * P.D[1] * P.D[2];
    S.table = dtable.ensure((size_t)(g.ntiles * g.ncell_tile * Dt));
    start_pass();
    zero_async(S.table, sizeof(int32_t) * (size_t)(g.ntiles * g.ncell_tile * Dt), s);
    {
      KScope k(&tm, "drift_rank", (double)np * (3 * esz() + 16));
      drift_rank(P, xpc(), vpc(), rhocc(), off.p, 0, (int)g.ntiles, S, s, true);
    }
    count_scan();
    if (fl[0] & DRIFT_SHIFT_BOUND) fused = false;
    else {
      ensure_set(other, nplocal);
      offsets_scatter(0, (int)g.ntiles);
    }
  }
  if (!fused) {
    start_pass();
    {
      KScope k(&tm, "drift_shift", (double)np * (3 * esz() + 12));
      drift_shift(P, xpc(), vpc(), rhocc(), off.p, g.ncells_ext, S, s);
    }
    count_scan();
    for (int d = 0; d < 3; ++d) {
      if (fl[1 + d] > kMaxShift) throw Error(CUBE_ERR_OUT_OF_RANGE, "drift shift beyond 16 coarse cells");
      P.M[d] = fl[1 + d];
      P.D[d] = 2 * P.M[d] + 1;
    }
    ensure_set(other, nplocal);
    const int64_t Dt = (int64_t)P.D[0] * P.D[1] * P.D[2];
    const int tg = table_tiles(Dt);
    S.table = dtable.ensure((size_t)(tg * g.ncell_tile * Dt));
    for (int t0 = 0; t0 < (int)g.ntiles; t0 += tg) {
      const int t1 = std::min<int>((int)g.ntiles, t0 + tg);
      zero_async(S.table, sizeof(int32_t) * (size_t)(t1 - t0) * g.ncell_tile * Dt, s);
      {
        KScope k(&tm, "drift_rank", (double)np * (3 * esz() + 16) * (t1 - t0) / g.ntiles);
        drift_rank(P, xpc(), vpc(), rhocc(), off.p, t0, t1, S, s);
      }
      offsets_scatter(t0, t1);
    }
  }
  int32_t* rnew = rhoc[1 - rcur].p;
  drift_write_rhoc(P, S, rnew, s);
  CUBE_HIP(hipMemcpyAsync(fl, S.flags, sizeof(fl), hipMemcpyDeviceToHost, s));
  CUBE_HIP(hipStreamSynchronize(s));
  if (fl[0] & 4) throw Error(CUBE_ERR_COUNT_MISMATCH, "warning: density different (rholocal /= rhoce)");
  if (fl[0] & DRIFT_PID_ZERO) throw Error(CUBE_ERR_PID, "found pid=0");
  cur = other;
  rcur = 1 - rcur;
  layout = 2;
  np = nplocal;
  off_valid = false;
  out.nplocal = nplocal;
}

// buffer_grid + buffer_x + buffer_v at nn=1.
void cube_p3m_ctx::buffer(int64_t* np_ext) {
  if (layout != 2) throw Error(CUBE_ERR_STATE, "buffer needs the physical layout produced by drift");
  const int64_t nt3 = g.nt * g.nt * g.nt, nphys_cells = g.ntiles * nt3;
  int32_t* rp = rhocc();
  int32_t* re = rhoc[1 - rcur].p;
  physcount.ensure(nphys_cells);
  physoff.ensure(nphys_cells + 1);
  {
    KScope k(&tm, "buffer_rhoc", (double)g.ncells_ext * 8);
    buffer_rhoc(g, rp, re, physcount.p, s);
  }
  rcur = 1 - rcur;
  scan(re, off.p, g.ncells_ext);
  scan(physcount.p, physoff.p, nphys_cells);
  int64_t n = 0;
  CUBE_HIP(hipMemcpyAsync(&n, off.p + g.ncells_ext, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  CUBE_HIP(hipStreamSynchronize(s));
  if (n > g.np_image_max)
    throw Error(CUBE_ERR_IMAGE_OVERFLOW, "too many particles in this image+buffer: " + std::to_string(n) + " > " +
                                             std::to_string(g.np_image_max) + "; please set image_buffer larger");
  ensure_set(1 - cur, n);
  const int other = 1 - cur;
  {
    KScope k(&tm, "buffer_copy", (double)n * (3 * esz() + 20) * 2);
    buffer_copy(g, re, off.p, physoff.p, xpc(), vpc(), pidc(), xp[other].p, vp[other].p,
                has_pid ? pid[other].p : nullptr, s);
  }
  cur = other;
  layout = 1;
  np = n;
  off_valid = true;
  if (np_ext) *np_ext = n;
}

// buffer_v at nn=1 on the resident extended layout.
void cube_p3m_ctx::buffer_vel() {
  if (layout != 1) throw Error(CUBE_ERR_STATE, "buffer_v needs the buffered (extended) particle layout");
  ensure_offsets();
  KScope k(&tm, "buffer_v", (double)np * 12);
  buffer_v(g, off.p, vpc(), nullptr, s);
}

// ---------------------------------------------------------------------------
// host spine (spine_image, variables.f90:225-295)
static void host_spine(const Geo& g, const int32_t* rhoc, int64_t* idx_b_l, int64_t* idx_b_r, int64_t* ppl0,
                       int64_t* pplr, int64_t* pprl, int64_t* ppr0, int64_t* ppl, int64_t* ppr) {
  const int64_t nt = g.nt, ncb = g.ncb, nte = g.nte;
  std::vector<int64_t> bl(g.ntiles * nte * nte), br(g.ntiles * nte * nte);
  int64_t nsum = 0;
  for (int64_t T = 0; T < g.ntiles; ++T)
    for (int64_t k = 1 - ncb; k <= nt + ncb; ++k)
      for (int64_t j = 1 - ncb; j <= nt + ncb; ++j) {
        const int64_t r = (T * nte + (k - 1 + ncb)) * nte + (j - 1 + ncb);
        bl[r] = nsum;
        for (int64_t i = 1 - ncb; i <= nt + ncb; ++i) nsum += rhoc[rc_index(i, j, k, T, nte, ncb)];
        br[r] = nsum;
      }
  if (idx_b_l) std::memcpy(idx_b_l, bl.data(), sizeof(int64_t) * bl.size());
  if (idx_b_r) std::memcpy(idx_b_r, br.data(), sizeof(int64_t) * br.size());
  int64_t nsum_p = 0;
  for (int64_t T = 0; T < g.ntiles; ++T)
    for (int64_t k = 1; k <= nt; ++k)
      for (int64_t j = 1; j <= nt; ++j) {
        const int64_t q = (T * nt + (k - 1)) * nt + (j - 1);
        const int64_t r = (T * nte + (k - 1 + ncb)) * nte + (j - 1 + ncb);
        int64_t np_phy = 0, lsum = 0, rsum = 0, rl = 0, lr = 0;
        for (int64_t i = 1; i <= nt; ++i) np_phy += rhoc[rc_index(i, j, k, T, nte, ncb)];
        for (int64_t i = 1 - ncb; i <= 0; ++i) lsum += rhoc[rc_index(i, j, k, T, nte, ncb)];
        for (int64_t i = nt + 1; i <= nt + ncb; ++i) rsum += rhoc[rc_index(i, j, k, T, nte, ncb)];
        for (int64_t i = nt - ncb + 1; i <= nt; ++i) rl += rhoc[rc_index(i, j, k, T, nte, ncb)];
        for (int64_t i = 1; i <= ncb; ++i) lr += rhoc[rc_index(i, j, k, T, nte, ncb)];
        if (ppl) ppl[q] = nsum_p;
        nsum_p += np_phy;
        if (ppr) ppr[q] = nsum_p;
        if (ppr0) ppr0[q] = br[r] - rsum;
        if (pprl) pprl[q] = br[r] - rsum - rl;
        if (ppl0) ppl0[q] = bl[r] + lsum;
        if (pplr) pplr[q] = bl[r] + lsum + lr;
      }
}

static int64_t host_sum(const int32_t* a, int64_t n) {
  int64_t s = 0;
  for (int64_t i = 0; i < n; ++i) s += a[i];
  return s;
}

template <class F>
static int guarded(cube_p3m_ctx* ctx, F&& f) {
  if (!ctx) return CUBE_ERR_ARG;
  try {
    CUBE_HIP(hipSetDevice(ctx->device));
    f();
    ctx->err.clear();
    return CUBE_OK;
  } catch (const Error& e) {
    ctx->err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return CUBE_ERR_HIP;
  }
}

static void upload_ext(cube_p3m_ctx* c, const void* xp, const float* vp, const int64_t* pid, const int32_t* rhoc,
                       int64_t np_arr) {
  const int64_t n = host_sum(rhoc, c->g.ncells_ext);
  if (n > np_arr) throw Error(CUBE_ERR_ARG, "particle arrays shorter than sum(rhoc)");
  c->ensure_set(0, std::max<int64_t>(n, 1));
  c->cur = 0;
  c->rcur = 0;
  CUBE_HIP(hipMemcpyAsync(c->rhoc[0].p, rhoc, sizeof(int32_t) * c->g.ncells_ext, hipMemcpyHostToDevice, c->s));
  CUBE_HIP(hipMemcpyAsync(c->xp[0].p, xp, (size_t)n * 3 * c->esz(), hipMemcpyHostToDevice, c->s));
  CUBE_HIP(hipMemcpyAsync(c->vp[0].p, vp, sizeof(float) * 3 * n, hipMemcpyHostToDevice, c->s));
  c->has_pid = pid != nullptr;
  if (pid) CUBE_HIP(hipMemcpyAsync(c->pid[0].p, pid, sizeof(int64_t) * n, hipMemcpyHostToDevice, c->s));
  c->layout = 1;
  c->np = n;
  c->off_valid = false;
  c->vmax_valid = false;
}

// The second species' outputs of a coupled kick: its PM1/PM2 maxima (its
// levels' f2max at its own positions, the Delta-t criteria from them, its
// velocity maxima) and its PM1 density sum.
template <typename R>
static void nu_outputs(cube_p3m_ctx* ctx, cube_p3m_ctx* nu, const cube_kick_in& in, int mask, cube_kick_out& o) {
  std::memset(&o, 0, sizeof(o));
  PMOut pn[4];
  CUBE_HIP(hipMemcpyAsync(pn, nu->pmout.p, sizeof(PMOut) * 4, hipMemcpyDeviceToHost, ctx->s));
  CUBE_HIP(hipStreamSynchronize(ctx->s));
  const bool f = std::is_same<R, float>::value;
  auto f2 = [&](const PMOut& p) { return f ? p.f2max : (float)p.f2max_d; };
  float vmax[3] = {0, 0, 0};
  for (int lv = 0; lv < 2; ++lv) {
    if (!(mask & (1 << lv))) continue;
    const PMOut& p = pn[lv];
    const float q = f2(p);
    if (lv == 0) {
      o.f2max_pm1 = q;
      o.dt_pm1 = dt_from_f2(0.5f, q, (float)in.a_mid);
      unsigned long long tfx = 0;
      double td = 0;
      for (int k = 0; k < kSumSlots; ++k) {
        tfx += p.testrho_fx[k];
        td += p.testrho_d[k];
      }
      o.testrho = f ? (double)tfx / 4294967296.0 : td;
    } else {
      o.f2max_pm2 = q;
      o.dt_pm2 = dt_from_f2(0.5f, q, (float)in.a_mid);
    }
    for (int d = 0; d < 3; ++d) vmax[d] = f ? p.vmax[d] : (float)p.vmax_d[d];
  }
  for (int d = 0; d < 3; ++d) o.vmax[d] = vmax[d];
  o.dt_vmax = (0.9f * 14.0f) / std::max({vmax[0], vmax[1], vmax[2]});
  o.vz_max = vmax[2];
  // PM2 interpolates and kicks every physical particle of the species, and
  // buffer_v copies their velocities: the maxima bound its resident set
  nu->vmax_valid = (mask & 2) != 0;
  for (int d = 0; d < 3; ++d) nu->last_vmax[d] = vmax[d];
}

extern "C" {

const char* cube_p3m_version(void) { return "cube_p3m 0.1 (gfx950)"; }

int cube_p3m_create(const cube_p3m_config* cfg, cube_p3m_ctx** out) {
  if (!cfg || !out) return CUBE_ERR_ARG;
  *out = nullptr;
  auto* c = new cube_p3m_ctx();
  c->cfg = *cfg;
  try {
    c->init();
  } catch (const Error& e) {
    int code = e.code;
    delete c;
    return code;
  } catch (...) {
    delete c;
    return CUBE_ERR_HIP;
  }
  *out = c;
  return CUBE_OK;
}

int cube_p3m_destroy(cube_p3m_ctx* ctx) {
  if (!ctx) return CUBE_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->s);
  delete ctx;
  return CUBE_OK;
}

const char* cube_p3m_last_error(const cube_p3m_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int cube_p3m_set_tf(cube_p3m_ctx* ctx, int32_t level, const float* tf, int64_t n) {
  return guarded(ctx, [&] {
    if (level < 1 || level > 6 || !tf) throw Error(CUBE_ERR_ARG, "bad tf level");
    const int64_t m = ctx->tfn[level];
    if (n != (m / 2 + 1) * m * m) throw Error(CUBE_ERR_ARG, "tf size mismatch");
    CUBE_HIP(hipMemcpyAsync(ctx->tf[level].p, tf, sizeof(float) * n, hipMemcpyHostToDevice, ctx->s));
    CUBE_HIP(hipStreamSynchronize(ctx->s));
  });
}

int cube_p3m_get_tf(cube_p3m_ctx* ctx, int32_t level, float* tf, int64_t n) {
  return guarded(ctx, [&] {
    if (level < 1 || level > 6 || !tf) throw Error(CUBE_ERR_ARG, "bad tf level");
    const int64_t m = ctx->tfn[level];
    if (n != (m / 2 + 1) * m * m) throw Error(CUBE_ERR_ARG, "tf size mismatch");
    CUBE_HIP(hipMemcpyAsync(tf, ctx->tf[level].p, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->s));
    CUBE_HIP(hipStreamSynchronize(ctx->s));
  });
}

static void check_spine(cube_p3m_ctx* c, const int32_t* rhoc, const int64_t* idx_b_r) {
  if (!idx_b_r) return;
  const Geo& g = c->g;
  std::vector<int64_t> br(g.ntiles * g.nte * g.nte);
  host_spine(g, rhoc, nullptr, br.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
  for (size_t i = 0; i < br.size(); ++i)
    if (br[i] != idx_b_r[i]) throw Error(CUBE_ERR_SPINE, "idx_b_r does not match the cumulative sum of rhoc");
}

int cube_p3m_particle_mesh(cube_p3m_ctx* ctx, const void* xp, float* vp, const int32_t* rhoc,
                           const int64_t* idx_b_r, int64_t np, const cube_kick_in* in, float* density,
                           cube_kick_out* out) {
  return guarded(ctx, [&] {
    if (!xp || !vp || !rhoc || !in) throw Error(CUBE_ERR_ARG, "null argument");
    check_spine(ctx, rhoc, idx_b_r);
    upload_ext(ctx, xp, vp, nullptr, rhoc, np);
    cube_kick_in k = *in;
    k.phase_mask = 1;
    cube_kick_out o;
    if (ctx->prec == 32) ctx->kick<float>(k, o, density);
    else ctx->kick<double>(k, o, density);
    CUBE_HIP(hipMemcpyAsync(vp, ctx->vpc(), sizeof(float) * 3 * ctx->np, hipMemcpyDeviceToHost, ctx->s));
    CUBE_HIP(hipStreamSynchronize(ctx->s));
    if (out) *out = o;
  });
}

int cube_p3m_kick(cube_p3m_ctx* ctx, const void* xp, float* vp, const int32_t* rhoc, const int64_t* idx_b_r,
                  int64_t np, const cube_kick_in* in, cube_kick_out* out) {
  return guarded(ctx, [&] {
    if (!xp || !vp || !rhoc || !in) throw Error(CUBE_ERR_ARG, "null argument");
    check_spine(ctx, rhoc, idx_b_r);
    upload_ext(ctx, xp, vp, nullptr, rhoc, np);
    cube_kick_out o;
    if (ctx->prec == 32) ctx->kick<float>(*in, o, nullptr);
    else ctx->kick<double>(*in, o, nullptr);
    CUBE_HIP(hipMemcpyAsync(vp, ctx->vpc(), sizeof(float) * 3 * ctx->np, hipMemcpyDeviceToHost, ctx->s));
    CUBE_HIP(hipStreamSynchronize(ctx->s));
    if (out) *out = o;
  });
}

int cube_p3m_drift(cube_p3m_ctx* ctx, void* xp, float* vp, int64_t* pid, int32_t* rhoc, int64_t np,
                   int64_t* idx_b_l, int64_t* idx_b_r, int64_t* ppl0, int64_t* pplr, int64_t* pprl, int64_t* ppr0,
                   int64_t* ppl, int64_t* ppr, const cube_drift_in* in, cube_drift_out* out) {
  return guarded(ctx, [&] {
    if (!xp || !vp || !rhoc || !in) throw Error(CUBE_ERR_ARG, "null argument");
    upload_ext(ctx, xp, vp, pid, rhoc, np);
    cube_drift_out o{};
    ctx->drift(*in, o);
    const int64_t n = o.nplocal;
    if (n > np) throw Error(CUBE_ERR_IMAGE_OVERFLOW, "output arrays too short for nplocal");
    CUBE_HIP(hipMemcpyAsync(xp, ctx->xpc(), (size_t)n * 3 * ctx->esz(), hipMemcpyDeviceToHost, ctx->s));
    CUBE_HIP(hipMemcpyAsync(vp, ctx->vpc(), sizeof(float) * 3 * n, hipMemcpyDeviceToHost, ctx->s));
    if (pid) CUBE_HIP(hipMemcpyAsync(pid, ctx->pidc(), sizeof(int64_t) * n, hipMemcpyDeviceToHost, ctx->s));
    CUBE_HIP(hipMemcpyAsync(rhoc, ctx->rhocc(), sizeof(int32_t) * ctx->g.ncells_ext, hipMemcpyDeviceToHost, ctx->s));
    CUBE_HIP(hipStreamSynchronize(ctx->s));
    // xp(:,nplocal+1:)=0 etc. (drift.f90:183-186)
    std::memset((char*)xp + (size_t)n * 3 * ctx->esz(), 0, (size_t)(np - n) * 3 * ctx->esz());
    std::memset(vp + 3 * n, 0, sizeof(float) * 3 * (np - n));
    if (pid) std::memset(pid + n, 0, sizeof(int64_t) * (np - n));
    host_spine(ctx->g, rhoc, idx_b_l, idx_b_r, ppl0, pplr, pprl, ppr0, ppl, ppr);
    if (out) *out = o;
  });
}

int cube_p3m_buffer(cube_p3m_ctx* ctx, void* xp, float* vp, int64_t* pid, int32_t* rhoc, int64_t nplocal,
                    int64_t np_cap, int64_t* idx_b_l, int64_t* idx_b_r, int64_t* np_ext) {
  return guarded(ctx, [&] {
    if (!xp || !vp || !rhoc) throw Error(CUBE_ERR_ARG, "null argument");
    const Geo& g = ctx->g;
    // physical counts only: buffers are ignored (filled here)
    std::vector<int32_t> rp(rhoc, rhoc + g.ncells_ext);
    int64_t nphys = 0;
    for (int64_t T = 0; T < g.ntiles; ++T)
      for (int64_t k = 1 - g.ncb; k <= g.nt + g.ncb; ++k)
        for (int64_t j = 1 - g.ncb; j <= g.nt + g.ncb; ++j)
          for (int64_t i = 1 - g.ncb; i <= g.nt + g.ncb; ++i) {
            const int64_t c = rc_index(i, j, k, T, g.nte, g.ncb);
            const bool ph = i >= 1 && i <= g.nt && j >= 1 && j <= g.nt && k >= 1 && k <= g.nt;
            if (!ph) rp[c] = 0;
            else nphys += rp[c];
          }
    if (nphys != nplocal) throw Error(CUBE_ERR_ARG, "nplocal differs from the physical sum of rhoc");
    ctx->ensure_set(0, std::max<int64_t>(nplocal, 1));
    ctx->cur = 0;
    ctx->rcur = 0;
    CUBE_HIP(hipMemcpyAsync(ctx->rhoc[0].p, rp.data(), sizeof(int32_t) * g.ncells_ext, hipMemcpyHostToDevice, ctx->s));
    CUBE_HIP(hipMemcpyAsync(ctx->xp[0].p, xp, (size_t)nplocal * 3 * ctx->esz(), hipMemcpyHostToDevice, ctx->s));
    CUBE_HIP(hipMemcpyAsync(ctx->vp[0].p, vp, sizeof(float) * 3 * nplocal, hipMemcpyHostToDevice, ctx->s));
    ctx->has_pid = pid != nullptr;
    if (pid) CUBE_HIP(hipMemcpyAsync(ctx->pid[0].p, pid, sizeof(int64_t) * nplocal, hipMemcpyHostToDevice, ctx->s));
    ctx->vmax_valid = false;
    ctx->layout = 2;
    ctx->np = nplocal;
    int64_t n = 0;
    ctx->buffer(&n);
    if (n > np_cap) throw Error(CUBE_ERR_IMAGE_OVERFLOW, "output arrays too short for the buffered layout");
    CUBE_HIP(hipMemcpyAsync(xp, ctx->xpc(), (size_t)n * 3 * ctx->esz(), hipMemcpyDeviceToHost, ctx->s));
    CUBE_HIP(hipMemcpyAsync(vp, ctx->vpc(), sizeof(float) * 3 * n, hipMemcpyDeviceToHost, ctx->s));
    if (pid) CUBE_HIP(hipMemcpyAsync(pid, ctx->pidc(), sizeof(int64_t) * n, hipMemcpyDeviceToHost, ctx->s));
    CUBE_HIP(hipMemcpyAsync(rhoc, ctx->rhocc(), sizeof(int32_t) * g.ncells_ext, hipMemcpyDeviceToHost, ctx->s));
    CUBE_HIP(hipStreamSynchronize(ctx->s));
    host_spine(g, rhoc, idx_b_l, idx_b_r, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    if (np_ext) *np_ext = n;
  });
}

int cube_p3m_buffer_v(cube_p3m_ctx* ctx, float* vp, int64_t* pid, const int32_t* rhoc, int64_t np) {
  return guarded(ctx, [&] {
    if (!vp || !rhoc) throw Error(CUBE_ERR_ARG, "null argument");
    const Geo& g = ctx->g;
    const int64_t n = host_sum(rhoc, g.ncells_ext);
    if (n > np) throw Error(CUBE_ERR_ARG, "particle arrays shorter than sum(rhoc)");
    ctx->ensure_set(0, std::max<int64_t>(n, 1));
    ctx->cur = 0;
    ctx->rcur = 0;
    CUBE_HIP(hipMemcpyAsync(ctx->rhoc[0].p, rhoc, sizeof(int32_t) * g.ncells_ext, hipMemcpyHostToDevice, ctx->s));
    CUBE_HIP(hipMemcpyAsync(ctx->vp[0].p, vp, sizeof(float) * 3 * n, hipMemcpyHostToDevice, ctx->s));
    if (pid) CUBE_HIP(hipMemcpyAsync(ctx->pid[0].p, pid, sizeof(int64_t) * n, hipMemcpyHostToDevice, ctx->s));
    ctx->layout = 1;
    ctx->np = n;
    ctx->off_valid = false;
    ctx->ensure_offsets();
    {
      KScope k(&ctx->tm, "buffer_v", (double)n * 12);
      buffer_v(g, ctx->off.p, ctx->vp[0].p, pid ? ctx->pid[0].p : nullptr, ctx->s);
    }
    CUBE_HIP(hipMemcpyAsync(vp, ctx->vp[0].p, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, ctx->s));
    if (pid) CUBE_HIP(hipMemcpyAsync(pid, ctx->pid[0].p, sizeof(int64_t) * n, hipMemcpyDeviceToHost, ctx->s));
    CUBE_HIP(hipStreamSynchronize(ctx->s));
    ctx->layout = 0;  // positions were not uploaded
    ctx->vmax_valid = false;
  });
}

int cube_p3m_buffer_v_resident(cube_p3m_ctx* ctx) {
  return guarded(ctx, [&] { ctx->buffer_vel(); });
}

int cube_p3m_state_upload(cube_p3m_ctx* ctx, const void* xp, const float* vp, const int64_t* pid,
                          const int32_t* rhoc, int64_t np) {
  return guarded(ctx, [&] {
    if (!xp || !vp || !rhoc) throw Error(CUBE_ERR_ARG, "null argument");
    upload_ext(ctx, xp, vp, pid, rhoc, np);
    CUBE_HIP(hipStreamSynchronize(ctx->s));
  });
}

int cube_p3m_state_download(cube_p3m_ctx* ctx, void* xp, float* vp, int64_t* pid, int32_t* rhoc, int64_t np_cap,
                            int64_t* np) {
  return guarded(ctx, [&] {
    if (ctx->layout == 0) throw Error(CUBE_ERR_STATE, "no resident state");
    const int64_t n = ctx->np;
    if (np) *np = n;
    if (n > np_cap) throw Error(CUBE_ERR_ARG, "download capacity too small");
    if (xp) CUBE_HIP(hipMemcpyAsync(xp, ctx->xpc(), (size_t)n * 3 * ctx->esz(), hipMemcpyDeviceToHost, ctx->s));
    if (vp) CUBE_HIP(hipMemcpyAsync(vp, ctx->vpc(), sizeof(float) * 3 * n, hipMemcpyDeviceToHost, ctx->s));
    if (pid && ctx->has_pid)
      CUBE_HIP(hipMemcpyAsync(pid, ctx->pidc(), sizeof(int64_t) * n, hipMemcpyDeviceToHost, ctx->s));
    if (rhoc)
      CUBE_HIP(hipMemcpyAsync(rhoc, ctx->rhocc(), sizeof(int32_t) * ctx->g.ncells_ext, hipMemcpyDeviceToHost, ctx->s));
    CUBE_HIP(hipStreamSynchronize(ctx->s));
  });
}

int cube_p3m_kick_resident(cube_p3m_ctx* ctx, const cube_kick_in* in, cube_kick_out* out) {
  return guarded(ctx, [&] {
    if (!in) throw Error(CUBE_ERR_ARG, "null argument");
    cube_kick_out o;
    if (ctx->prec == 32) ctx->kick<float>(*in, o, nullptr);
    else ctx->kick<double>(*in, o, nullptr);
    if (out) *out = o;
  });
}

int cube_p3m_kick_resident_nu(cube_p3m_ctx* ctx, cube_p3m_ctx* nu, const cube_kick_in* in, float nu_mass,
                              cube_kick_out* out, cube_kick_out* out_nu) {
  return guarded(ctx, [&] {
    if (!in || !nu) throw Error(CUBE_ERR_ARG, "null argument");
    if (nu == ctx) throw Error(CUBE_ERR_ARG, "the second species needs a context of its own");
    const cube_p3m_config &a = ctx->cfg, &b = nu->cfg;
    if (nu->device != ctx->device || a.ng != b.ng || a.nnt != b.nnt || a.nns != b.nns || a.ratio_cs != b.ratio_cs ||
        a.ncb != b.ncb || a.ngb != b.ngb || a.izipx != b.izipx || a.precision != b.precision)
      throw Error(CUBE_ERR_ARG, "the two species' contexts differ in device, geometry, coordinates or precision");
    if (nu->layout != 1) throw Error(CUBE_ERR_STATE, "the second species needs the buffered (extended) layout");
    if (!(nu_mass >= 0.0f)) throw Error(CUBE_ERR_ARG, "nu_mass must be >= 0");
    // the second species' arrays were last written on its own stream
    CUBE_HIP(hipSetDevice(nu->device));
    nu->ensure_offsets();
    CUBE_HIP(hipStreamSynchronize(nu->s));
    struct Detach {
      cube_p3m_ctx* c;
      ~Detach() { c->nu = nullptr; }
    } detach{ctx};
    ctx->nu = nu;
    ctx->nu_mass = nu_mass;
    const int mask = in->phase_mask ? in->phase_mask : 15;
    cube_kick_out o, on;
    if (ctx->prec == 32) {
      ctx->kick<float>(*in, o, nullptr);
      nu_outputs<float>(ctx, nu, *in, mask, on);
    } else {
      ctx->kick<double>(*in, o, nullptr);
      nu_outputs<double>(ctx, nu, *in, mask, on);
    }
    if (out) *out = o;
    if (out_nu) *out_nu = on;
  });
}

int cube_p3m_add_thermal_velocities(cube_p3m_ctx* ctx, double v_scale, uint64_t seed, float* vz_max) {
  return guarded(ctx, [&] {
    if (ctx->layout == 0) throw Error(CUBE_ERR_STATE, "no resident particles");
    if (!ctx->has_pid) throw Error(CUBE_ERR_STATE, "thermal velocities are keyed by particle ID: no resident IDs");
    if (!(v_scale >= 0.0)) throw Error(CUBE_ERR_ARG, "v_scale must be >= 0");
    const float vz = add_thermal_velocities(ctx->vpc(), ctx->pidc(), ctx->np, v_scale, seed, ctx->s);
    ctx->vmax_valid = false;
    if (vz_max) *vz_max = vz;
  });
}

int cube_p3m_drift_resident(cube_p3m_ctx* ctx, const cube_drift_in* in, cube_drift_out* out) {
  return guarded(ctx, [&] {
    if (!in) throw Error(CUBE_ERR_ARG, "null argument");
    cube_drift_out o{};
    ctx->drift(*in, o);
    if (out) *out = o;
  });
}

int cube_p3m_buffer_resident(cube_p3m_ctx* ctx, int64_t* np_ext) {
  return guarded(ctx, [&] { ctx->buffer(np_ext); });
}

int cube_p3m_step_resident(cube_p3m_ctx* ctx, const cube_drift_in* din, const cube_kick_in* kin,
                           cube_drift_out* dout, cube_kick_out* kout) {
  return guarded(ctx, [&] {
    if (!din || !kin) throw Error(CUBE_ERR_ARG, "null argument");
    cube_drift_out d{};
    ctx->drift(*din, d);
    int64_t n = 0;
    ctx->buffer(&n);
    cube_kick_out k;
    if (ctx->prec == 32) ctx->kick<float>(*kin, k, nullptr, true);  // buffer_v rides on the kick
    else ctx->kick<double>(*kin, k, nullptr, true);
    if (dout) *dout = d;
    if (kout) *kout = k;
  });
}

int cube_p3m_make_ic(cube_p3m_ctx* ctx, const cube_ic_params* ic, float* vz_max, int64_t* nplocal) {
  return guarded(ctx, [&] {
    if (!ic) throw Error(CUBE_ERR_ARG, "null argument");
    const Geo& g = ctx->g;
    ctx->ensure_set(ctx->cur, std::max<int64_t>(g.np_image, 1));
    ctx->has_pid = true;
    ICResult r = make_ic(g, *ic, ctx->xp[ctx->cur].p, ctx->vp[ctx->cur].p, ctx->pid[ctx->cur].p,
                         ctx->rhoc[ctx->rcur].p, ctx->s);
    ctx->layout = 2;
    ctx->np = r.nplocal;
    ctx->off_valid = false;
    ctx->vmax_valid = false;
    int64_t n = 0;
    ctx->buffer(&n);
    if (vz_max) *vz_max = r.vz_max;
    if (nplocal) *nplocal = r.nplocal;
  });
}

int cube_p3m_cicpower(cube_p3m_ctx* ctx, int32_t nw, float* xi, int32_t nbin, float* delta, double* mean) {
  return guarded(ctx, [&] {
    if (!xi) throw Error(CUBE_ERR_ARG, "null argument");
    if (ctx->layout == 0) throw Error(CUBE_ERR_STATE, "no resident particles");
    const Geo& g = ctx->g;
    ctx->ensure_offsets();
    // physical particle count (n_particle of auto_power)
    std::vector<int32_t> rc((size_t)g.ncells_ext);
    CUBE_HIP(hipMemcpyAsync(rc.data(), ctx->rhocc(), sizeof(int32_t) * rc.size(), hipMemcpyDeviceToHost, ctx->s));
    CUBE_HIP(hipStreamSynchronize(ctx->s));
    int64_t npart = 0;
    for (int64_t T = 0; T < g.ntiles; ++T)
      for (int64_t k = 1; k <= g.nt; ++k)
        for (int64_t j = 1; j <= g.nt; ++j)
          for (int64_t i = 1; i <= g.nt; ++i) npart += rc[rc_index(i, j, k, T, g.nte, g.ncb)];
    if (npart == 0) throw Error(CUBE_ERR_STATE, "no physical particles");
    KScope k(&ctx->tm, "cicpower", 0);
    cic_power(g, nw, ctx->cfg.box, ctx->xpc(), ctx->rhocc(), ctx->off.p, npart, xi, nbin, delta, mean, ctx->s);
  });
}

int cube_p3m_synchronize(cube_p3m_ctx* ctx) {
  return guarded(ctx, [&] { CUBE_HIP(hipStreamSynchronize(ctx->s)); });
}

int cube_p3m_graph_enable(cube_p3m_ctx* ctx, int32_t enable) {
  return guarded(ctx, [&] {
    ctx->graph_on = enable != 0;
    if (!ctx->graph_on) ctx->graph_drop();
  });
}

int cube_p3m_timing_enable(cube_p3m_ctx* ctx, int32_t enable) {
  return guarded(ctx, [&] { ctx->tm.enabled = enable != 0; });
}

int cube_p3m_timing_reset(cube_p3m_ctx* ctx) {
  return guarded(ctx, [&] { ctx->tm.reset(); });
}

int cube_p3m_timing_report(cube_p3m_ctx* ctx, char* buf, int64_t len) {
  return guarded(ctx, [&] {
    std::string r = ctx->tm.report();
    char extra[128];
    snprintf(extra, sizeof(extra), "pp_pairs_last %lld 0 0 0\n", (long long)ctx->last_pairs);
    r += extra;
    snprintf(extra, sizeof(extra), "graph_captures %lld 0 0 0\ngraph_launches %lld 0 0 0\n",
             (long long)ctx->graph_captures, (long long)ctx->graph_launches);
    r += extra;
    for (const auto& kv : ctx->tune_picks)  // picks read from CUBE_AUTOTUNE_FILE, not timed here
      if (!ctx->tuned.count(kv.first)) {
        snprintf(extra, sizeof(extra), "autotune:%s:%d 1 0 0 0\n", kv.first.c_str(), kv.second);
        r += extra;
      }
    for (const auto& kv : ctx->tuned) {
      // "autotune:<key> <pick> <ms of the pick> 0 0", then one line per candidate
      const auto& t = kv.second;
      for (size_t ci = 0; ci < t.cand.size(); ++ci) {
        snprintf(extra, sizeof(extra), "autotune:%s:%d %d %.4f 0 0\n", kv.first.c_str(), t.cand[ci],
                 t.cand[ci] == t.pick ? 1 : 0, t.ms[ci]);
        r += extra;
      }
    }
    if (buf && len > 0) {
      std::strncpy(buf, r.c_str(), (size_t)len - 1);
      buf[len - 1] = 0;
    }
  });
}

}  // extern "C"
