// Internal definitions shared by the libhipfdtd translation units.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "hipfdtd.h"

namespace hfdtd {

// Thread-local last-error text behind hfdtd_last_error().
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
int hip_fail(hipError_t err, const char *what);

#define HF_HIP(call)                                                                         \
  do {                                                                                       \
    hipError_t _e = (call);                                                                  \
    if (_e != hipSuccess) return ::hfdtd::hip_fail(_e, #call);                               \
  } while (0)

// Physical constants (constants.py:11-14 restated).
constexpr double kC0 = 299792458.0;
constexpr double kMU0 = 4e-7 * 3.14159265358979323846;
constexpr double kEPS0 = 1.0 / (kMU0 * (kC0 * kC0));

// psi stripe slots, in the reference's naming psi_<field>_<axis>
// (engine.py:305-318).  Slots 0..5 belong to the H update, 6..11 to E.
enum PsiSlot {
  PSI_HY_X = 0, PSI_HZ_X, PSI_HX_Y, PSI_HZ_Y, PSI_HX_Z, PSI_HY_Z,
  PSI_EY_X = 6, PSI_EZ_X, PSI_EX_Y, PSI_EZ_Y, PSI_EX_Z, PSI_EY_Z,
  PSI_COUNT = 12
};

// Axis of the stripe each slot lives in (0 = x, 1 = y, 2 = z).
constexpr int kPsiAxis[PSI_COUNT] = {0, 0, 1, 1, 2, 2, 0, 0, 1, 1, 2, 2};

constexpr int kMarchSources = 4;  // sources the march sweep applies inside the sweep
constexpr int kMatMax = 48;  // materials of an indexed grid at most (march sweep)

// Where the march sweep takes an edge's ce from on lossless FP32 grids: the
// three ce volumes; the material index (a byte per cell and a table of ce
// per pair of materials); or a volume of eps_r (an FP32 value per cell,
// exact copies of the values the coefficients were made from), ce
// recomputed per edge in FP64 and rounded, the arithmetic of the
// coefficient kernel, so the values are those of the volumes.
constexpr int kCeVolumes = 0, kCeIndex = 1, kCeEps = 2;

// The z-face record of the march sweep.  A Huygens-box monitor whose
// z faces lie at k = kf[0] and kf[1] (each at least 1, two or more apart)
// asks the engine to keep, for every i in [i0, i1] and j in [j0, j1] (its
// face cells and their +1 neighbours), the values its co-location reads
// there: (Ex, Ey) at k and (Hx, Hy) at k and k - 1.  March steps write
// (Ex, Ey, Hx, Hy) of both columns as float quads into buf,
// [face][column: k - 1 | k][i - i0][j - j0]; the monitor's gather then reads
// them contiguously instead of one cache line per sample and field.  The
// record matches the state while `epoch` equals the engine's state_epoch,
// which every change of the state advances.  buf holds two records of
// `bytes` each, one per slab parity: a sweep that writes state[p] writes
// record p, so the gather of the previous step may still read the other.
struct ZRec {
  const void *owner = nullptr;  // the monitor it serves (nullptr: off)
  int kf[2] = {};
  int i0 = 0, i1 = -1, j0 = 0, j1 = -1;
  float *buf = nullptr;
  size_t bytes = 0;  // of one record
  uint64_t epoch = ~0ull;
  // quads per (face, column) array
  size_t quads() const { return (size_t)(i1 - i0 + 1) * (size_t)(j1 - j0 + 1); }
  float *half(int parity) const {
    return buf ? buf + (size_t)parity * (bytes / sizeof(float)) : nullptr;
  }
};

struct GatherTail;  // n2f_gather.h

struct Engine {
  int nx = 0, ny = 0, nz = 0, npml = 0;
  int fp64 = 0;
  int device = 0;
  double dl = 0, dt = 0;
  double dtm_dl = 0, dtm = 0, inv_dl = 0;  // folded on the host in FP64
  size_t ncell = 0;
  size_t vstride = 0;  // elements between volumes of a slab (ncell + 64, rounded up to 64)
  // elements between j rows of the volumes the march sweep addresses: nz, or
  // the padded row length during a row-padded run
  int sk = 0;
  size_t elem = 4;     // bytes per value
  hipStream_t stream = nullptr;

  // Slabs.  state[cur] holds Ex Ey Ez Hx Hy Hz (each vstride long); the split
  // sweeps update it in place, the fused sweep reads state[cur], writes
  // state[cur ^ 1] and flips cur.  psi[] follows the same parity.
  void *state[2] = {};
  void *coef = nullptr;  // ce_x ce_y ce_z ca_x ca_y ca_z
  void *psi[2] = {};     // 12 face stripes, slot order of PsiSlot
  size_t psi_off[PSI_COUNT + 1] = {};  // element offset of each slot
  void *prof = nullptr;  // [stagger 0=E,1=H][b, c, ik][x | y | z]
  int cur = 0;

  bool lossless = true;  // ca == 1 everywhere (set_epsilon without sigma)
  int tile_rows = 8;     // fused sweep: j rows per block
  int tile_planes = 0;   // fused sweep: i planes per block (0 = chosen from the grid)
  int xcd = 8;           // XCDs for the work-order remap (1 = hardware order)
  int split_run = 128;   // split sweeps: consecutive blocks per XCD run (1 = round robin)
  int cus = 256;         // compute units of the device
  // march sweep (register-pipelined fused sweep): rows per wave (1 or 2;
  // 4 rows per wave-row group), i planes per block (0 = chosen from the
  // grid), consecutive blocks per XCD run, non-temporal streams (-1 = when
  // the slabs exceed the last-level cache, 0 = never, 1 = always)
  int march_rows = 0;  // 0: one row on material-indexed grids, else two
  // pair sweep (two steps per launch): i planes per block; the watched
  // boxes whose state between the two steps goes to the shadow slab, and
  // whether field reads see that slab
  int pair_planes = 24;
  int watch_n = 0;
  int watch_b[8][6] = {};
  int watch_c[8] = {};  // component masks
  void *shadow = nullptr;
  bool view_shadow = false;
  int march_planes = 0;
  int march_run = 16;
  bool march_per_volume = false;  // per-volume addressing on any FP32 slab (set by hand)
  int march_nt = -1;
  // march tiles of 64 or 32 columns of k (FP32; 64 / W rows of j per
  // wavefront slot)
  int march_width = 64;
  // resident march blocks per CU by [width 64/32][lossy][rows], and the
  // same for the index / eps_r sweeps by [width][ce source][rows]
  int march_occ[2][2][3] = {};
  int march_occ_mat[2][3][3] = {};
  // Lossless FP32 grids with at most kMatMax distinct eps_r values: a byte of
  // material per cell and the table of ce for every pair of materials (an
  // edge's ce depends only on its two cells), which the march sweep reads
  // instead of the three ce volumes.  mat_ok while they match the ce volumes.
  void *mat = nullptr;
  float *mtab = nullptr;
  int nmat = 0;
  bool mat_ok = false;
  // Other lossless FP32 grids whose eps_r values are exact in FP32: the
  // eps_r volume (ncell + 64 floats), from which the march sweep recomputes
  // ce.  eps_ok while it matches the ce volumes.
  void *epsv = nullptr;
  bool eps_ok = false;
  int ce_source() const {
    return !lossless || fp64 ? kCeVolumes : mat_ok ? kCeIndex : eps_ok ? kCeEps : kCeVolumes;
  }

  uint64_t bytes = 0;

  // Row-padded run (hfdtd_padded_begin / hfdtd_padded_end): the march sweep
  // steps a copy of the state and psi (and of the ce source it reads) whose
  // rows are padded to whole 128-byte lines; state[], psi[], vstride and sk
  // point at the copy meanwhile, and the engine's own slabs are kept here.
  struct PaddedRun {
    bool live = false;
    void *buf = nullptr;
    size_t bytes = 0;
    void *state[2] = {}, *psi[2] = {};
    void *coef = nullptr, *mat = nullptr, *epsv = nullptr;
    size_t vstride = 0;
    struct PadJob *d_jobs = nullptr;
    int jobs_cap = 0;
  } prun;

  // Advanced by every call that changes E, H or psi (or what field() reads);
  // a march step that writes the z-face record stamps it with the new value.
  uint64_t state_epoch = 0;
  ZRec zrec;

  // A K-buffered near-to-far monitor's gather of step n, held back so that
  // the march sweep of step n + 1 runs it in blocks after its own (both read
  // state[p] only; the sweep writes the other slab and the other record).
  // Every other call on the engine first runs a held gather on its own
  // (HF_SETTLE).  pgather is owned by n2f.hip.
  bool fuse_gather = true;
  GatherTail *pgather = nullptr;
  void touch() { ++state_epoch; }
  bool zrec_valid(const void *owner) const {
    return zrec.owner == owner && zrec.buf && zrec.epoch == state_epoch && !view_shadow;
  }

  // HIP-event timing of the Yee sweeps (hfdtd_timing / hfdtd_timing_read):
  // while enabled, every timing_every-th sweep launch records a start and a
  // stop event around its kernel (timing_events hands them out).
  bool timing = false;
  int timing_every = 1;
  long long timing_launches = 0;  // sweep launches since timing was enabled
  std::vector<hipEvent_t> events;  // pool; [2q, 2q+1] bracket timed launch q
  size_t events_used = 0;

  char *vol(int slab_parity, int c) const {
    return (char *)state[slab_parity] + (size_t)c * vstride * elem;
  }
  void *field(int c) const {
    if (view_shadow && shadow) return (char *)shadow + (size_t)c * vstride * elem;
    return vol(cur, c);
  }
  void *coef_vol(int c) const { return (char *)coef + (size_t)c * vstride * elem; }
  // The fused sweeps address slabs with 32-bit byte offsets.
  bool fused_fits() const {
    const double lim = 4294967296.0;
    return 6.0 * (double)vstride * elem < lim && ((double)psi_off[PSI_COUNT] + 64.0) * elem < lim;
  }
  size_t prof_len() const { return (size_t)nx + ny + nz; }
  // Resident fused blocks per CU, by rows (index) and lossy (0/1); 0 = not
  // yet queried.
  int fused_occ[2][13] = {};
  // i planes per fused block: the set value, or the count P in [8, 64] that
  // minimises the modelled sweep time (waves of resident blocks) x (P + 2):
  // each wave lasts as long as its blocks' P planes plus the two halo planes
  // every block re-reads; slots = resident blocks per CU x CUs.  Ties go to
  // more planes per block.
  // i planes per march block (even): the set value, or the even P in
  // [8, 64] minimising (waves of resident blocks) x (P + 3), the 3 for the
  // plane before the chunk and its prologue (measured at 256^3: short
  // chunks cost more than one plane each).  With PML the x-stripes form
  // chunks of their own and P chunks the planes between.
  int march_planes_for(int rows_per_wave, int per_cu, int width = 64) const {
    if (march_planes > 0) return (march_planes + 1) / 2 * 2;
    const long long slots = (long long)(per_cu > 0 ? per_cu : 1) * cus;
    const int tj = 4 * rows_per_wave * (64 / width);
    const long long tiles = (long long)((nz + width - 1) / width) * ((ny + tj - 1) / tj);
    int best = 64;
    long long best_t = -1;
    const bool split = npml > 0 && nx > 2 * npml;  // x-stripes in chunks of their own
    const int inner = split ? nx - 2 * npml : nx;
    for (int P = 64; P >= 8; P -= 2) {
      const long long blocks = tiles * ((inner + P - 1) / P + (split ? 2 : 0));
      const long long t = (blocks + slots - 1) / slots * (P + 3);
      if (best_t < 0 || t < best_t) {
        best_t = t;
        best = P;
      }
    }
    return best;
  }
  // Non-temporal march streams pay off once the slabs no longer fit the
  // device's last-level cache (256 MB on MI355X): the stores then stop
  // evicting the halo rows other blocks still need.
  bool march_streaming() const {
    if (march_nt >= 0) return march_nt > 0;
    const double slabs = (12.0 + (lossless ? 3.0 : 6.0)) * (double)vstride * elem;
    return slabs > 256.0 * 1024 * 1024;
  }
  int fused_planes(int rows, int per_cu) const {
    if (tile_planes > 0) return tile_planes;
    const long long slots = (long long)(per_cu > 0 ? per_cu : 1) * cus;
    const long long tiles = (long long)((nz + 63) / 64) * ((ny + rows - 1) / rows);
    int best = 64;
    long long best_t = -1;
    for (int P = 64; P >= 8; --P) {
      const long long blocks = tiles * ((nx + P - 1) / P);
      const long long t = (blocks + slots - 1) / slots * (P + 2);
      if (best_t < 0 || t < best_t) {
        best_t = t;
        best = P;
      }
    }
    return best;
  }
};

// The start / stop events of the next sweep launch when timing samples it
// (both null otherwise); the launch itself records them
// (hipExtLaunchKernelGGL), around the kernel's execution.
int timing_events(Engine *e, hipEvent_t &start, hipEvent_t &stop);

// Launch the split H / E sweeps and the source kernel (yee.hip).
int launch_update_H(Engine *e);
int launch_update_E(Engine *e);
int launch_sources(Engine *e, const hfdtd_source *src, int n);
int launch_fused_step(Engine *e, const hfdtd_source *src, int n);
int launch_march_step(Engine *e, const hfdtd_source *src, int n);
int launch_pair_steps(Engine *e, const hfdtd_source *sa, int na, const hfdtd_source *sb, int nb);
// ensure the ping-pong partners exist (engine.hip)
int ensure_alt(Engine *e);

// One row copy between two layouts of a volume: rows x w floats from src
// (row pitch sp) to dst (row pitch dp).
struct PadJob {
  const float *src;
  float *dst;
  long long rows;
  int w, sp, dp;
};
// Run row copies on a stream (batch.hip); d_jobs / cap hold the device list.
int run_pad_jobs(hipStream_t stream, const std::vector<PadJob> &jobs, PadJob *&d_jobs, int &cap);
// End a row-padded run: the copy's state and psi back into the engine's own
// slabs (engine.hip); a no-op when none is live.
int end_padded_run(Engine *e);
// The held gather (Engine::pgather; n2f.hip): flush_gather runs it on the
// engine's stream if one is held; free_gather drops the holder.
int flush_gather(Engine *e);
void free_gather(Engine *e);
// Before any entry point other than a march step and the per-step monitor
// kernels: a held gather run, then the row-padded run ended.
int settle(Engine *e);
#define HF_SETTLE(e)                             \
  do {                                           \
    const int _src = ::hfdtd::settle(e);         \
    if (_src) return _src;                       \
  } while (0)

inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace hfdtd

struct hfdtd_engine {
  hfdtd::Engine s;
};
