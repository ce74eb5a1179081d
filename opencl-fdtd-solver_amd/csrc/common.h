// Internal definitions shared by the libhipfdtd translation units.
#pragma once

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

struct Engine {
  int nx = 0, ny = 0, nz = 0, npml = 0;
  int fp64 = 0;
  int device = 0;
  double dl = 0, dt = 0;
  double dtm_dl = 0, dtm = 0, inv_dl = 0;  // folded on the host in FP64
  size_t ncell = 0;
  size_t vstride = 0;  // elements between volumes of a slab (ncell + 64, rounded up to 64)
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

  uint64_t bytes = 0;

  // HIP-event timing of the Yee sweeps (hfdtd_timing / hfdtd_timing_read):
  // events are recorded on `stream` around every sweep launch while enabled.
  bool timing = false;
  std::vector<hipEvent_t> events;  // pool; [2q, 2q+1] bracket launch q
  size_t events_used = 0;

  char *vol(int slab_parity, int c) const {
    return (char *)state[slab_parity] + (size_t)c * vstride * elem;
  }
  void *field(int c) const { return vol(cur, c); }
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

// Record the next timing event on the engine stream (no-op when disabled).
int timing_mark(Engine *e);

// Launch the split H / E sweeps and the source kernel (yee.hip).
int launch_update_H(Engine *e);
int launch_update_E(Engine *e);
int launch_sources(Engine *e, const hfdtd_source *src, int n);
int launch_fused_step(Engine *e, const hfdtd_source *src, int n);
// ensure the ping-pong partners exist (engine.hip)
int ensure_alt(Engine *e);

inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace hfdtd

struct hfdtd_engine {
  hfdtd::Engine s;
};
