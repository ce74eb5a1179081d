// Point probe and the batched sweep of many independent simulations.
//
// The batched sweeps advance a set of engines of uneven sizes together: a
// flat work list holds every (engine, tile) pair of the path (fused tiles,
// 256-cell blocks of the split sweeps' flattened planes, or march tiles in
// batch_march.h).  The step index lives in device memory: launch q of a run
// reads *step + q, and a one-thread kernel adds the run's length after it,
// so a run of K steps is captured once as a HIP graph and replayed.  Source
// amplitudes and probe phases come from per-engine tables indexed by that
// step, which keeps every launch argument fixed across replays.

#include <algorithm>
#include <vector>

#include "batch_march.h"
#include "fdtd_core.h"
#include "launch_common.h"

struct hfdtd_probe {
  hfdtd::Engine *e = nullptr;
  int i = 0, j = 0, k = 0, comp = 0, nfreq = 0;
  double *acc = nullptr;  // nfreq x {re, im}, device
};

namespace hfdtd {

constexpr int kProbeFreqPerLaunch = 16;

struct ProbePhase {
  double v[kProbeFreqPerLaunch][2];
};

template <typename T>
__global__ void k_probe_accumulate(const T *__restrict__ F, size_t idx, int f0, int nf,
                                   ProbePhase ph, double *__restrict__ acc) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double v = (double)F[idx];
  for (int q = 0; q < nf; ++q) {
    acc[2 * (f0 + q)] += v * ph.v[q][0];
    acc[2 * (f0 + q) + 1] += v * ph.v[q][1];
  }
}

// ── batch ─────────────────────────────────────────────────────────────────

template <typename T>
struct SimDesc {
  Geo g;
  T *st[2];   // state slab, by step parity
  T *psi[2];  // psi slab, by step parity
  const T *coef;
  const T *prof;
  // lossless FP32 engines with a material index (split path): one byte of
  // material per cell and the ce table; null otherwise
  const unsigned char *mat;
  const float *mtab;
  int nmat;
  int tile_begin, tiles_k, tiles_j;
  int lossy;
  int has_src;
  Src<T> src;
  int has_probe, pi, pj, pk;
};

template <typename T>
struct BatchArgs {
  const SimDesc<T> *sims;
  const int *tile_sim;
  int n_tiles;
  const int *step;  // steps taken since the tables were set
  int step0;        // value of *step when `sims` was built (parity origin)
  int planes;
  T dtm_dl, dtm, inv_dl;
  const double *amp;  // n x n_table
  const double *phase;  // n x n_table x nfreq x 2
  int n_table;
  int nfreq, comp;
  double *probe_acc;  // n x nfreq x 2
  int xcd;            // XCD count for the work-order remap (1 = off)
  int run;            // split path: consecutive items per XCD run
  int n_sims;
  // split path: runs of kFlatBlock cells per item, and the items whose
  // plane holds their engine's source or probe (their tile_sim entry also
  // carries kHookBit)
  int runs;
  const int *hook_items;
  int n_hook;
};

// A launch never writes its work list or descriptors: reading them through
// the constant address space lets them come in by scalar loads into SGPRs
// (read through a generic pointer, part of them come in by vector loads
// after the first stores, each wavefront's copy in VGPRs).
#define HF_CONST __attribute__((address_space(4)))
template <typename S>
__device__ __forceinline__ const HF_CONST S *as_const(const S *p) {
  return (const HF_CONST S *)p;
}
template <typename S>
__device__ __forceinline__ S load_const(const HF_CONST S *p) {
  static_assert(sizeof(S) % 4 == 0, "read in dwords");
  S out;
  const HF_CONST unsigned *q = (const HF_CONST unsigned *)p;
  unsigned *o = reinterpret_cast<unsigned *>(&out);
#pragma unroll
  for (int w = 0; w < (int)(sizeof(S) / 4); ++w) o[w] = q[w];
  return out;
}

template <typename T>
struct BatchSrcP {
  const Src<T> *s;
  T a;
  __device__ int count() const { return 1; }
  __device__ const Src<T> &at(int) const { return *s; }
  __device__ T amp(int) const { return a; }
};

// LOSSY selects the update with the ca load; it is exact for lossless
// engines too (ca = 1), so one variant serves a mixed batch.
template <typename T, int TJ, bool LOSSY>
__global__ __launch_bounds__(64 * (TJ + 1)) void k_batch_step(BatchArgs<T> a, int q) {
  const int par = (*a.step + q - a.step0) & 1;
  for (int n = blockIdx.x; n < a.n_tiles; n += gridDim.x) {
    const int t = xcd_remap(n, a.n_tiles, a.xcd);
    const int s = a.tile_sim[t];
    const SimDesc<T> &d = a.sims[s];
    const Geo g = d.g;
    const int lt = t - d.tile_begin;
    const int bx = lt % d.tiles_k;
    const int rest = lt / d.tiles_k;
    const int by = rest % d.tiles_j;
    const int bz = rest / d.tiles_j;
    lds_barrier();  // LDS of the previous work item is free
    fused_tile<T, TJ, LOSSY>(g, a.planes, a.dtm_dl, a.dtm, a.inv_dl, d.st[par], d.st[par ^ 1],
                             d.coef, d.prof, d.psi[par], d.psi[par ^ 1], bx, by, bz);
  }
}

// After the batched sweep: one block per engine redoes E on the cells of its
// source box (source_fixup_cell), then records its probe cell.
template <typename T>
__global__ __launch_bounds__(256) void k_batch_fixup(BatchArgs<T> a, int q) {
  const int step = *a.step + q;
  const int par = (step - a.step0) & 1;
  const int s = blockIdx.x;
  const SimDesc<T> &d = a.sims[s];
  const Geo g = d.g;
  const size_t row = (size_t)s * a.n_table + step;
  const T *s0 = d.st[par];
  T *s1 = d.st[par ^ 1];
  if (d.has_src) {
    const BatchSrcP<T> sp{&d.src, T(a.amp[row])};
    const int ni = d.src.ci1 - d.src.ci0, nj = d.src.cj1 - d.src.cj0;
    const int nk = d.src.ck1 - d.src.ck0;
    const long long total = (long long)ni * nj * nk;
    for (long long t = threadIdx.x; t < total; t += blockDim.x) {
      const int k = d.src.ck0 + (int)(t % nk);
      const int j = d.src.cj0 + (int)((t / nk) % nj);
      const int i = d.src.ci0 + (int)(t / ((long long)nk * nj));
      if (d.lossy)
        source_fixup_cell<T, true>(g, a.inv_dl, s0, s1, d.coef, d.prof, d.psi[par],
                                   d.psi[par ^ 1], sp, i, j, k);
      else
        source_fixup_cell<T, false>(g, a.inv_dl, s0, s1, d.coef, d.prof, d.psi[par],
                                    d.psi[par ^ 1], sp, i, j, k);
    }
  }
  if (!d.has_probe) return;
  __syncthreads();  // the block's E writes are visible to thread 0
  if (threadIdx.x != 0) return;
  const size_t idx = ((size_t)d.pi * g.ny + d.pj) * g.nz + d.pk;
  const double v = (double)s1[a.comp * g.vs + idx];
  const double *ph = a.phase + row * a.nfreq * 2;
  double *acc = a.probe_acc + (size_t)s * a.nfreq * 2;
  for (int f = 0; f < a.nfreq; ++f) {
    acc[2 * f] += v * ph[2 * f];
    acc[2 * f + 1] += v * ph[2 * f + 1];
  }
}

// ── batched split sweeps ──────────────────────────────────────────────────
//
// A work item is `runs` (1 to kSplitRunsMax, a launch setting) runs of
// kFlatBlock consecutive cells of one i-plane.  A block looks its engine up
// once and walks the runs, so the work-list lookup and descriptor loads that
// precede its first field load are paid once per item (with one run per
// item the batched H and E sweeps took 7 % and 15 % longer than the engine's
// own on a 160^3 grid, with four 0 % and 9 %); longer items leave more
// threads idle in the last item of each plane.
constexpr int kSplitRunsMax = 8;
// split path: tile_sim entries of items left to the hook pass
constexpr int kHookBit = 0x40000000;
//
// Work item n is a block of 256 consecutive cells of one i-plane of one
// engine (tiles_k holds the blocks per plane), so engines of any extents
// pack the wavefronts without idle lanes.  With one block per item the items
// are remapped so every XCD runs runs of consecutive items (neighbour rows
// share that XCD's L2).  The state is updated in place, in
// the reference's order: H everywhere, sources, E everywhere, probes.

template <typename T>
__global__ __launch_bounds__(kFlatBlock) void k_batch_split_H(BatchArgs<T> a) {
  for (int m = blockIdx.x; m < a.n_tiles; m += gridDim.x) {
    const int n = gridDim.x == a.n_tiles ? xcd_runs(m, a.n_tiles, a.xcd, a.run) : m;
    const int ts = as_const(a.tile_sim)[n];
    const HF_CONST SimDesc<T> *d = as_const(a.sims) + (ts & ~kHookBit);
    const Geo g = load_const(&d->g);
    T *const st = d->st[0];
    const T *const prof = d->prof;
    T *const psi = d->psi[0];
    const int tiles_k = d->tiles_k;
    const int lt = n - d->tile_begin;
    const int i = lt / tiles_k;
    const int q0 = (lt - i * tiles_k) * (kFlatBlock * a.runs) + threadIdx.x;
    const int ncell = g.ny * g.nz;
    for (int r = 0; r < a.runs; ++r) {
      const int q = q0 + r * kFlatBlock;
      if (q >= ncell) break;
      const int j = q / g.nz;
      split_h_cell(g, a.dtm_dl, a.dtm, st, prof, psi, i, j, q - j * g.nz);
    }
  }
}

// The engine's source of the step (added to E before the cell's update, as
// the single-engine split path adds it between its sweeps: sources.cl:7-49
// via src_term, with the cell's ce) and its probe (value x phase added to the
// probe's FP64 DFT after the update, by the one thread that owns the cell).
template <typename T>
struct BatchEHook {
  bool on;   // the source applies to this item
  Src<T> s;
  T amp;
  int pi, pj, pk, comp, nf;  // pi < 0: no probe
  const double *ph;
  double *acc;
  __device__ __forceinline__ void pre(int i, int j, int k, T cx, T cy, T cz, T &ex, T &ey,
                                      T &ez) const {
    if (!on) return;
    if (i < s.ci0 || i >= s.ci1 || j < s.cj0 || j >= s.cj1 || k < s.ck0 || k >= s.ck1) return;
    const T v = src_term<T>(s, i, j, k, s.comp == 0 ? cx : s.comp == 1 ? cy : cz, amp);
    if (s.comp == 0) ex += v;
    else if (s.comp == 1) ey += v;
    else ez += v;
  }
  __device__ __forceinline__ void post(int i, int j, int k, T ex, T ey, T ez) const {
    if (i != pi || j != pj || k != pk) return;
    const double v = (double)(comp == 0 ? ex : comp == 1 ? ey : ez);
    for (int f = 0; f < nf; ++f) {
      acc[2 * f] += v * ph[2 * f];
      acc[2 * f + 1] += v * ph[2 * f + 1];
    }
  }
};

// E half of the batched split step.  A block walks the runs of
// its work item, as the H sweep does; items whose plane holds their engine's
// source or probe are left to k_batch_split_Ehook, so this kernel carries no
// hook code and keeps the plain update's registers.
template <typename T, bool LOSSY>
__global__ __launch_bounds__(kFlatBlock) void k_batch_split_E(BatchArgs<T> a) {
  for (int m = blockIdx.x; m < a.n_tiles; m += gridDim.x) {
    const int n = gridDim.x == a.n_tiles ? xcd_runs(m, a.n_tiles, a.xcd, a.run) : m;
    const int ts = as_const(a.tile_sim)[n];
    if (ts & kHookBit) continue;
    const HF_CONST SimDesc<T> *d = as_const(a.sims) + ts;
    const Geo g = load_const(&d->g);
    T *const st = d->st[0];
    const T *const coef = d->coef;
    const T *const prof = d->prof;
    T *const psi = d->psi[0];
    const int tiles_k = d->tiles_k;
    const int lt = n - d->tile_begin;
    const int i = lt / tiles_k;
    const int q0 = (lt - i * tiles_k) * (kFlatBlock * a.runs) + threadIdx.x;
    const int ncell = g.ny * g.nz;
    if constexpr (sizeof(T) == 4) {
      const unsigned char *const mat = d->mat;
      if (mat) {  // lossless (ca = 1), whichever variant the batch runs
        const float *const mtab = d->mtab;
        const int nmat = d->nmat;
        for (int r = 0; r < a.runs; ++r) {
          const int q = q0 + r * kFlatBlock;
          if (q >= ncell) break;
          const int j = q / g.nz, k = q - j * g.nz;
          split_e_cell_mat(g, a.inv_dl, st, coef, mat, mtab, nmat, prof, psi, i, j, k);
        }
        continue;
      }
    }
    for (int r = 0; r < a.runs; ++r) {
      const int q = q0 + r * kFlatBlock;
      if (q >= ncell) break;
      const int j = q / g.nz, k = q - j * g.nz;
      split_e_cell<T, LOSSY>(g, a.inv_dl, st, coef, prof, psi, i, j, k);
    }
  }
}

// E on the items whose plane holds their engine's source or probe (after the
// H sweep, beside k_batch_split_E, which skips them): one pass per block,
// the source added before the cell's update and the probe recorded after it.
template <typename T, bool LOSSY>
__global__ __launch_bounds__(kFlatBlock) void k_batch_split_Ehook(BatchArgs<T> a, int qs) {
  const int step = *a.step + qs;
  const int n = a.hook_items[blockIdx.x / a.runs];
  const int pass = blockIdx.x % a.runs;
  const int sim = __builtin_amdgcn_readfirstlane(a.tile_sim[n]) & ~kHookBit;
  const SimDesc<T> &d = a.sims[sim];
  const Geo g = d.g;
  const int lt = n - d.tile_begin;
  const int i = lt / d.tiles_k;
  const int q = (lt - i * d.tiles_k) * (kFlatBlock * a.runs) + pass * kFlatBlock + threadIdx.x;
  if (q >= g.ny * g.nz) return;
  const int j = q / g.nz, k = q - j * g.nz;
  const bool src_here = a.amp && d.has_src && i >= d.src.ci0 && i < d.src.ci1;
  const bool prb_here = a.probe_acc && d.has_probe && i == d.pi;
  const size_t row = (size_t)sim * a.n_table + step;
  BatchEHook<T> hk;
  hk.on = src_here;
  hk.s = d.src;
  hk.amp = src_here ? T(a.amp[row]) : T(0);
  hk.pi = prb_here ? d.pi : -1;
  hk.pj = d.pj;
  hk.pk = d.pk;
  hk.comp = a.comp;
  hk.nf = a.nfreq;
  hk.ph = a.phase + row * a.nfreq * 2;
  hk.acc = a.probe_acc + (size_t)sim * a.nfreq * 2;
  if constexpr (sizeof(T) == 4) {
    if (d.mat) {
      split_e_cell_mat(g, a.inv_dl, d.st[0], d.coef, d.mat, d.mtab, d.nmat, d.prof, d.psi[0], i, j,
                       k, hk);
      return;
    }
  }
  split_e_cell<T, LOSSY>(g, a.inv_dl, d.st[0], d.coef, d.prof, d.psi[0], i, j, k, hk);
}

// After a run of n batched launches (which read *step + q).
__global__ void k_step_add(int *step, int n) {
  if (threadIdx.x == 0 && blockIdx.x == 0) atomicAdd(step, n);
}

// One engine's part of the batched march's arena: its geometry there (row
// stride sk, volume stride vs) and the element offsets of its pieces.
struct ArenaPart {
  Geo g;
  long long st, psi, coef, prof, mat;  // mat in bytes
  long long slab, psi_len;             // elements of one state slab, one psi slab
};

// Job blockIdx.y, grid-stride over its elements.
__global__ __launch_bounds__(256) void k_pad_rows(const PadJob *__restrict__ jobs) {
  const PadJob J = jobs[blockIdx.y];
  const long long n = J.rows * J.w;
  if (n < 0x7fffffffll) {  // 32-bit index arithmetic
    const unsigned w = (unsigned)J.w, nn = (unsigned)n;
    for (unsigned t = blockIdx.x * 256u + threadIdx.x; t < nn; t += gridDim.x * 256u) {
      const unsigned r = t / w, c = t - r * w;
      J.dst[(size_t)r * J.dp + c] = J.src[(size_t)r * J.sp + c];
    }
    return;
  }
  for (long long t = blockIdx.x * 256ll + threadIdx.x; t < n; t += (long long)gridDim.x * 256) {
    const long long r = t / J.w;
    const int c = (int)(t - r * J.w);
    J.dst[r * J.dp + c] = J.src[r * J.sp + c];
  }
}

}  // namespace hfdtd

struct hfdtd_batch {
  std::vector<hfdtd::Engine *> eng;
  hipStream_t stream = nullptr;
  int fp64 = 0;
  void *d_desc = nullptr;
  int *d_tile_sim = nullptr;
  int n_tiles = 0;
  int *d_hook_items = nullptr;  // split path: items holding a source or probe
  int n_hook = 0;
  int *d_step = nullptr;
  int step_host = 0;  // mirror of *d_step
  std::vector<hfdtd_source> src;
  std::vector<int> has_src;
  double *d_amp = nullptr;
  int n_amp = 0;
  std::vector<int> probe;  // n x 3, -1 = none
  int nfreq = 0, comp = 0;
  double *d_phase = nullptr;
  int n_phase = 0;
  double *d_probe_acc = nullptr;
  int blocks = 0, rows = 4, planes = 16;
  int split_runs = 2;  // split path: runs of 256 cells per work item
  int path = HFDTD_PATH_FUSED;  // fused tiles, the flat split sweeps, or march tiles
  // march path: rows per wavefront (1 or 2), i planes per chunk, consecutive
  // work items per XCD run; the medium kind and loss of the kernel it runs
  int mrows = 1, mplanes = 32, mrun = 16, mwidth = 64;
  int mce = 0;
  bool mlossy = false;
  bool lossy = false;
  // march path: every engine runs on a copy in one batch-owned arena
  // (state and psi of both parities, coefficients, profiles, material
  // index), filled from the engine before a run and written back after it;
  // rows padded to whole 128-byte lines where nz does not fill them (mpad)
  int mpad = 1;
  void *arena = nullptr;
  size_t arena_bytes = 0;
  std::vector<hfdtd::ArenaPart> parts;  // per engine
  hfdtd::PadJob *d_jobs = nullptr;      // the row copies of one direction
  int n_jobs_cap = 0;
  hipGraphExec_t gexec = nullptr;
};

namespace hfdtd {

static int batch_free_graph(hfdtd_batch *b) {
  if (b->gexec) {
    HF_HIP(hipGraphExecDestroy(b->gexec));
    b->gexec = nullptr;
  }
  return HFDTD_OK;
}

// Build the per-engine descriptors (parity origin = current step) and the
// tile work list.
template <typename T>
static int batch_build(hfdtd_batch *b) {
  const int n = (int)b->eng.size();
  const int TJ = b->rows;
  const int P = b->planes;
  std::vector<SimDesc<T>> desc(n);
  std::vector<int> tile_sim;
  int begin = 0;
  b->lossy = false;
  for (int s = 0; s < n; ++s) {
    Engine *e = b->eng[s];
    b->lossy = b->lossy || !e->lossless;
    if (b->path == HFDTD_PATH_FUSED) {
      if (!e->fused_fits())
        return fail(HFDTD_EUNSUPPORTED, "batched engines need field and psi slabs below 4 GB");
      int rc = ensure_alt(e);
      if (rc) return rc;
    }
    SimDesc<T> &d = desc[s];
    d.g = geo_of(e);
    // parity 0 (relative to step0) is the engine's current state
    d.st[0] = (T *)e->state[e->cur];
    d.st[1] = (T *)e->state[e->cur ^ 1];
    d.psi[0] = (T *)e->psi[e->cur];
    d.psi[1] = (T *)e->psi[e->cur ^ 1];
    d.coef = (const T *)e->coef;
    d.prof = (const T *)e->prof;
    const bool mat = b->path == HFDTD_PATH_SPLIT && sizeof(T) == 4 && e->lossless && e->mat_ok;
    d.mat = mat ? (const unsigned char *)e->mat : nullptr;
    d.mtab = mat ? (const float *)e->mtab : nullptr;
    d.nmat = mat ? e->nmat : 0;
    d.tile_begin = begin;
    int nt;
    if (b->path == HFDTD_PATH_SPLIT) {  // items of runs x 256 cells per plane
      d.tiles_k = (int)ceil_div((size_t)e->ny * e->nz, (size_t)kFlatBlock * b->split_runs);
      d.tiles_j = 1;
      nt = d.tiles_k * e->nx;
    } else {
      d.tiles_k = (int)ceil_div(e->nz, 64);
      d.tiles_j = (int)ceil_div(e->ny, TJ);
      nt = d.tiles_k * d.tiles_j * (int)ceil_div(e->nx, P);
    }
    tile_sim.insert(tile_sim.end(), nt, s);
    begin += nt;
    d.lossy = e->lossless ? 0 : 1;
    d.has_src = b->has_src.empty() ? 0 : b->has_src[s];
    if (d.has_src) {
      Src<T> o;
      const hfdtd_source &in = b->src[s];
      o.comp = in.comp;
      o.mode = in.mode;
      o.taper = in.taper & 7;
      o.i0 = in.i0; o.i1 = in.i1; o.j0 = in.j0; o.j1 = in.j1; o.k0 = in.k0; o.k1 = in.k1;
      auto clip = [](int a, int m) { return a < 0 ? 0 : (a > m ? m : a); };
      o.ci0 = clip(in.i0, e->nx); o.ci1 = std::max(o.ci0, clip(in.i1, e->nx));
      o.cj0 = clip(in.j0, e->ny); o.cj1 = std::max(o.cj0, clip(in.j1, e->ny));
      o.ck0 = clip(in.k0, e->nz); o.ck1 = std::max(o.ck0, clip(in.k1, e->nz));
      o.amp = T(0);
      o.re = T(in.rim_edge);
      d.src = o;
    }
    d.has_probe = 0;
    d.pi = d.pj = d.pk = -1;
    if (!b->probe.empty() && b->probe[3 * s] >= 0) {
      d.has_probe = 1;
      d.pi = b->probe[3 * s];
      d.pj = b->probe[3 * s + 1];
      d.pk = b->probe[3 * s + 2];
    }
  }
  // split path: the items whose plane holds their engine's source (when the
  // amplitude table is set) or probe (when the probes are set)
  std::vector<int> hook_items;
  if (b->path == HFDTD_PATH_SPLIT) {
    for (int s = 0; s < n; ++s) {
      const SimDesc<T> &d = desc[s];
      auto mark = [&](int i) {
        for (int t = 0; t < d.tiles_k; ++t) {
          const int it = d.tile_begin + i * d.tiles_k + t;
          if (!(tile_sim[it] & kHookBit)) hook_items.push_back(it);
          tile_sim[it] |= kHookBit;
        }
      };
      if (d.has_src && b->d_amp)
        for (int i = d.src.ci0; i < d.src.ci1; ++i) mark(i);
      if (d.has_probe && b->d_probe_acc && b->nfreq > 0) mark(d.pi);
    }
  }
  int rc = batch_free_graph(b);
  if (rc) return rc;
  if (b->d_desc) HF_HIP(hipFree(b->d_desc));
  if (b->d_tile_sim) HF_HIP(hipFree(b->d_tile_sim));
  if (b->d_hook_items) HF_HIP(hipFree(b->d_hook_items));
  b->d_desc = nullptr;
  b->d_tile_sim = nullptr;
  b->d_hook_items = nullptr;
  b->n_hook = (int)hook_items.size();
  if (b->n_hook) {
    HF_HIP(hipMalloc((void **)&b->d_hook_items, sizeof(int) * hook_items.size()));
    HF_HIP(hipMemcpyAsync(b->d_hook_items, hook_items.data(), sizeof(int) * hook_items.size(),
                          hipMemcpyHostToDevice, b->stream));
  }
  HF_HIP(hipMalloc(&b->d_desc, sizeof(SimDesc<T>) * n));
  HF_HIP(hipMalloc((void **)&b->d_tile_sim, sizeof(int) * tile_sim.size()));
  HF_HIP(hipMemcpyAsync(b->d_desc, desc.data(), sizeof(SimDesc<T>) * n, hipMemcpyHostToDevice,
                        b->stream));
  HF_HIP(hipMemcpyAsync(b->d_tile_sim, tile_sim.data(), sizeof(int) * tile_sim.size(),
                        hipMemcpyHostToDevice, b->stream));
  HF_HIP(hipStreamSynchronize(b->stream));
  b->n_tiles = (int)tile_sim.size();
  return HFDTD_OK;
}

// March path: per-grid descriptors (parity origin = current step) and the
// work list of march tiles.  The kernel reads the material index when every
// grid has one, else the ce volumes (with ca when some grid is lossy).
// Element offset of psi slot s in a psi slab of geometry g.
static long long psi_slot_host(const Geo &g, int s) { return (long long)psi_slot(g, s); }

// The arena's layout (in elements of T, every piece aligned to 64): per
// engine the state slabs of both parities, the psi slabs of both parities,
// the coefficient slab (ce volumes kernels), the profiles and the material
// index (indexed kernels).  Rows are padded to whole 128-byte lines when
// pad is set and nz does not fill them.
template <typename T>
static size_t arena_layout(hfdtd_batch *b, bool indexed) {
  const size_t n = b->eng.size();
  b->parts.assign(n, ArenaPart{});
  const long long per_line = 128 / (long long)sizeof(T);
  auto up = [](long long v) { return (v + 63) / 64 * 64; };
  long long at = 0;
  for (size_t s = 0; s < n; ++s) {
    const Engine *e = b->eng[s];
    ArenaPart &a = b->parts[s];
    a.g = geo_of(e);
    if (b->mpad) a.g.sk = (int)((e->nz + per_line - 1) / per_line * per_line);
    a.g.vs = up((long long)e->nx * e->ny * a.g.sk + 64);
    a.slab = 6 * a.g.vs;
    a.psi_len = up(psi_slot_host(a.g, PSI_COUNT) + 64);
    a.st = at;
    at += 2 * a.slab;
    a.psi = at;
    at += 2 * a.psi_len;
    a.coef = at;
    if (!indexed) at += a.slab;
    a.prof = at;
    at += up(6 * (long long)e->prof_len());
    a.mat = at * (long long)sizeof(T);
    if (indexed) at += up(((long long)e->nx * e->ny * a.g.sk + 64 + sizeof(T) - 1) / sizeof(T));
  }
  return (size_t)at * sizeof(T);
}

// Row copies between engine s's slabs of parity ep (its state[cur ^ ep] and
// psi[cur ^ ep]) and its arena part of parity q, in the direction to_arena;
// with coef, also its coefficient slab and profiles (into the arena).
template <typename T>
static void arena_jobs(const hfdtd_batch *b, int s, int ep, int q, bool to_arena, bool coef,
                       std::vector<PadJob> &jobs) {
  const Engine *e = b->eng[s];
  const ArenaPart &a = b->parts[s];
  T *base = (T *)b->arena;
  constexpr int F = (int)(sizeof(T) / sizeof(float));  // floats per element
  auto add = [&](const void *eng_p, T *ar_p, long long rows, long long w, long long ew,
                 long long aw) {
    if (rows <= 0 || w <= 0) return;
    float *ep_ = (float *)const_cast<void *>(eng_p), *ap = (float *)ar_p;
    jobs.push_back(to_arena ? PadJob{ep_, ap, rows, (int)(w * F), (int)(ew * F), (int)(aw * F)}
                            : PadJob{ap, ep_, rows, (int)(w * F), (int)(aw * F), (int)(ew * F)});
  };
  const long long rows = (long long)e->nx * e->ny;
  for (int c = 0; c < 6; ++c)
    add((const T *)e->state[e->cur ^ ep] + (size_t)c * e->vstride,
        base + a.st + q * a.slab + c * a.g.vs, rows, e->nz, e->nz, a.g.sk);
  const long long n2 = 2ll * e->npml;
  for (int t = 0; t < PSI_COUNT; ++t) {
    const T *eng_p = (const T *)e->psi[e->cur ^ ep] + e->psi_off[t];
    T *ar_p = base + a.psi + q * a.psi_len + psi_slot_host(a.g, t);
    const int u = t % 6;
    if (u < 2) add(eng_p, ar_p, n2 * e->ny, e->nz, e->nz, a.g.sk);
    else if (u < 4) add(eng_p, ar_p, (long long)e->nx * n2, e->nz, e->nz, a.g.sk);
    else add(eng_p, ar_p, 1, rows * n2, rows * n2, rows * n2);
  }
  if (coef) {
    if (a.prof != a.coef)  // the ce volumes kernels read the coefficient slab
      for (int c = 0; c < 6; ++c)
        add((const T *)e->coef + (size_t)c * e->vstride, base + a.coef + c * a.g.vs, rows, e->nz,
            e->nz, a.g.sk);
    const long long pl = 6 * (long long)e->prof_len();
    add(e->prof, base + a.prof, 1, pl, pl, pl);
  }
}

int run_pad_jobs(hipStream_t stream, const std::vector<PadJob> &jobs, PadJob *&d_jobs, int &cap) {
  if (jobs.empty()) return HFDTD_OK;
  if ((int)jobs.size() > cap) {
    if (d_jobs) HF_HIP(hipFree(d_jobs));
    d_jobs = nullptr;
    cap = 0;
    HF_HIP(hipMalloc((void **)&d_jobs, sizeof(PadJob) * jobs.size()));
    cap = (int)jobs.size();
  }
  HF_HIP(hipMemcpyAsync(d_jobs, jobs.data(), sizeof(PadJob) * jobs.size(), hipMemcpyHostToDevice,
                        stream));
  long long most = 0;
  for (const PadJob &j : jobs) most = std::max(most, j.rows * j.w);
  const unsigned gx = (unsigned)std::min<long long>((most + 255) / 256, 2048);
  for (size_t j0 = 0; j0 < jobs.size(); j0 += 65535) {  // grid y of one launch: 65535 jobs
    const unsigned gy = (unsigned)std::min<size_t>(jobs.size() - j0, 65535);
    hipLaunchKernelGGL(k_pad_rows, dim3(gx, gy), dim3(256), 0, stream,
                       (const PadJob *)d_jobs + j0);
    HF_HIP(hipGetLastError());
  }
  return HFDTD_OK;
}

static int pad_run(hfdtd_batch *b, const std::vector<PadJob> &jobs) {
  return run_pad_jobs(b->stream, jobs, b->d_jobs, b->n_jobs_cap);
}

// Lay the arena out, (re)allocate it when it grows (zeroed: pad columns stay
// zero) and fill it from the engines' current state, coefficients,
// profiles and material index.
template <typename T>
static int arena_setup(hfdtd_batch *b, bool indexed) {
  const size_t bytes = arena_layout<T>(b, indexed);
  if (bytes > b->arena_bytes) {
    if (b->arena) HF_HIP(hipFree(b->arena));
    b->arena = nullptr;
    b->arena_bytes = 0;
    HF_HIP(hipMalloc(&b->arena, bytes));
    HF_HIP(hipMemsetAsync(b->arena, 0, bytes, b->stream));
    b->arena_bytes = bytes;
  }
  std::vector<PadJob> jobs;
  for (int s = 0; s < (int)b->eng.size(); ++s) {
    const Engine *e = b->eng[s];
    const ArenaPart &a = b->parts[s];
    if (indexed)
      HF_HIP(hipMemcpy2DAsync((char *)b->arena + a.mat, (size_t)a.g.sk, e->mat, (size_t)e->nz,
                              (size_t)e->nz, (size_t)e->nx * e->ny, hipMemcpyDeviceToDevice,
                              b->stream));
    arena_jobs<T>(b, s, 0, 0, true, true, jobs);
  }
  return pad_run(b, jobs);
}

// After n steps: the state and psi of parity n & 1 back into each engine's
// slabs of that parity (the caller then flips cur when n is odd).
template <typename T>
static int arena_writeback(hfdtd_batch *b, int n_steps) {
  if (n_steps <= 0) return HFDTD_OK;
  std::vector<PadJob> jobs;
  for (int s = 0; s < (int)b->eng.size(); ++s)
    arena_jobs<T>(b, s, n_steps & 1, n_steps & 1, false, false, jobs);
  int rc = pad_run(b, jobs);
  if (rc) return rc;
  HF_HIP(hipStreamSynchronize(b->stream));
  return HFDTD_OK;
}

template <typename T>
static int batch_build_march(hfdtd_batch *b) {
  const int n = (int)b->eng.size();
  const int RW = b->mrows;
  std::vector<BMSim<T>> desc(n);
  std::vector<int> item_sim;
  int begin = 0;
  bool all_mat = sizeof(T) == 4;
  b->mlossy = false;
  for (int s = 0; s < n; ++s) {
    Engine *e = b->eng[s];
    if (!e->fused_fits() || ((double)e->psi_off[PSI_COUNT] + 64.0) * e->elem >= 2147483648.0)
      return fail(HFDTD_EUNSUPPORTED,
                  "batched march tiles need field slabs below 4 GB and psi slabs below 2 GB");
    int rc = ensure_alt(e);
    if (rc) return rc;
    b->mlossy = b->mlossy || !e->lossless;
    all_mat = all_mat && e->ce_source() == kCeIndex;
  }
  b->mce = all_mat ? kCeIndex : kCeVolumes;
  arena_layout<T>(b, all_mat);
  for (const ArenaPart &a : b->parts)  // the sweep's 32-bit slab offsets
    if ((double)a.slab * sizeof(T) >= 4294967296.0 || (double)a.psi_len * sizeof(T) >= 2147483648.0)
      return fail(HFDTD_EUNSUPPORTED,
                  "batched march tiles need field slabs below 4 GB and psi slabs below 2 GB");
  int prc = arena_setup<T>(b, all_mat);
  if (prc) return prc;
  for (int s = 0; s < n; ++s) {
    Engine *e = b->eng[s];
    const ArenaPart &a = b->parts[s];
    BMSim<T> &d = desc[s];
    d.g = a.g;
    d.st[0] = a.st;
    d.st[1] = a.st + a.slab;
    d.psi[0] = a.psi;
    d.psi[1] = a.psi + a.psi_len;
    d.coef = a.coef;
    d.prof = a.prof;
    d.mat = a.mat;
    d.mtab = all_mat ? e->mtab : nullptr;
    d.nmat = all_mat ? e->nmat : 0;
    d.planes = (b->mplanes + 1) / 2 * 2;
    d.tiles_k = (int)ceil_div(e->nz, b->mwidth);
    d.tiles_j = (int)ceil_div(e->ny, 4 * RW * (64 / b->mwidth));
    d.ch.split = e->npml > 0 && e->nx > 2 * e->npml ? 1 : 0;
    d.ch.inner = (int)ceil_div(d.ch.split ? e->nx - 2 * e->npml : e->nx, d.planes);
    const int nt = d.tiles_k * d.tiles_j * (d.ch.inner + 2 * d.ch.split);
    d.item0 = begin;
    item_sim.insert(item_sim.end(), nt, s);
    begin += nt;
    d.has_src = b->has_src.empty() ? 0 : b->has_src[s];
    if (d.has_src) {
      hfdtd_source in = b->src[s];
      in.amp = 0.0;  // from the table
      if (!make_src<T>(e, in, d.src)) return fail(HFDTD_EINVAL, "invalid source descriptor");
    } else {
      d.src = Src<T>{};
    }
    d.has_probe = 0;
    d.pi = d.pj = d.pk = -1;
    if (!b->probe.empty() && b->probe[3 * s] >= 0) {
      d.has_probe = 1;
      d.pi = b->probe[3 * s];
      d.pj = b->probe[3 * s + 1];
      d.pk = b->probe[3 * s + 2];
    }
  }
  int rc = batch_free_graph(b);
  if (rc) return rc;
  if (b->d_desc) HF_HIP(hipFree(b->d_desc));
  if (b->d_tile_sim) HF_HIP(hipFree(b->d_tile_sim));
  b->d_desc = nullptr;
  b->d_tile_sim = nullptr;
  HF_HIP(hipMalloc(&b->d_desc, sizeof(BMSim<T>) * n));
  HF_HIP(hipMalloc((void **)&b->d_tile_sim, sizeof(int) * item_sim.size()));
  HF_HIP(hipMemcpyAsync(b->d_desc, desc.data(), sizeof(BMSim<T>) * n, hipMemcpyHostToDevice,
                        b->stream));
  HF_HIP(hipMemcpyAsync(b->d_tile_sim, item_sim.data(), sizeof(int) * item_sim.size(),
                        hipMemcpyHostToDevice, b->stream));
  HF_HIP(hipStreamSynchronize(b->stream));
  b->n_tiles = (int)item_sim.size();
  return HFDTD_OK;
}

template <typename T>
static BMArgs<T> batch_march_args(const hfdtd_batch *b) {
  const Engine *e = b->eng[0];
  BMArgs<T> a;
  a.sims = (const BMSim<T> *)b->d_desc;
  a.base = b->arena;
  a.item_sim = b->d_tile_sim;
  a.n_items = b->n_tiles;
  a.step = b->d_step;
  a.step0 = b->step_host;
  a.dtm_dl = T(e->dtm_dl);
  a.dtm = T(e->dtm);
  a.inv_dl = T(e->inv_dl);
  a.amp = b->d_amp;
  a.phase = b->d_phase;
  a.probe_acc = b->nfreq > 0 ? b->d_probe_acc : nullptr;
  a.n_table = std::max(b->n_amp, b->n_phase);
  a.nfreq = b->nfreq;
  a.comp = b->comp;
  a.xcd = e->xcd;
  a.run = b->mrun;
  return a;
}

// Launch q of a run of batched march steps.
template <typename T>
static int batch_launch_march(const hfdtd_batch *b, const BMArgs<T> &a, int q) {
  if constexpr (sizeof(T) == 8) {
    return b->mlossy ? launch_batch_march<T, 1, true, kCeVolumes, 64>(a, q, b->stream)
                     : launch_batch_march<T, 1, false, kCeVolumes, 64>(a, q, b->stream);
  } else {
#define HF_BM(RW, W)                                                                     \
  return b->mce == kCeIndex ? launch_batch_march<T, RW, false, kCeIndex, W>(a, q, b->stream)  \
         : b->mlossy        ? launch_batch_march<T, RW, true, kCeVolumes, W>(a, q, b->stream) \
                            : launch_batch_march<T, RW, false, kCeVolumes, W>(a, q, b->stream)
    if (b->mrows == 2) HF_BM(2, 64);
    if (b->mwidth == 32) HF_BM(1, 32);
    HF_BM(1, 64);
#undef HF_BM
  }
}

template <typename T>
static BatchArgs<T> batch_args(const hfdtd_batch *b) {
  const Engine *e = b->eng[0];
  BatchArgs<T> a;
  a.sims = (const SimDesc<T> *)b->d_desc;
  a.tile_sim = b->d_tile_sim;
  a.n_tiles = b->n_tiles;
  a.step = b->d_step;
  a.step0 = b->step_host;
  a.planes = b->planes;
  a.dtm_dl = T(e->dtm_dl);
  a.dtm = T(e->dtm);
  a.inv_dl = T(e->inv_dl);
  a.amp = b->d_amp;
  a.phase = b->d_phase;
  a.n_table = std::max(b->n_amp, b->n_phase);
  a.nfreq = b->nfreq;
  a.comp = b->comp;
  a.probe_acc = b->d_probe_acc;
  a.xcd = e->xcd;
  a.run = e->split_run;
  a.n_sims = (int)b->eng.size();
  a.runs = b->split_runs;
  a.hook_items = b->d_hook_items;
  a.n_hook = b->n_hook;
  return a;
}

// Step q of a run on the split path: the H sweep, the E sweep, and E on the
// items that hold the engines' sources and probes.
template <typename T>
static int batch_launch_split(hfdtd_batch *b, const BatchArgs<T> &a, int q) {
  const int grid = b->blocks > 0 ? std::min(b->blocks, b->n_tiles) : b->n_tiles;
  hipLaunchKernelGGL(k_batch_split_H<T>, dim3(grid), dim3(kFlatBlock), 0, b->stream, a);
  HF_HIP(hipGetLastError());
  if (b->lossy)
    hipLaunchKernelGGL((k_batch_split_E<T, true>), dim3(grid), dim3(kFlatBlock), 0, b->stream, a);
  else
    hipLaunchKernelGGL((k_batch_split_E<T, false>), dim3(grid), dim3(kFlatBlock), 0, b->stream, a);
  HF_HIP(hipGetLastError());
  if (a.n_hook) {
    const dim3 hg((unsigned)a.n_hook * a.runs);
    if (b->lossy)
      hipLaunchKernelGGL((k_batch_split_Ehook<T, true>), hg, dim3(kFlatBlock), 0, b->stream, a, q);
    else
      hipLaunchKernelGGL((k_batch_split_Ehook<T, false>), hg, dim3(kFlatBlock), 0, b->stream, a,
                         q);
    HF_HIP(hipGetLastError());
  }
  return HFDTD_OK;
}

// Step q of a run on the fused path: the fused tiles, then one block per
// engine for its source and probe.
template <typename T>
static int batch_launch_step(hfdtd_batch *b, const BatchArgs<T> &a, int q) {
  if (b->path == HFDTD_PATH_SPLIT) return batch_launch_split<T>(b, a, q);
  const int grid = b->blocks > 0 ? std::min(b->blocks, b->n_tiles) : b->n_tiles;
#define HF_BATCH(TJ)                                                                          \
  if (b->lossy)                                                                               \
    hipLaunchKernelGGL((k_batch_step<T, TJ, true>), dim3(grid), dim3(64, TJ + 1), 0, b->stream, a, q); \
  else                                                                                        \
    hipLaunchKernelGGL((k_batch_step<T, TJ, false>), dim3(grid), dim3(64, TJ + 1), 0, b->stream, a, q)
  if constexpr (sizeof(T) == 8) {  // FP64: tiles of at most 4 rows fit the LDS
    if (b->rows == 2) HF_BATCH(2);
    else HF_BATCH(4);
  } else {
    switch (b->rows) {
      case 2: HF_BATCH(2); break;
      case 6: HF_BATCH(6); break;
      case 8: HF_BATCH(8); break;
      default: HF_BATCH(4); break;
    }
  }
#undef HF_BATCH
  HF_HIP(hipGetLastError());
  if (a.amp || a.probe_acc) {
    hipLaunchKernelGGL(k_batch_fixup<T>, dim3((unsigned)b->eng.size()), dim3(256), 0, b->stream,
                       a, q);
    HF_HIP(hipGetLastError());
  }
  return HFDTD_OK;
}

// n_steps launches of launch(q), q counting within each graph or direct run,
// with *step advanced after each run; graph_steps > 0 captures that many
// steps once and replays them (remainder launched directly).
template <typename F>
static int batch_run_steps(hfdtd_batch *b, int n_steps, int graph_steps, F launch) {
  int done = 0, rc = HFDTD_OK;
  if (graph_steps > 0 && n_steps >= graph_steps) {
    hipGraph_t graph;
    HF_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    for (int q = 0; q < graph_steps && rc == HFDTD_OK; ++q) rc = launch(q);
    if (rc == HFDTD_OK) {
      hipLaunchKernelGGL(k_step_add, dim3(1), dim3(64), 0, b->stream, b->d_step, graph_steps);
      if (hipGetLastError() != hipSuccess) rc = fail(HFDTD_EHIP, "k_step_add launch");
    }
    if (rc) {
      hipGraph_t junk;
      (void)hipStreamEndCapture(b->stream, &junk);
      if (junk) (void)hipGraphDestroy(junk);
      return rc;
    }
    HF_HIP(hipStreamEndCapture(b->stream, &graph));
    hipError_t err = hipGraphInstantiate(&b->gexec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (err != hipSuccess) return hip_fail(err, "hipGraphInstantiate");
    for (; done + graph_steps <= n_steps; done += graph_steps)
      HF_HIP(hipGraphLaunch(b->gexec, b->stream));
  }
  if (done < n_steps) {
    for (int q = 0; q < n_steps - done; ++q)
      if ((rc = launch(q))) return rc;
    hipLaunchKernelGGL(k_step_add, dim3(1), dim3(64), 0, b->stream, b->d_step, n_steps - done);
    HF_HIP(hipGetLastError());
  }
  HF_HIP(hipStreamSynchronize(b->stream));
  return HFDTD_OK;
}

template <typename T>
static int batch_run_t(hfdtd_batch *b, int n_steps, int graph_steps) {
  int rc;
  if (b->path == HFDTD_PATH_MARCH) {
    if ((rc = batch_build_march<T>(b))) return rc;
    const BMArgs<T> a = batch_march_args<T>(b);
    rc = batch_run_steps(b, n_steps, graph_steps,
                         [&](int q) { return batch_launch_march<T>(b, a, q); });
    if (rc == HFDTD_OK) rc = arena_writeback<T>(b, n_steps);
  } else {
    if ((rc = batch_build<T>(b))) return rc;
    const BatchArgs<T> a = batch_args<T>(b);
    rc = batch_run_steps(b, n_steps, graph_steps,
                         [&](int q) { return batch_launch_step<T>(b, a, q); });
  }
  if (rc) return rc;
  b->step_host += n_steps;
  if (b->path != HFDTD_PATH_SPLIT && (n_steps & 1))
    for (Engine *e : b->eng) e->cur ^= 1;
  return HFDTD_OK;
}

}  // namespace hfdtd

using namespace hfdtd;

extern "C" {

int hfdtd_probe_create(hfdtd_engine *h, int i, int j, int k, int comp, int nfreq,
                       hfdtd_probe **out) {
  if (!h || !out) return fail(HFDTD_EINVAL, "NULL argument");
  *out = nullptr;
  Engine *e = &h->s;
  HF_SETTLE(e);
  if (i < 0 || i >= e->nx || j < 0 || j >= e->ny || k < 0 || k >= e->nz)
    return fail(HFDTD_EINVAL, "probe cell outside the grid");
  if (comp < 0 || comp > 5 || nfreq < 1) return fail(HFDTD_EINVAL, "bad component or nfreq");
  HF_HIP(hipSetDevice(e->device));
  hfdtd_probe *p = new hfdtd_probe();
  p->e = e;
  p->i = i;
  p->j = j;
  p->k = k;
  p->comp = comp;
  p->nfreq = nfreq;
  hipError_t err = hipMalloc((void **)&p->acc, sizeof(double) * 2 * nfreq);
  if (err == hipSuccess) {
    std::vector<double> z(2 * nfreq, 0.0);
    err = hipMemcpy(p->acc, z.data(), sizeof(double) * 2 * nfreq, hipMemcpyHostToDevice);
  }
  if (err != hipSuccess) {
    if (p->acc) (void)hipFree(p->acc);
    delete p;
    return hip_fail(err, "hfdtd_probe_create");
  }
  *out = p;
  return HFDTD_OK;
}

int hfdtd_probe_destroy(hfdtd_probe *p) {
  if (!p) return HFDTD_OK;
  (void)hipSetDevice(p->e->device);
  (void)hipStreamSynchronize(p->e->stream);
  if (p->acc) (void)hipFree(p->acc);
  delete p;
  return HFDTD_OK;
}

int hfdtd_probe_accumulate(hfdtd_probe *p, const double *phase) {
  if (!p || !phase) return fail(HFDTD_EINVAL, "NULL argument");
  Engine *e = p->e;  // the field volume as the engine holds it (a row-padded run's too)
  HF_HIP(hipSetDevice(e->device));
  const size_t idx = ((size_t)p->i * e->ny + p->j) * e->sk + p->k;
  for (int f0 = 0; f0 < p->nfreq; f0 += kProbeFreqPerLaunch) {
    const int nf = std::min(kProbeFreqPerLaunch, p->nfreq - f0);
    ProbePhase ph;
    for (int q = 0; q < nf; ++q) {
      ph.v[q][0] = phase[2 * (f0 + q)];
      ph.v[q][1] = phase[2 * (f0 + q) + 1];
    }
    if (e->fp64)
      hipLaunchKernelGGL(k_probe_accumulate<double>, dim3(1), dim3(64), 0, e->stream,
                         (const double *)e->field(p->comp), idx, f0, nf, ph, p->acc);
    else
      hipLaunchKernelGGL(k_probe_accumulate<float>, dim3(1), dim3(64), 0, e->stream,
                         (const float *)e->field(p->comp), idx, f0, nf, ph, p->acc);
    HF_HIP(hipGetLastError());
  }
  return HFDTD_OK;
}

int hfdtd_probe_read(hfdtd_probe *p, double *out) {
  if (!p || !out) return fail(HFDTD_EINVAL, "NULL argument");
  HF_HIP(hipSetDevice(p->e->device));
  HF_HIP(hipMemcpyAsync(out, p->acc, sizeof(double) * 2 * p->nfreq, hipMemcpyDeviceToHost,
                        p->e->stream));
  HF_HIP(hipStreamSynchronize(p->e->stream));
  return HFDTD_OK;
}

int hfdtd_batch_create(hfdtd_engine *const *engines, int n, hfdtd_batch **out) {
  if (!engines || !out || n < 1) return fail(HFDTD_EINVAL, "bad engine list");
  *out = nullptr;
  const Engine *e0 = &engines[0]->s;
  for (int s = 0; s < n; ++s) {
    if (!engines[s]) return fail(HFDTD_EINVAL, "NULL engine");
    const Engine *e = &engines[s]->s;
    if (e->fp64 != e0->fp64 || e->dl != e0->dl || e->dt != e0->dt || e->device != e0->device)
      return fail(HFDTD_EINVAL, "batched engines need the same dl, dt, dtype and device");
  }
  hfdtd_batch *b = new hfdtd_batch();
  for (int s = 0; s < n; ++s) b->eng.push_back(&engines[s]->s);
  b->fp64 = e0->fp64;
  hipError_t err = hipSetDevice(e0->device);
  if (err == hipSuccess) err = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (err == hipSuccess) err = hipMalloc((void **)&b->d_step, sizeof(int));
  if (err == hipSuccess) {
    const int zero = 0;
    err = hipMemcpy(b->d_step, &zero, sizeof(int), hipMemcpyHostToDevice);
  }
  // the engines' own streams must be idle before batch work starts
  for (int s = 0; s < n && err == hipSuccess; ++s) err = hipStreamSynchronize(b->eng[s]->stream);
  if (err != hipSuccess) {
    int rc = hip_fail(err, "hfdtd_batch_create");
    hfdtd_batch_destroy(b);
    return rc;
  }
  *out = b;
  return HFDTD_OK;
}

int hfdtd_batch_destroy(hfdtd_batch *b) {
  if (!b) return HFDTD_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  (void)batch_free_graph(b);
  for (void *p : {(void *)b->d_desc, (void *)b->d_tile_sim, (void *)b->d_hook_items,
                  (void *)b->d_step, (void *)b->d_amp,
                  (void *)b->d_phase, (void *)b->d_probe_acc, (void *)b->d_jobs})
    if (p) (void)hipFree(p);
  if (b->arena) (void)hipFree(b->arena);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return HFDTD_OK;
}

static int reset_step(hfdtd_batch *b) {
  const int zero = 0;
  HF_HIP(hipMemcpy(b->d_step, &zero, sizeof(int), hipMemcpyHostToDevice));
  b->step_host = 0;
  return HFDTD_OK;
}

int hfdtd_batch_set_sources(hfdtd_batch *b, const hfdtd_source *src, const double *amp,
                            int n_steps) {
  if (!b || !src || !amp || n_steps < 1) return fail(HFDTD_EINVAL, "bad source tables");
  const int n = (int)b->eng.size();
  for (int s = 0; s < n; ++s) {
    if (src[s].comp < 0 || src[s].comp > 2 ||
        (src[s].mode != HFDTD_SRC_SOFT && src[s].mode != HFDTD_SRC_CURRENT))
      return fail(HFDTD_EINVAL, "invalid source descriptor");
  }
  if (b->n_phase > 0 && b->n_phase != n_steps)
    return fail(HFDTD_EINVAL, "source and probe tables must cover the same steps");
  HF_HIP(hipSetDevice(b->eng[0]->device));
  b->src.assign(src, src + n);
  b->has_src.assign(n, 1);
  if (b->d_amp) HF_HIP(hipFree(b->d_amp));
  b->d_amp = nullptr;
  HF_HIP(hipMalloc((void **)&b->d_amp, sizeof(double) * n * n_steps));
  HF_HIP(hipMemcpy(b->d_amp, amp, sizeof(double) * n * n_steps, hipMemcpyHostToDevice));
  b->n_amp = n_steps;
  return reset_step(b);
}

int hfdtd_batch_set_probes(hfdtd_batch *b, const int *ijk, int comp, int nfreq,
                           const double *phase, int n_steps) {
  if (!b || !ijk || !phase || nfreq < 1 || n_steps < 1) return fail(HFDTD_EINVAL, "bad probes");
  if (comp < 0 || comp > 2) return fail(HFDTD_EINVAL, "batched probes record Ex, Ey or Ez");
  if (b->n_amp > 0 && b->n_amp != n_steps)
    return fail(HFDTD_EINVAL, "source and probe tables must cover the same steps");
  const int n = (int)b->eng.size();
  for (int s = 0; s < n; ++s) {
    const Engine *e = b->eng[s];
    if (ijk[3 * s] < 0) continue;
    if (ijk[3 * s] >= e->nx || ijk[3 * s + 1] < 0 || ijk[3 * s + 1] >= e->ny ||
        ijk[3 * s + 2] < 0 || ijk[3 * s + 2] >= e->nz)
      return fail(HFDTD_EINVAL, "probe cell outside its grid");
  }
  HF_HIP(hipSetDevice(b->eng[0]->device));
  b->probe.assign(ijk, ijk + 3 * n);
  b->comp = comp;
  b->nfreq = nfreq;
  if (b->d_phase) HF_HIP(hipFree(b->d_phase));
  if (b->d_probe_acc) HF_HIP(hipFree(b->d_probe_acc));
  b->d_phase = nullptr;
  b->d_probe_acc = nullptr;
  const size_t np = (size_t)n * n_steps * nfreq * 2;
  HF_HIP(hipMalloc((void **)&b->d_phase, sizeof(double) * np));
  HF_HIP(hipMemcpy(b->d_phase, phase, sizeof(double) * np, hipMemcpyHostToDevice));
  HF_HIP(hipMalloc((void **)&b->d_probe_acc, sizeof(double) * n * nfreq * 2));
  std::vector<double> z((size_t)n * nfreq * 2, 0.0);
  HF_HIP(hipMemcpy(b->d_probe_acc, z.data(), sizeof(double) * z.size(), hipMemcpyHostToDevice));
  b->n_phase = n_steps;
  return reset_step(b);
}

int hfdtd_batch_run(hfdtd_batch *b, int n_steps, int graph_steps) {
  if (!b || n_steps < 0 || graph_steps < 0) return fail(HFDTD_EINVAL, "bad run arguments");
  const int table = std::max(b->n_amp, b->n_phase);
  if (table > 0 && b->step_host + n_steps > table)
    return fail(HFDTD_EINVAL, "run goes past the end of the source/probe tables");
  if (b->nfreq > 0 && b->d_probe_acc == nullptr) return fail(HFDTD_EINVAL, "probes not set");
  HF_HIP(hipSetDevice(b->eng[0]->device));
  for (Engine *e : b->eng) {
    HF_SETTLE(e);
    HF_HIP(hipStreamSynchronize(e->stream));
    e->touch();
  }
  return b->fp64 ? batch_run_t<double>(b, n_steps, graph_steps)
                 : batch_run_t<float>(b, n_steps, graph_steps);
}

int hfdtd_batch_read_probes(hfdtd_batch *b, double *out) {
  if (!b || !out) return fail(HFDTD_EINVAL, "NULL argument");
  if (!b->d_probe_acc) return fail(HFDTD_EINVAL, "probes not set");
  HF_HIP(hipSetDevice(b->eng[0]->device));
  const size_t n = b->eng.size() * (size_t)b->nfreq * 2;
  HF_HIP(hipMemcpyAsync(out, b->d_probe_acc, sizeof(double) * n, hipMemcpyDeviceToHost,
                        b->stream));
  HF_HIP(hipStreamSynchronize(b->stream));
  return HFDTD_OK;
}

int hfdtd_batch_set_path(hfdtd_batch *b, int path) {
  if (!b || (path != HFDTD_PATH_FUSED && path != HFDTD_PATH_SPLIT && path != HFDTD_PATH_MARCH))
    return fail(HFDTD_EINVAL,
                "path must be HFDTD_PATH_FUSED, HFDTD_PATH_SPLIT or HFDTD_PATH_MARCH");
  b->path = path;
  return batch_free_graph(b);
}

int hfdtd_batch_set_launch(hfdtd_batch *b, int blocks, int rows_j, int planes_i) {
  if (!b || blocks < 0 || planes_i < 0) return fail(HFDTD_EINVAL, "bad launch shape");
  if (rows_j != 0 && rows_j != 2 && rows_j != 4 && rows_j != 6 && rows_j != 8)
    return fail(HFDTD_EINVAL, "rows_j must be 2, 4, 6 or 8");
  if (b->fp64 && rows_j > 4) return fail(HFDTD_EINVAL, "FP64 batches use rows_j 2 or 4");
  b->blocks = blocks;
  if (rows_j) b->rows = rows_j;
  if (planes_i) b->planes = planes_i;
  return HFDTD_OK;
}

int hfdtd_batch_set_split(hfdtd_batch *b, int runs) {
  if (!b || runs < 0 || runs > kSplitRunsMax)
    return fail(HFDTD_EINVAL, "split runs per work item must be 1..8 (0 = keep)");
  if (runs) b->split_runs = runs;
  return batch_free_graph(b);
}

int hfdtd_batch_set_march(hfdtd_batch *b, int rows, int planes, int run, int width) {
  if (!b || rows < 0 || rows > 2 || planes < 0 || planes > 256 || run < 0 ||
      (width != 0 && width != 64 && width != 32))
    return fail(HFDTD_EINVAL, "march rows must be 1 or 2, planes in [2, 256], run >= 1, width "
                              "64 or 32 (0 = keep)");
  if (b->fp64 && (rows == 2 || (width && width != 64)))
    return fail(HFDTD_EINVAL, "FP64 batches use one row per wavefront and 64-column tiles");
  const int r = rows ? rows : b->mrows, w = width ? width : b->mwidth;
  if (r == 2 && w != 64) return fail(HFDTD_EINVAL, "two rows per wavefront need 64-column tiles");
  if (width) b->mwidth = width;
  if (rows) b->mrows = rows;
  if (planes) b->mplanes = std::max(2, planes);
  if (run) b->mrun = run;
  return batch_free_graph(b);
}

int hfdtd_batch_set_march_padding(hfdtd_batch *b, int on) {
  if (!b || (on != 0 && on != 1)) return fail(HFDTD_EINVAL, "padding must be 0 or 1");
  b->mpad = on;
  return batch_free_graph(b);
}

}  // extern "C"
