// Batched march sweep: many independent grids advanced by one launch per
// step, each block running one march tile (march.h) of one grid.
//
// The work list holds every grid's march tiles (k tiles fastest, then j
// tiles, then i chunks, grid after grid); block n takes item
// xcd_runs(n, ...), so each XCD works through runs of consecutive tiles of
// one grid, whose halo rows then share that XCD's L2.  A block reads its
// grid's descriptor, applies the grid's source of the step inside the sweep
// (amplitude from the grid's table) and, when it owns the grid's probe
// cell, adds the probe's DFT term after its planes.  The step index is
// *step + q, where q is the launch's place in its run or graph and one
// small kernel advances *step after the run: every launch argument stays
// fixed across graph replays.
//
// The sweep runs on the batch's arena, one allocation holding a copy of
// every grid's state and psi (both parities), coefficients, profiles and
// material index, filled from the engines before a run and written back
// after it; rows are padded to whole 128-byte lines where nz does not fill
// them (an unaligned row makes a 64-column tile row touch three lines
// instead of two).  A grid's pieces are element offsets into the arena,
// whose address comes in as separate __restrict__ kernel arguments, one
// per role (state read, state written, psi read, psi written,
// coefficients, profiles, material index); the probe sums come in as one
// more.  The compiler can then tell the tile's stores from its loads, as in
// the single-engine sweep, and keeps the next plane's loads in flight
// across the current plane's stores: with pointers read from the
// descriptors it must assume they alias, and the sweep waited for every
// load (15 % slower on one 256^3 grid).  Within a launch the roles never
// share memory: the state and psi written are the other parity of the
// ones read.
#pragma once

#include "march.h"

namespace hfdtd {

// One grid of the batch.
template <typename T>
struct BMSim {
  Geo g;
  // element offsets into the arena (bytes for mat)
  long long st[2];   // state slab, by step parity (relative to step0)
  long long psi[2];  // psi slab, by step parity
  long long coef;
  long long prof;
  long long mat;      // material index (kCeIndex kernels)
  const float *mtab;  // its ce table (nmat x nmat)
  int nmat;
  int tiles_k, tiles_j, planes;
  MarchChunks ch;
  int item0;  // first work item of the grid
  int has_src;
  Src<T> src;  // amplitude from the table
  int has_probe, pi, pj, pk;
};

template <typename T>
struct BMArgs {
  const BMSim<T> *sims;
  const void *base;     // the arena
  const int *item_sim;  // grid of each work item
  int n_items;
  const int *step;  // steps taken since the tables were set, before this run
  int step0;        // value of *step when the descriptors were built (parity origin)
  T dtm_dl, dtm, inv_dl;
  const double *amp;    // n_sims x n_table source amplitudes, or null
  const double *phase;  // n_sims x n_table x nfreq x {re, im} probe phases, or null
  double *probe_acc;    // n_sims x nfreq x {re, im}
  int n_table, nfreq, comp;
  int xcd, run;
};

template <typename T, int RW, bool LOSSY, int CEM, int W>
__global__ __launch_bounds__(kMarchThreads, sizeof(T) == 4 && RW == 1 ? 4 : 2) void k_batch_march(
    BMArgs<T> a, int q, const T *__restrict__ rb, T *__restrict__ wb, const T *__restrict__ qrb,
    T *__restrict__ qwb, const T *__restrict__ cb, const T *__restrict__ pb,
    const unsigned char *__restrict__ mb, double *__restrict__ pacc) {
  constexpr int TJ = MarchShape<RW, W>::TJ;
  __shared__ MarchLds<T, RW, W> L;
  __shared__ Src<T> lsrc[1];
  __shared__ float ltab[CEM == kCeIndex ? kMatMax * kMatMax : 1];
  const int G = xcd_runs(blockIdx.x, gridDim.x, a.xcd, a.run);
  const int s = __builtin_amdgcn_readfirstlane(a.item_sim[G]);
  const BMSim<T> &d = a.sims[s];
  const Geo g = d.g;
  const int step = *a.step + q;
  const int par = (step - a.step0) & 1;
  const size_t row = (size_t)s * a.n_table + step;
  const bool src_on = d.has_src && a.amp;
  if (threadIdx.x == 0 && src_on) {
    Src<T> t = d.src;
    t.amp = T(a.amp[row]);
    lsrc[0] = t;
  }  // visible after the barrier that ends the prologue
  if constexpr (CEM == kCeIndex) {
    for (int t = threadIdx.x; t < d.nmat * d.nmat; t += kMarchThreads) ltab[t] = d.mtab[t];
  }
  int k0, j0, i0, i1;
  march_tile_of<RW, W>(g, G - d.item0, d.planes, d.tiles_k, d.tiles_j, d.ch, a.run, k0, j0, i0, i1);
  int si0 = g.nx, si1 = 0;
  march_src_window<T, RW, W>(d.src, src_on, k0, j0, si0, si1);
  MarchProbe pr{};
  const bool prb = d.has_probe && a.probe_acc && d.pi >= i0 && d.pi < i1 && d.pj >= j0 &&
                   d.pj < j0 + TJ && d.pk >= k0 && d.pk < k0 + W;
  if (prb) {
    pr.i = d.pi;
    pr.j = d.pj;
    pr.k = d.pk;
    pr.comp = a.comp;
    pr.nf = a.nfreq;
    pr.ph = a.phase + row * a.nfreq * 2;
    pr.acc = pacc + (size_t)s * a.nfreq * 2;
  }
  // the source count as the engine's flag (0 or 1, an int the compiler
  // cannot bound), so the sweep keeps its source loop a loop: with a count
  // known to be at most one the compiler peels it into the plane body,
  // where it raises register pressure past the budget (VGPR spills)
  march_tile<T, RW, LOSSY, true, CEM, false, 2, W>(
      L, lsrc, a.amp ? d.has_src : 0, si0, si1, ltab, g, k0, j0, i0, i1, a.dtm_dl, a.dtm, a.inv_dl,
      rb + d.st[par], wb + d.st[par ^ 1], cb + d.coef, pb + d.prof, qrb + d.psi[par],
      qwb + d.psi[par ^ 1], mb + d.mat, d.nmat, 0.0, ZCap{}, prb, pr);
}

// Launch one batched march step (instantiated in batch_march_*.hip).
template <typename T, int RW, bool LOSSY, int CEM, int W>
int launch_batch_march(const BMArgs<T> &a, int q, hipStream_t stream) {
  const T *r = (const T *)a.base;
  hipLaunchKernelGGL((k_batch_march<T, RW, LOSSY, CEM, W>), dim3(a.n_items), dim3(kMarchThreads),
                     0, stream, a, q, r, (T *)r, r, (T *)r, r, r, (const unsigned char *)a.base,
                     a.probe_acc);
  HF_HIP(hipGetLastError());
  return HFDTD_OK;
}

#define HF_BATCH_MARCH(X, T, RW, LOSSY, CEM, W) \
  X template int launch_batch_march<T, RW, LOSSY, CEM, W>(const BMArgs<T> &, int, hipStream_t)
#define HF_BATCH_MARCH_F32(X, RW, W)                    \
  HF_BATCH_MARCH(X, float, RW, false, kCeIndex, W);     \
  HF_BATCH_MARCH(X, float, RW, false, kCeVolumes, W);   \
  HF_BATCH_MARCH(X, float, RW, true, kCeVolumes, W)
#define HF_BATCH_MARCH_F64(X, W)                         \
  HF_BATCH_MARCH(X, double, 1, false, kCeVolumes, W);    \
  HF_BATCH_MARCH(X, double, 1, true, kCeVolumes, W)

}  // namespace hfdtd
