// Yee H/E updates with face-stripe CPML, source injection, and the fused
// H+sources+E sweep.
//
// Per-cell arithmetic follows the reference kernels term for term
// (kernels/yee_update.cl:11-238, kernels/sources.cl:7-49): interior cells use
// the folded dtm_dl / inv_dl form, PML-shell cells the kappa-scaled form with
// psi recursions on the face stripes.  The split sweeps and the fused sweep
// call the same h_math / e_math device functions, so both paths produce the
// same values.
//
// Layout: C-order (Nx, Ny, Nz), idx = (i*Ny + j)*Nz + k, k contiguous.  One
// wavefront (64 lanes) spans 64 consecutive k, so every field load and store
// is a 256-byte coalesced segment (FP32).  The six fields, the six
// coefficients and the twelve psi stripes each live in one slab, so a kernel
// receives a few base pointers instead of dozens.

#include <algorithm>

#include "fdtd_core.h"
#include "launch_common.h"

namespace hfdtd {

// ── split sweeps (reference order: all of H, then sources, then all of E) ──
//
// One thread per cell over the flattened (j, k) plane: a block of 256 threads
// covers 256 consecutive cells of one i-plane, so rows of any length pack the
// wavefronts without idle lanes (64-lane k-rows would leave up to 63 lanes
// idle per row).  Logical block L works on plane L / bpp, bpp = blocks per
// plane; xcd_runs hands every XCD runs of consecutive L, so most j+-1
// neighbour rows a cell reads come through the same XCD's L2, while all
// XCDs stream the same region of the grid.

template <typename T>
__global__ __launch_bounds__(kFlatBlock) void k_update_H(Geo g, T dtm_dl, T dtm,
                                                          T *__restrict__ st,
                                                          const T *__restrict__ prof,
                                                          T *__restrict__ psi, int bpp,
                                                          int xcd, int run) {
  const int L = xcd_runs(blockIdx.x, gridDim.x, xcd, run);
  const int i = L / bpp;
  const int q = (L - i * bpp) * kFlatBlock + threadIdx.x;
  if (q >= g.ny * g.nz) return;
  const int j = q / g.nz;
  split_h_cell(g, dtm_dl, dtm, st, prof, psi, i, j, q - j * g.nz);
}

// MAT: ce from the material index (lossless FP32 grids), not the volumes.
template <typename T, bool LOSSY, bool MAT = false>
__global__ __launch_bounds__(kFlatBlock) void k_update_E(Geo g, T inv_dl, T *__restrict__ st,
                                                          const T *__restrict__ coef,
                                                          const T *__restrict__ prof,
                                                          T *__restrict__ psi, int bpp,
                                                          int xcd, int run,
                                                          const unsigned char *__restrict__ mat,
                                                          const float *__restrict__ mtab,
                                                          int nmat) {
  static_assert(!MAT || (sizeof(T) == 4 && !LOSSY), "the material index is lossless FP32");
  const int L = xcd_runs(blockIdx.x, gridDim.x, xcd, run);
  const int i = L / bpp;
  const int q = (L - i * bpp) * kFlatBlock + threadIdx.x;
  if (q >= g.ny * g.nz) return;
  const int j = q / g.nz;
  if constexpr (MAT)
    split_e_cell_mat(g, inv_dl, st, coef, mat, mtab, nmat, prof, psi, i, j, q - j * g.nz);
  else
    split_e_cell<T, LOSSY>(g, inv_dl, st, coef, prof, psi, i, j, q - j * g.nz);
}

// ── sources ──────────────────────────────────────────────────────────────

template <typename T>
__global__ void k_add_source(Geo g, Src<T> s, T *__restrict__ F, const T *__restrict__ ce) {
  const int ni = s.ci1 - s.ci0, nj = s.cj1 - s.cj0, nk = s.ck1 - s.ck0;
  const size_t total = (size_t)ni * nj * nk;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int k = s.ck0 + (int)(t % nk);
  const int j = s.cj0 + (int)((t / nk) % nj);
  const int i = s.ci0 + (int)(t / ((size_t)nk * nj));
  const size_t idx = ((size_t)i * g.ny + j) * g.nz + k;
  F[idx] += src_term<T>(s, i, j, k, ce[idx], s.amp);
}

// ── fused sweep kernel ────────────────────────────────────────────────────

template <typename T, int TJ, bool LOSSY>
__global__ __launch_bounds__(64 * (TJ + 1)) void k_fused(Geo g, int planes, T dtm_dl, T dtm,
                                                   T inv_dl, const T *__restrict__ s0,
                                                   T *__restrict__ s1,
                                                   const T *__restrict__ coef,
                                                   const T *__restrict__ prof,
                                                   const T *__restrict__ q0,
                                                   T *__restrict__ q1, int tiles_k,
                                                   int tiles_j, int xcd) {
  // 1-D grid of tiles, k fastest, then j, then the i chunk
  const int L = xcd_remap(blockIdx.x, gridDim.x, xcd);
  const int bx = L % tiles_k, by = (L / tiles_k) % tiles_j, bz = L / (tiles_k * tiles_j);
  fused_tile<T, TJ, LOSSY>(g, planes, dtm_dl, dtm, inv_dl, s0, s1, coef, prof, q0, q1, bx, by,
                           bz);
}

// Source cells after a fused sweep: blockIdx.y picks the source whose box
// this block walks; each cell gets every source that covers it (a cell in
// two boxes is recomputed twice, to the same value).
template <typename T, bool LOSSY>
__global__ __launch_bounds__(256) void k_source_fixup(Geo g, T inv_dl, const T *__restrict__ s0,
                                                      T *__restrict__ s1,
                                                      const T *__restrict__ coef,
                                                      const T *__restrict__ prof,
                                                      const T *__restrict__ q0,
                                                      T *__restrict__ q1, SrcList<T> src) {
  const Src<T> &b = src.s[blockIdx.y];
  const int ni = b.ci1 - b.ci0, nj = b.cj1 - b.cj0, nk = b.ck1 - b.ck0;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)ni * nj * nk) return;
  const int k = b.ck0 + (int)(t % nk);
  const int j = b.cj0 + (int)((t / nk) % nj);
  const int i = b.ci0 + (int)(t / ((long long)nk * nj));
  source_fixup_cell<T, LOSSY>(g, inv_dl, s0, s1, coef, prof, q0, q1, SrcListP<T>{src}, i, j, k);
}

// ── launchers ─────────────────────────────────────────────────────────────

template <typename T>
static int update_H_t(Engine *e) {
  const unsigned bpp = ceil_div((size_t)e->ny * e->nz, kFlatBlock);
  hipEvent_t t0, t1;
  int rc = timing_events(e, t0, t1);
  if (rc) return rc;
  hipExtLaunchKernelGGL(k_update_H<T>, dim3(bpp * (unsigned)e->nx), dim3(kFlatBlock), 0,
                        e->stream, t0, t1, 0, geo_of(e), T(e->dtm_dl), T(e->dtm),
                        (T *)e->state[e->cur], (const T *)e->prof, (T *)e->psi[e->cur], (int)bpp,
                        e->xcd, e->split_run);
  HF_HIP(hipGetLastError());
  return HFDTD_OK;
}

template <typename T>
static int update_E_t(Engine *e) {
  const unsigned bpp = ceil_div((size_t)e->ny * e->nz, kFlatBlock);
  auto kern = e->lossless ? k_update_E<T, false> : k_update_E<T, true>;
  const bool mat = sizeof(T) == 4 && e->lossless && e->mat_ok;
  if constexpr (sizeof(T) == 4)
    if (mat) kern = k_update_E<float, false, true>;
  hipEvent_t t0, t1;
  int rc = timing_events(e, t0, t1);
  if (rc) return rc;
  hipExtLaunchKernelGGL(kern, dim3(bpp * (unsigned)e->nx), dim3(kFlatBlock), 0, e->stream, t0, t1,
                        0, geo_of(e), T(e->inv_dl), (T *)e->state[e->cur], (const T *)e->coef,
                        (const T *)e->prof, (T *)e->psi[e->cur], (int)bpp, e->xcd, e->split_run,
                        mat ? (const unsigned char *)e->mat : nullptr,
                        mat ? (const float *)e->mtab : nullptr, mat ? e->nmat : 0);
  HF_HIP(hipGetLastError());
  return HFDTD_OK;
}

template <typename T>
static int sources_t(Engine *e, const hfdtd_source *src, int n) {
  for (int q = 0; q < n; ++q) {
    Src<T> s;
    if (!make_src<T>(e, src[q], s)) return fail(HFDTD_EINVAL, "invalid source descriptor");
    const size_t total = (size_t)(s.ci1 - s.ci0) * (s.cj1 - s.cj0) * (s.ck1 - s.ck0);
    if (total == 0) continue;
    hipLaunchKernelGGL(k_add_source<T>, dim3(ceil_div(total, 256)), dim3(256), 0, e->stream,
                       geo_of(e), s, (T *)e->field(s.comp), (const T *)e->coef_vol(s.comp));
    HF_HIP(hipGetLastError());
  }
  return HFDTD_OK;
}

// E on the source cells of a ping-pong sweep, redone with the sources added.
template <typename T, bool LOSSY>
int source_fixup(Engine *e, const SrcList<T> &sl) {
  if (sl.n == 0) return HFDTD_OK;
  const int o = e->cur ^ 1;
  long long most = 0;
  for (int q = 0; q < sl.n; ++q) {
    const Src<T> &s = sl.s[q];
    most = std::max(most, (long long)(s.ci1 - s.ci0) * (s.cj1 - s.cj0) * (s.ck1 - s.ck0));
  }
  auto fix = k_source_fixup<T, LOSSY>;
  hipLaunchKernelGGL(fix, dim3((unsigned)ceil_div(most, 256), sl.n), dim3(256), 0, e->stream,
                     geo_of(e), T(e->inv_dl), (const T *)e->state[e->cur], (T *)e->state[o],
                     (const T *)e->coef, (const T *)e->prof, (const T *)e->psi[e->cur],
                     (T *)e->psi[o], sl);
  HF_HIP(hipGetLastError());
  return HFDTD_OK;
}
template int source_fixup<float, false>(Engine *, const SrcList<float> &);
template int source_fixup<float, true>(Engine *, const SrcList<float> &);
template int source_fixup<double, false>(Engine *, const SrcList<double> &);
template int source_fixup<double, true>(Engine *, const SrcList<double> &);

template <typename T, int TJ, bool LOSSY>
static int fused_tj(Engine *e, const SrcList<T> &sl) {
  auto kern = k_fused<T, TJ, LOSSY>;
  int &occ = e->fused_occ[LOSSY ? 1 : 0][TJ];
  if (occ == 0) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * (TJ + 1), 0) !=
            hipSuccess ||
        per_cu < 1)
      per_cu = 1;
    occ = per_cu;
  }
  const int planes = e->fused_planes(TJ, occ);
  const int tiles_k = (int)ceil_div(e->nz, 64), tiles_j = (int)ceil_div(e->ny, TJ);
  const dim3 grid(tiles_k * tiles_j * ceil_div(e->nx, planes));
  hipEvent_t t0, t1;
  int rc = timing_events(e, t0, t1);
  if (rc) return rc;
  const int o = e->cur ^ 1;
  hipExtLaunchKernelGGL(kern, grid, dim3(64, TJ + 1), 0, e->stream, t0, t1, 0, geo_of(e), planes,
                        T(e->dtm_dl), T(e->dtm), T(e->inv_dl), (const T *)e->state[e->cur],
                        (T *)e->state[o], (const T *)e->coef, (const T *)e->prof,
                        (const T *)e->psi[e->cur], (T *)e->psi[o], tiles_k, tiles_j, e->xcd);
  HF_HIP(hipGetLastError());
  return source_fixup<T, LOSSY>(e, sl);
}

template <typename T>
static int fused_t(Engine *e, const hfdtd_source *src, int n) {
  SrcList<T> sl;
  int rc = build_srclist<T>(e, src, n, sl);
  if (rc) return rc;
  if (!e->fused_fits())
    return fail(HFDTD_EUNSUPPORTED, "the fused sweep needs field and psi slabs below 4 GB; "
                                    "use the split sweeps for this grid");
  if ((rc = ensure_alt(e))) return rc;
  // FP64 slots are twice as large: tiles of at most 4 rows fit the LDS.
  // FP32 tiles of 12 rows fit only without the ca rows (lossy grids use 10).
  const bool ll = e->lossless;
  if constexpr (sizeof(T) == 8) {
    switch (e->tile_rows) {
      case 1: rc = ll ? fused_tj<T, 1, false>(e, sl) : fused_tj<T, 1, true>(e, sl); break;
      case 2: rc = ll ? fused_tj<T, 2, false>(e, sl) : fused_tj<T, 2, true>(e, sl); break;
      default: rc = ll ? fused_tj<T, 4, false>(e, sl) : fused_tj<T, 4, true>(e, sl); break;
    }
  } else {
    switch (e->tile_rows) {
      case 1: rc = ll ? fused_tj<T, 1, false>(e, sl) : fused_tj<T, 1, true>(e, sl); break;
      case 2: rc = ll ? fused_tj<T, 2, false>(e, sl) : fused_tj<T, 2, true>(e, sl); break;
      case 6: rc = ll ? fused_tj<T, 6, false>(e, sl) : fused_tj<T, 6, true>(e, sl); break;
      case 8: rc = ll ? fused_tj<T, 8, false>(e, sl) : fused_tj<T, 8, true>(e, sl); break;
      case 10: rc = ll ? fused_tj<T, 10, false>(e, sl) : fused_tj<T, 10, true>(e, sl); break;
      case 12: rc = ll ? fused_tj<T, 12, false>(e, sl) : fused_tj<T, 10, true>(e, sl); break;
      default: rc = ll ? fused_tj<T, 4, false>(e, sl) : fused_tj<T, 4, true>(e, sl); break;
    }
  }
  if (rc) return rc;
  e->cur ^= 1;
  return HFDTD_OK;
}

int launch_update_H(Engine *e) {
  HF_HIP(hipSetDevice(e->device));
  e->touch();
  return e->fp64 ? update_H_t<double>(e) : update_H_t<float>(e);
}

int launch_update_E(Engine *e) {
  HF_HIP(hipSetDevice(e->device));
  e->touch();
  return e->fp64 ? update_E_t<double>(e) : update_E_t<float>(e);
}

int launch_sources(Engine *e, const hfdtd_source *src, int n) {
  HF_HIP(hipSetDevice(e->device));
  e->touch();
  return e->fp64 ? sources_t<double>(e, src, n) : sources_t<float>(e, src, n);
}

int launch_fused_step(Engine *e, const hfdtd_source *src, int n) {
  HF_HIP(hipSetDevice(e->device));
  e->touch();
  return e->fp64 ? fused_t<double>(e, src, n) : fused_t<float>(e, src, n);
}

}  // namespace hfdtd
