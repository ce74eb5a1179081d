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
#include "march.h"
#include "pair.h"

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

template <typename T, bool LOSSY>
__global__ __launch_bounds__(kFlatBlock) void k_update_E(Geo g, T inv_dl, T *__restrict__ st,
                                                          const T *__restrict__ coef,
                                                          const T *__restrict__ prof,
                                                          T *__restrict__ psi, int bpp,
                                                          int xcd, int run) {
  const int L = xcd_runs(blockIdx.x, gridDim.x, xcd, run);
  const int i = L / bpp;
  const int q = (L - i * bpp) * kFlatBlock + threadIdx.x;
  if (q >= g.ny * g.nz) return;
  const int j = q / g.nz;
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

template <typename T>
static bool make_src(const Engine *e, const hfdtd_source &in, Src<T> &o) {
  if (in.comp < 0 || in.comp > 2) return false;
  if (in.mode != HFDTD_SRC_SOFT && in.mode != HFDTD_SRC_CURRENT) return false;
  o.comp = in.comp;
  o.mode = in.mode;
  o.taper = in.taper & 7;
  o.i0 = in.i0; o.i1 = in.i1; o.j0 = in.j0; o.j1 = in.j1; o.k0 = in.k0; o.k1 = in.k1;
  auto clip = [](int a, int n) { return a < 0 ? 0 : (a > n ? n : a); };
  o.ci0 = clip(in.i0, e->nx); o.ci1 = clip(in.i1, e->nx);
  o.cj0 = clip(in.j0, e->ny); o.cj1 = clip(in.j1, e->ny);
  o.ck0 = clip(in.k0, e->nz); o.ck1 = clip(in.k1, e->nz);
  if (o.ci1 < o.ci0) o.ci1 = o.ci0;
  if (o.cj1 < o.cj0) o.cj1 = o.cj0;
  if (o.ck1 < o.ck0) o.ck1 = o.ck0;
  o.amp = T(in.amp);
  o.re = T(in.rim_edge);
  return true;
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
  int rc = timing_mark(e);
  if (rc) return rc;
  hipLaunchKernelGGL(k_update_H<T>, dim3(bpp * (unsigned)e->nx), dim3(kFlatBlock), 0, e->stream,
                     geo_of(e), T(e->dtm_dl), T(e->dtm), (T *)e->state[e->cur],
                     (const T *)e->prof, (T *)e->psi[e->cur], (int)bpp, e->xcd, e->split_run);
  HF_HIP(hipGetLastError());
  return timing_mark(e);
}

template <typename T>
static int update_E_t(Engine *e) {
  const unsigned bpp = ceil_div((size_t)e->ny * e->nz, kFlatBlock);
  auto kern = e->lossless ? k_update_E<T, false> : k_update_E<T, true>;
  int rc = timing_mark(e);
  if (rc) return rc;
  hipLaunchKernelGGL(kern, dim3(bpp * (unsigned)e->nx), dim3(kFlatBlock), 0, e->stream,
                     geo_of(e), T(e->inv_dl), (T *)e->state[e->cur], (const T *)e->coef,
                     (const T *)e->prof, (T *)e->psi[e->cur], (int)bpp, e->xcd, e->split_run);
  HF_HIP(hipGetLastError());
  return timing_mark(e);
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
static int source_fixup(Engine *e, const SrcList<T> &sl) {
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
  int rc = timing_mark(e);
  if (rc) return rc;
  const int o = e->cur ^ 1;
  hipLaunchKernelGGL(kern, grid, dim3(64, TJ + 1), 0, e->stream, geo_of(e), planes, T(e->dtm_dl),
                     T(e->dtm), T(e->inv_dl), (const T *)e->state[e->cur], (T *)e->state[o],
                     (const T *)e->coef, (const T *)e->prof, (const T *)e->psi[e->cur],
                     (T *)e->psi[o], tiles_k, tiles_j, e->xcd);
  HF_HIP(hipGetLastError());
  rc = timing_mark(e);
  if (rc) return rc;
  return source_fixup<T, LOSSY>(e, sl);
}

template <typename T, int RW, bool LOSSY, bool NT, bool MAT = false>
static int march_rw(Engine *e, const SrcList<T> &sl) {
  auto kern = k_march<T, RW, LOSSY, NT, MAT>;
  int &occ = MAT ? e->march_occ_mat[RW] : e->march_occ[LOSSY ? 1 : 0][RW];
  if (occ == 0) {
    // resident blocks per CU from the kernel's own resource use: a block puts
    // one wavefront on each of the four SIMDs, each SIMD holds 512 VGPRs per
    // lane (allocated in granules of 8), a CU has 160 KB of LDS
    hipFuncAttributes fa;
    int per_cu = 1;
    if (hipFuncGetAttributes(&fa, (const void *)kern) == hipSuccess && fa.numRegs > 0) {
      const int vg = (fa.numRegs + 7) / 8 * 8;
      per_cu = 512 / vg;
      if (fa.sharedSizeBytes > 0)
        per_cu = std::min(per_cu, (int)(160 * 1024 / fa.sharedSizeBytes));
      per_cu = std::max(per_cu, 1);
    }
    occ = per_cu;
  }
  const int planes = e->march_planes_for(RW, occ);
  const int tiles_k = (int)ceil_div(e->nz, 64), tiles_j = (int)ceil_div(e->ny, 4 * RW);
  MarchChunks ch;
  ch.split = e->npml > 0 && e->nx > 2 * e->npml ? 1 : 0;
  ch.inner = (int)ceil_div(ch.split ? e->nx - 2 * e->npml : e->nx, planes);
  const dim3 grid(tiles_k * tiles_j * (ch.inner + 2 * ch.split));
  MarchSources<T> ms;
  const bool inside = sl.n <= kMarchSources;
  ms.n = inside ? sl.n : 0;
  for (int q = 0; q < ms.n; ++q) ms.s[q] = sl.s[q];
  // the z-face record, if a monitor keeps one (FP32)
  ZCap zc{};
  const ZRec &z = e->zrec;
  const bool record = sizeof(T) == 4 && z.owner && z.buf;
  if (record) {
    zc.buf = z.buf;
    zc.kf0 = z.kf[0];
    zc.kf1 = z.kf[1];
    zc.i0 = z.i0;
    zc.i1 = z.i1;
    zc.j0 = z.j0;
    zc.j1 = z.j1;
    zc.bytes = (unsigned)z.bytes;
  }
  int rc = timing_mark(e);
  if (rc) return rc;
  const int o = e->cur ^ 1;
  hipLaunchKernelGGL(kern, grid, dim3(kMarchThreads), 0, e->stream, geo_of(e), planes,
                     T(e->dtm_dl), T(e->dtm), T(e->inv_dl), (const T *)e->state[e->cur],
                     (T *)e->state[o], (const T *)e->coef, (const T *)e->prof,
                     (const T *)e->psi[e->cur], (T *)e->psi[o], tiles_k, tiles_j, ch, e->xcd,
                     e->march_run, ms, MAT ? (const unsigned char *)e->mat : nullptr,
                     MAT ? (const float *)e->mtab : nullptr, MAT ? e->nmat : 0, zc);
  HF_HIP(hipGetLastError());
  rc = timing_mark(e);
  if (rc) return rc;
  // the record holds this step's values unless the fix-up kernel below
  // changes E on source cells afterwards
  if (record && inside) e->zrec.epoch = e->state_epoch;
  if (inside) return rc;
  return source_fixup<T, LOSSY>(e, sl);
}

template <typename T>
static int build_srclist(Engine *e, const hfdtd_source *src, int n, SrcList<T> &sl) {
  if (n > kMaxSources) return fail(HFDTD_EINVAL, "at most 16 sources per fused step");
  sl.n = 0;
  for (int q = 0; q < n; ++q) {
    Src<T> s;
    if (!make_src<T>(e, src[q], s)) return fail(HFDTD_EINVAL, "invalid source descriptor");
    if (s.ci1 > s.ci0 && s.cj1 > s.cj0 && s.ck1 > s.ck0) sl.s[sl.n++] = s;
  }
  return HFDTD_OK;
}

template <typename T>
static int march_t(Engine *e, const hfdtd_source *src, int n) {
  SrcList<T> sl;
  int rc = build_srclist<T>(e, src, n, sl);
  if (rc) return rc;
  if (!e->fused_fits() || ((double)e->psi_off[PSI_COUNT] + 64.0) * e->elem >= 2147483648.0)
    return fail(HFDTD_EUNSUPPORTED, "the march sweep needs field slabs below 4 GB and psi "
                                    "slabs below 2 GB; use the split sweeps for this grid");
  if ((rc = ensure_alt(e))) return rc;
  const bool ll = e->lossless, nt = e->march_streaming();
  if constexpr (sizeof(T) == 8) {
    rc = ll ? (nt ? march_rw<T, 1, false, true>(e, sl) : march_rw<T, 1, false, false>(e, sl))
            : (nt ? march_rw<T, 1, true, true>(e, sl) : march_rw<T, 1, true, false>(e, sl));
  } else if (e->march_rows == 1 || (e->march_rows == 0 && ll && e->mat_ok)) {
    // one row per wavefront: four blocks per CU; with the material index's
    // light loads that occupancy pays (256^3: 233 against 249 us)
    if (ll && e->mat_ok)
      rc = nt ? march_rw<T, 1, false, true, true>(e, sl) : march_rw<T, 1, false, false, true>(e, sl);
    else
      rc = ll ? (nt ? march_rw<T, 1, false, true>(e, sl) : march_rw<T, 1, false, false>(e, sl))
              : (nt ? march_rw<T, 1, true, true>(e, sl) : march_rw<T, 1, true, false>(e, sl));
  } else if (ll && e->mat_ok) {  // lossless with a material index: ce from its table
    rc = nt ? march_rw<T, 2, false, true, true>(e, sl) : march_rw<T, 2, false, false, true>(e, sl);
  } else {
    rc = ll ? (nt ? march_rw<T, 2, false, true>(e, sl) : march_rw<T, 2, false, false>(e, sl))
            : (nt ? march_rw<T, 2, true, true>(e, sl) : march_rw<T, 2, true, false>(e, sl));
  }
  if (rc) return rc;
  e->cur ^= 1;
  return HFDTD_OK;
}

// ── pair sweep: two steps per launch ──

template <typename T, int NW, int RW, bool LOSSY>
static int pair_launch(Engine *e, const PairSources<T> &ps) {
  constexpr int TJ = NW * RW - 4;
  const int tiles_k = (int)ceil_div(e->nz, kPairCols), tiles_j = (int)ceil_div(e->ny, TJ);
  const int planes = e->pair_planes > 0 ? e->pair_planes : 24;
  const int chunks = (int)ceil_div(e->nx, planes);
  const dim3 grid(tiles_k * tiles_j * chunks);
  PairWatch pw{};
  pw.n = e->shadow ? e->watch_n : 0;
  for (int b = 0; b < pw.n; ++b) {
    for (int c = 0; c < 6; ++c) pw.b[b][c] = e->watch_b[b][c];
    pw.c[b] = e->watch_c[b];
  }
  int rc = timing_mark(e);
  if (rc) return rc;
  const int o = e->cur ^ 1;
  hipLaunchKernelGGL((k_pair<T, NW, RW, LOSSY>), grid, dim3(NW * 64), 0, e->stream, geo_of(e),
                     planes, T(e->dtm_dl), T(e->dtm), T(e->inv_dl), (const T *)e->state[e->cur],
                     (T *)e->state[o], (const T *)e->coef, (const T *)e->prof,
                     (const T *)e->psi[e->cur], (T *)e->psi[o], pw.n ? (T *)e->shadow : nullptr,
                     tiles_k, tiles_j, e->xcd, e->march_run, ps, pw);
  HF_HIP(hipGetLastError());
  return timing_mark(e);
}

// 16 wavefronts of one row each: every variant fits 128 VGPRs, so a CU
// keeps 16 wavefronts resident (8 x 2 rows needs up to 248 VGPRs on PML
// blocks: 8 wavefronts, 25 % slower at 256^3)
template <typename T, bool LOSSY>
static int pair_shape(Engine *e, const PairSources<T> &ps) {
  return pair_launch<T, 16, 1, LOSSY>(e, ps);
}

template <typename T>
static int pair_t(Engine *e, const hfdtd_source *src_a, int na, const hfdtd_source *src_b,
                  int nb) {
  SrcList<T> l[2];
  int rc = build_srclist<T>(e, src_a, na, l[0]);
  if (rc || (rc = build_srclist<T>(e, src_b, nb, l[1]))) return rc;
  if (l[0].n > kMarchSources || l[1].n > kMarchSources)
    return fail(HFDTD_EUNSUPPORTED, "the pair sweep takes at most 4 sources per step");
  if (!e->fused_fits() || 6.0 * (double)e->vstride * e->elem >= 2147483648.0 ||
      ((double)e->psi_off[PSI_COUNT] + 64.0) * e->elem >= 2147483648.0)
    return fail(HFDTD_EUNSUPPORTED, "the pair sweep needs field and psi slabs below 2 GB");
  if ((rc = ensure_alt(e))) return rc;
  PairSources<T> ps{};
  for (int t = 0; t < 2; ++t) {
    ps.n[t] = l[t].n;
    for (int q = 0; q < l[t].n; ++q) ps.s[t][q] = l[t].s[q];
  }
  rc = e->lossless ? pair_shape<T, false>(e, ps) : pair_shape<T, true>(e, ps);
  if (rc) return rc;
  e->cur ^= 1;
  return HFDTD_OK;
}

int launch_pair_steps(Engine *e, const hfdtd_source *sa, int na, const hfdtd_source *sb, int nb) {
  HF_HIP(hipSetDevice(e->device));
  e->touch();
  if (e->fp64) return fail(HFDTD_EUNSUPPORTED, "the pair sweep is FP32 only");
  return pair_t<float>(e, sa, na, sb, nb);
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

int launch_march_step(Engine *e, const hfdtd_source *src, int n) {
  HF_HIP(hipSetDevice(e->device));
  e->touch();
  return e->fp64 ? march_t<double>(e, src, n) : march_t<float>(e, src, n);
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
