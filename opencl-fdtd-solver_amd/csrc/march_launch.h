// Launcher of the march sweep (k_march, march.h) for one tiling and medium
// kind.  Defined here and instantiated in march_f32_r1.hip, march_f32_r2.hip
// and march_f64.hip, so the many sweep bodies compile in parallel; march.hip
// picks the instantiation for a step.
#pragma once

#include "launch_common.h"
#include "march.h"

namespace hfdtd {

template <typename T, int RW, bool LOSSY, bool NT, int CEM = kCeVolumes, bool BIG = false,
          int W = 64>
int march_rw(Engine *e, const SrcList<T> &sl) {
  auto kern = k_march<T, RW, LOSSY, NT, CEM, BIG, W>;
  constexpr int wi = W == 64 ? 0 : 1;
  int big_occ = 0;
  int &occ = BIG                 ? big_occ
             : CEM != kCeVolumes ? e->march_occ_mat[wi][CEM][RW]
                                 : e->march_occ[wi][LOSSY ? 1 : 0][RW];
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
  const int planes = e->march_planes_for(RW, occ, W);
  const int tiles_k = (int)ceil_div(e->nz, W);
  const int tiles_j = (int)ceil_div(e->ny, MarchShape<RW, W>::TJ);
  MarchChunks ch;
  ch.split = e->npml > 0 && e->nx > 2 * e->npml ? 1 : 0;
  ch.inner = (int)ceil_div(ch.split ? e->nx - 2 * e->npml : e->nx, planes);
  // a held Huygens-box gather of the step before runs in blocks after the
  // sweep's (FP32 slabs below 4 GB; otherwise on its own, first)
  GatherTail gt;
  int gblocks = 0;
  if (e->pgather && e->pgather->held) {
    if constexpr (sizeof(T) == 4 && !BIG) {
      gt = *e->pgather;
      gblocks = gt.blocks();
      e->pgather->held = false;
    } else {
      int rg = flush_gather(e);
      if (rg) return rg;
    }
  }
  const int nsweep = tiles_k * tiles_j * (ch.inner + 2 * ch.split);
  const dim3 grid(nsweep + gblocks);
  MarchSources<T> ms;
  const bool inside = sl.n <= kMarchSources;
  ms.n = inside ? sl.n : 0;
  for (int q = 0; q < ms.n; ++q) ms.s[q] = sl.s[q];
  // the z-face record, if a monitor keeps one (FP32)
  ZCap zc{};
  const ZRec &z = e->zrec;
  const bool record = sizeof(T) == 4 && !BIG && z.owner && z.buf;
  if (record) {
    zc.buf = z.half(e->cur ^ 1);  // the record of the slab this sweep writes
    zc.kf0 = z.kf[0];
    zc.kf1 = z.kf[1];
    zc.i0 = z.i0;
    zc.i1 = z.i1;
    zc.j0 = z.j0;
    zc.j1 = z.j1;
    zc.bytes = (unsigned)z.bytes;
  }
  hipEvent_t t0, t1;
  int rc = timing_events(e, t0, t1);
  if (rc) return rc;
  const int o = e->cur ^ 1;
  hipExtLaunchKernelGGL(kern, grid, dim3(kMarchThreads), 0, e->stream, t0, t1, 0, geo_of(e),
                        planes, T(e->dtm_dl), T(e->dtm), T(e->inv_dl),
                        (const T *)e->state[e->cur], (T *)e->state[o], (const T *)e->coef,
                        (const T *)e->prof, (const T *)e->psi[e->cur], (T *)e->psi[o], tiles_k,
                        tiles_j, ch, e->xcd, e->march_run, ms,
                        CEM == kCeIndex ? (const void *)e->mat
                                        : CEM == kCeEps ? (const void *)e->epsv : nullptr,
                        CEM == kCeIndex ? (const float *)e->mtab : nullptr,
                        CEM == kCeIndex ? e->nmat : 0, e->dt / kEPS0, zc, nsweep, gt);
  HF_HIP(hipGetLastError());
  // the record holds this step's values unless the fix-up kernel below
  // changes E on source cells afterwards
  if (record && inside) e->zrec.epoch = e->state_epoch;
  if (inside) return rc;
  return source_fixup<T, LOSSY>(e, sl);
}

#define HF_MARCH_RW(T, RW, LOSSY, NT, CEM, W) \
  template int march_rw<T, RW, LOSSY, NT, CEM, false, W>(Engine *, const SrcList<T> &)
// slabs of 4 GB and more (FP32, streaming)
#define HF_MARCH_RW_BIG1(X, RW)                                                                   \
  X template int march_rw<float, RW, false, true, kCeVolumes, true>(Engine *, const SrcList<float> &); \
  X template int march_rw<float, RW, true, true, kCeVolumes, true>(Engine *, const SrcList<float> &);  \
  X template int march_rw<float, RW, false, true, kCeIndex, true>(Engine *, const SrcList<float> &);   \
  X template int march_rw<float, RW, false, true, kCeEps, true>(Engine *, const SrcList<float> &)
#define HF_MARCH_RW_BIG(X) \
  HF_MARCH_RW_BIG1(X, 1);  \
  HF_MARCH_RW_BIG1(X, 2)
// W: the tile width (64 or 32 columns)
#define HF_MARCH_RW_SET(X, T, RW, W)        \
  X HF_MARCH_RW(T, RW, false, true, kCeVolumes, W);  \
  X HF_MARCH_RW(T, RW, false, false, kCeVolumes, W); \
  X HF_MARCH_RW(T, RW, true, true, kCeVolumes, W);   \
  X HF_MARCH_RW(T, RW, true, false, kCeVolumes, W)
#define HF_MARCH_RW_MAT(X, T, RW, W)       \
  X HF_MARCH_RW(T, RW, false, true, kCeIndex, W);  \
  X HF_MARCH_RW(T, RW, false, false, kCeIndex, W); \
  X HF_MARCH_RW(T, RW, false, true, kCeEps, W);    \
  X HF_MARCH_RW(T, RW, false, false, kCeEps, W)

}  // namespace hfdtd
