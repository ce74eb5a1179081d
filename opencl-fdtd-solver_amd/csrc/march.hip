// March sweep: the step launcher (tiling and medium kind pick one of the
// launchers instantiated in march_f32_r1.hip, march_f32_r2.hip,
// march_f32_w32.hip, march_f64.hip, march_f32_big.hip).

#include "march_launch.h"

namespace hfdtd {

HF_MARCH_RW_SET(extern, float, 1, 64);
HF_MARCH_RW_MAT(extern, float, 1, 64);
HF_MARCH_RW_SET(extern, float, 2, 64);
HF_MARCH_RW_MAT(extern, float, 2, 64);
HF_MARCH_RW_SET(extern, double, 1, 64);
HF_MARCH_RW_SET(extern, float, 1, 32);
HF_MARCH_RW_MAT(extern, float, 1, 32);
HF_MARCH_RW_BIG(extern);

// slabs of 4 GB and more: FP32 volumes below 2 GB each, psi slab below 2 GB
static bool march_big_fits(const Engine *e) {
  return !e->fp64 && (double)e->vstride * e->elem < 2147483648.0 &&
         ((double)e->psi_off[PSI_COUNT] + 64.0) * e->elem < 2147483648.0;
}

// the launcher for slabs of 4 GB and more
static int march_big(Engine *e, const SrcList<float> &sl) {
  const bool ll = e->lossless;
  const int ce = e->ce_source();
  if (e->march_rows == 1 || (e->march_rows == 0 && ce == kCeIndex)) {
    if (ce == kCeIndex) return march_rw<float, 1, false, true, kCeIndex, true>(e, sl);
    if (ce == kCeEps) return march_rw<float, 1, false, true, kCeEps, true>(e, sl);
    return ll ? march_rw<float, 1, false, true, kCeVolumes, true>(e, sl)
              : march_rw<float, 1, true, true, kCeVolumes, true>(e, sl);
  }
  if (ce == kCeIndex) return march_rw<float, 2, false, true, kCeIndex, true>(e, sl);
  if (ce == kCeEps) return march_rw<float, 2, false, true, kCeEps, true>(e, sl);
  return ll ? march_rw<float, 2, false, true, kCeVolumes, true>(e, sl)
            : march_rw<float, 2, true, true, kCeVolumes, true>(e, sl);
}

// FP32 tiles narrower than 64 columns (one row per wavefront slot)
template <int W>
static int march_narrow(Engine *e, const SrcList<float> &sl, bool nt) {
  const int ce = e->ce_source();
  if (ce == kCeIndex)
    return nt ? march_rw<float, 1, false, true, kCeIndex, false, W>(e, sl)
              : march_rw<float, 1, false, false, kCeIndex, false, W>(e, sl);
  if (ce == kCeEps)
    return nt ? march_rw<float, 1, false, true, kCeEps, false, W>(e, sl)
              : march_rw<float, 1, false, false, kCeEps, false, W>(e, sl);
  if (e->lossless)
    return nt ? march_rw<float, 1, false, true, kCeVolumes, false, W>(e, sl)
              : march_rw<float, 1, false, false, kCeVolumes, false, W>(e, sl);
  return nt ? march_rw<float, 1, true, true, kCeVolumes, false, W>(e, sl)
            : march_rw<float, 1, true, false, kCeVolumes, false, W>(e, sl);
}

template <typename T>
static int march_t(Engine *e, const hfdtd_source *src, int n) {
  SrcList<T> sl;
  int rc = build_srclist<T>(e, src, n, sl);
  if (rc) return rc;
  const bool big = !e->fused_fits() || (sizeof(T) == 4 && e->march_per_volume);
  if (big && !march_big_fits(e))
    return fail(HFDTD_EUNSUPPORTED, "the march sweep needs field slabs below 4 GB (FP32: volumes "
                                    "below 2 GB) and psi slabs below 2 GB; use the split sweeps "
                                    "for this grid");
  if (((double)e->psi_off[PSI_COUNT] + 64.0) * e->elem >= 2147483648.0)
    return fail(HFDTD_EUNSUPPORTED, "the march sweep needs psi slabs below 2 GB; use the split "
                                    "sweeps for this grid");
  if ((rc = ensure_alt(e))) return rc;
  if constexpr (sizeof(T) == 4) {
    if (big) {
      if ((rc = march_big(e, sl))) return rc;
      e->cur ^= 1;
      return HFDTD_OK;
    }
  }
  const bool ll = e->lossless, nt = e->march_streaming();
  if constexpr (sizeof(T) == 8) {
    rc = ll ? (nt ? march_rw<T, 1, false, true>(e, sl) : march_rw<T, 1, false, false>(e, sl))
            : (nt ? march_rw<T, 1, true, true>(e, sl) : march_rw<T, 1, true, false>(e, sl));
  } else if (e->march_width != 64) {
    rc = march_narrow<32>(e, sl, nt);
  } else if (e->march_rows == 1 || (e->march_rows == 0 && e->ce_source() == kCeIndex)) {
    // one row per wavefront: four blocks per CU; with the material index's
    // light loads that occupancy pays (256^3: 233 against 249 us)
    if (e->ce_source() == kCeIndex)
      rc = nt ? march_rw<T, 1, false, true, kCeIndex>(e, sl)
              : march_rw<T, 1, false, false, kCeIndex>(e, sl);
    else if (e->ce_source() == kCeEps)
      rc = nt ? march_rw<T, 1, false, true, kCeEps>(e, sl) : march_rw<T, 1, false, false, kCeEps>(e, sl);
    else
      rc = ll ? (nt ? march_rw<T, 1, false, true>(e, sl) : march_rw<T, 1, false, false>(e, sl))
              : (nt ? march_rw<T, 1, true, true>(e, sl) : march_rw<T, 1, true, false>(e, sl));
  } else if (e->ce_source() == kCeIndex) {  // lossless with a material index: ce from its table
    rc = nt ? march_rw<T, 2, false, true, kCeIndex>(e, sl)
            : march_rw<T, 2, false, false, kCeIndex>(e, sl);
  } else if (e->ce_source() == kCeEps) {  // lossless, eps_r volume: ce recomputed per edge
    rc = nt ? march_rw<T, 2, false, true, kCeEps>(e, sl) : march_rw<T, 2, false, false, kCeEps>(e, sl);
  } else {
    rc = ll ? (nt ? march_rw<T, 2, false, true>(e, sl) : march_rw<T, 2, false, false>(e, sl))
            : (nt ? march_rw<T, 2, true, true>(e, sl) : march_rw<T, 2, true, false>(e, sl));
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

}  // namespace hfdtd
