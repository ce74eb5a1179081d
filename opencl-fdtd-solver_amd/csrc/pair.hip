// Pair sweep (pair.h): two Yee steps per launch, the step launcher.

#include "launch_common.h"
#include "pair.h"

namespace hfdtd {

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
  hipEvent_t t0, t1;
  int rc = timing_events(e, t0, t1);
  if (rc) return rc;
  const int o = e->cur ^ 1;
  hipExtLaunchKernelGGL((k_pair<T, NW, RW, LOSSY>), grid, dim3(NW * 64), 0, e->stream, t0, t1, 0,
                        geo_of(e), planes, T(e->dtm_dl), T(e->dtm), T(e->inv_dl),
                        (const T *)e->state[e->cur], (T *)e->state[o], (const T *)e->coef,
                        (const T *)e->prof, (const T *)e->psi[e->cur], (T *)e->psi[o],
                        pw.n ? (T *)e->shadow : nullptr, tiles_k, tiles_j, e->xcd, e->march_run,
                        ps, pw);
  HF_HIP(hipGetLastError());
  return HFDTD_OK;
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

}  // namespace hfdtd
