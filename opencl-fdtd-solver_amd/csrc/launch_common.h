// Host-side pieces shared by the step launchers (yee.hip, march*.hip,
// pair.hip): source descriptors and the source fix-up after a ping-pong
// sweep.
#pragma once

#include "fdtd_core.h"

namespace hfdtd {

template <typename T>
inline bool make_src(const Engine *e, const hfdtd_source &in, Src<T> &o) {
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

template <typename T>
inline int build_srclist(Engine *e, const hfdtd_source *src, int n, SrcList<T> &sl) {
  if (n > kMaxSources) return fail(HFDTD_EINVAL, "at most 16 sources per fused step");
  sl.n = 0;
  for (int q = 0; q < n; ++q) {
    Src<T> s;
    if (!make_src<T>(e, src[q], s)) return fail(HFDTD_EINVAL, "invalid source descriptor");
    if (s.ci1 > s.ci0 && s.cj1 > s.cj0 && s.ck1 > s.ck0) sl.s[sl.n++] = s;
  }
  return HFDTD_OK;
}

// E on the source cells of a ping-pong sweep, redone with the sources added
// (yee.hip; instantiated there for float and double, lossless and lossy).
template <typename T, bool LOSSY>
int source_fixup(Engine *e, const SrcList<T> &sl);

}  // namespace hfdtd
