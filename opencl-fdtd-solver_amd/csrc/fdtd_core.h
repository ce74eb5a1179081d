// Device-side Yee/CPML math shared by the single-engine and batched sweeps.
//
// Per-cell arithmetic follows the reference kernels term for term
// (kernels/yee_update.cl:11-238, kernels/sources.cl:7-49): interior cells use
// the folded dtm_dl / inv_dl form, PML-shell cells the kappa-scaled form with
// psi recursions on the face stripes.  Every sweep (split, fused, batched)
// calls the same h_math / e_math functions, so all paths produce the same
// values.
#pragma once

#include "common.h"

namespace hfdtd {

// Grid geometry and slab offsets (kernel argument, passed by value).
struct Geo {
  int nx, ny, nz, npml;
  long long vs;                // elements between volumes of a slab
  int plen;                    // nx + ny + nz (profile slab row length)
  int psi_off[PSI_COUNT + 1];  // element offsets of the psi slots
};

// CPML profile accessor: stagger s (0 = E nodes, 1 = H nodes), array q
// (0 = b, 1 = c, 2 = 1/(kappa dl)), axis a, index along the axis.
template <typename T>
__device__ __forceinline__ T prof_at(const Geo &g, const T *__restrict__ prof, int s, int q, int a,
                                     int idx) {
  const int axoff = a == 0 ? 0 : (a == 1 ? g.nx : g.nx + g.ny);
  return prof[(s * 3 + q) * g.plen + axoff + idx];
}

// local index inside a face stripe (yee_update.cl:152, 160, 168)
__device__ __forceinline__ int stripe_local(int i, int n, int N) { return i < n ? i : n + i - (N - n); }

__device__ __forceinline__ bool is_interior(const Geo &g, int i, int j, int k) {
  const int n = g.npml;
  return n > 0 && i >= n && i < g.nx - n && j >= n && j < g.ny - n && k >= n && k < g.nz - n;
}

// Element offsets of a cell inside the x-, y- and z-stripes (only meaningful
// when the cell lies in that stripe).
struct StripeIdx {
  bool inx, iny, inz;
  int qx, qy, qz;
};

__device__ __forceinline__ StripeIdx stripe_idx(const Geo &g, int i, int j, int k) {
  StripeIdx s;
  const int n = g.npml;
  s.inx = i < n || i >= g.nx - n;
  s.iny = j < n || j >= g.ny - n;
  s.inz = k < n || k >= g.nz - n;
  s.qx = s.inx ? (stripe_local(i, n, g.nx) * g.ny + j) * g.nz + k : 0;
  s.qy = s.iny ? (i * 2 * n + stripe_local(j, n, g.ny)) * g.nz + k : 0;
  s.qz = s.inz ? (i * g.ny + j) * (2 * n) + stripe_local(k, n, g.nz) : 0;
  return s;
}

// Load the six old psi values of a cell for the H (base 0) or E (base 6)
// slots, in slot order x0 x1 y0 y1 z0 z1; zero outside the stripes.
template <typename T>
__device__ __forceinline__ void psi_load(const Geo &g, const T *__restrict__ psi, int base,
                                         const StripeIdx &s, T p[6]) {
  p[0] = s.inx ? psi[g.psi_off[base + 0] + s.qx] : T(0);
  p[1] = s.inx ? psi[g.psi_off[base + 1] + s.qx] : T(0);
  p[2] = s.iny ? psi[g.psi_off[base + 2] + s.qy] : T(0);
  p[3] = s.iny ? psi[g.psi_off[base + 3] + s.qy] : T(0);
  p[4] = s.inz ? psi[g.psi_off[base + 4] + s.qz] : T(0);
  p[5] = s.inz ? psi[g.psi_off[base + 5] + s.qz] : T(0);
}

template <typename T>
__device__ __forceinline__ void psi_store(const Geo &g, T *__restrict__ psi, int base,
                                          const StripeIdx &s, const T p[6]) {
  if (s.inx) {
    psi[g.psi_off[base + 0] + s.qx] = p[0];
    psi[g.psi_off[base + 1] + s.qx] = p[1];
  }
  if (s.iny) {
    psi[g.psi_off[base + 2] + s.qy] = p[2];
    psi[g.psi_off[base + 3] + s.qy] = p[3];
  }
  if (s.inz) {
    psi[g.psi_off[base + 4] + s.qz] = p[4];
    psi[g.psi_off[base + 5] + s.qz] = p[5];
  }
}

// H at (i,j,k) from E.  E at the cell, ex_j1/ez_j1 at j+1, ex_k1/ey_k1 at k+1,
// ey_i1/ez_i1 at i+1 (values past the last index are ignored: forward
// difference 0, yee_update.cl:136-141).  p[] holds the old psi of the H
// slots (Hy_x Hz_x Hx_y Hz_y Hx_z Hy_z) and receives the new ones.
template <typename T>
__device__ __forceinline__ void h_math(const Geo &g, const T *__restrict__ prof, int i, int j,
                                       int k, const StripeIdx &s, T ex, T ey, T ez, T ex_j1,
                                       T ez_j1, T ex_k1, T ey_k1, T ey_i1, T ez_i1, T &hx, T &hy,
                                       T &hz, T p[6], T dtm_dl, T dtm) {
  if (is_interior(g, i, j, k)) {  // yee_update.cl:11-40
    const T dEz_dy = ez_j1 - ez;
    const T dEy_dz = ey_k1 - ey;
    const T dEx_dz = ex_k1 - ex;
    const T dEz_dx = ez_i1 - ez;
    const T dEy_dx = ey_i1 - ey;
    const T dEx_dy = ex_j1 - ex;
    hx -= dtm_dl * (dEz_dy - dEy_dz);
    hy -= dtm_dl * (dEx_dz - dEz_dx);
    hz -= dtm_dl * (dEy_dx - dEx_dy);
    return;
  }
  // PML shell, or npml == 0 (yee_update.cl:42-118)
  const T dEz_dy = j < g.ny - 1 ? ez_j1 - ez : T(0);
  const T dEy_dz = k < g.nz - 1 ? ey_k1 - ey : T(0);
  const T dEx_dz = k < g.nz - 1 ? ex_k1 - ex : T(0);
  const T dEz_dx = i < g.nx - 1 ? ez_i1 - ez : T(0);
  const T dEy_dx = i < g.nx - 1 ? ey_i1 - ey : T(0);
  const T dEx_dy = j < g.ny - 1 ? ex_j1 - ex : T(0);
  if (s.inx) {
    const T b = prof_at(g, prof, 1, 0, 0, i), c = prof_at(g, prof, 1, 1, 0, i);
    p[0] = b * p[0] + c * dEz_dx;  // psi_Hy_x
    p[1] = b * p[1] + c * dEy_dx;  // psi_Hz_x
  }
  if (s.iny) {
    const T b = prof_at(g, prof, 1, 0, 1, j), c = prof_at(g, prof, 1, 1, 1, j);
    p[2] = b * p[2] + c * dEz_dy;  // psi_Hx_y
    p[3] = b * p[3] + c * dEx_dy;  // psi_Hz_y
  }
  if (s.inz) {
    const T b = prof_at(g, prof, 1, 0, 2, k), c = prof_at(g, prof, 1, 1, 2, k);
    p[4] = b * p[4] + c * dEy_dz;  // psi_Hx_z
    p[5] = b * p[5] + c * dEx_dz;  // psi_Hy_z
  }
  const T ikx = prof_at(g, prof, 1, 2, 0, i), iky = prof_at(g, prof, 1, 2, 1, j),
          ikz = prof_at(g, prof, 1, 2, 2, k);
  hx -= dtm * (dEz_dy * iky + p[2] - dEy_dz * ikz - p[4]);
  hy -= dtm * (dEx_dz * ikz + p[5] - dEz_dx * ikx - p[0]);
  hz -= dtm * (dEy_dx * ikx + p[1] - dEx_dy * iky - p[3]);
}

// E at (i,j,k) from the advanced H.  H at the cell, hx_jm/hz_jm at j-1,
// hx_km/hy_km at k-1, hy_im/hz_im at i-1 (backward difference 0 on index 0,
// yee_update.cl:256-261).  p[] holds the old psi of the E slots
// (Ey_x Ez_x Ex_y Ez_y Ex_z Ey_z) and receives the new ones.
template <typename T>
__device__ __forceinline__ void e_math(const Geo &g, const T *__restrict__ prof, int i, int j,
                                       int k, const StripeIdx &s, T hx, T hy, T hz, T hx_jm,
                                       T hz_jm, T hx_km, T hy_km, T hy_im, T hz_im, T &ex, T &ey,
                                       T &ez, T cax, T cay, T caz, T cex, T cey, T cez, T p[6],
                                       T inv_dl) {
  if (is_interior(g, i, j, k)) {  // yee_update.cl:120-155
    const T dHz_dy = hz - hz_jm;
    const T dHy_dz = hy - hy_km;
    const T dHx_dz = hx - hx_km;
    const T dHz_dx = hz - hz_im;
    const T dHy_dx = hy - hy_im;
    const T dHx_dy = hx - hx_jm;
    ex = cax * ex + cex * inv_dl * (dHz_dy - dHy_dz);
    ey = cay * ey + cey * inv_dl * (dHx_dz - dHz_dx);
    ez = caz * ez + cez * inv_dl * (dHy_dx - dHx_dy);
    return;
  }
  const T dHz_dy = j > 0 ? hz - hz_jm : T(0);  // yee_update.cl:157-238
  const T dHy_dz = k > 0 ? hy - hy_km : T(0);
  const T dHx_dz = k > 0 ? hx - hx_km : T(0);
  const T dHz_dx = i > 0 ? hz - hz_im : T(0);
  const T dHy_dx = i > 0 ? hy - hy_im : T(0);
  const T dHx_dy = j > 0 ? hx - hx_jm : T(0);
  if (s.inx) {
    const T b = prof_at(g, prof, 0, 0, 0, i), c = prof_at(g, prof, 0, 1, 0, i);
    p[0] = b * p[0] + c * dHz_dx;  // psi_Ey_x
    p[1] = b * p[1] + c * dHy_dx;  // psi_Ez_x
  }
  if (s.iny) {
    const T b = prof_at(g, prof, 0, 0, 1, j), c = prof_at(g, prof, 0, 1, 1, j);
    p[2] = b * p[2] + c * dHz_dy;  // psi_Ex_y
    p[3] = b * p[3] + c * dHx_dy;  // psi_Ez_y
  }
  if (s.inz) {
    const T b = prof_at(g, prof, 0, 0, 2, k), c = prof_at(g, prof, 0, 1, 2, k);
    p[4] = b * p[4] + c * dHy_dz;  // psi_Ex_z
    p[5] = b * p[5] + c * dHx_dz;  // psi_Ey_z
  }
  const T ikx = prof_at(g, prof, 0, 2, 0, i), iky = prof_at(g, prof, 0, 2, 1, j),
          ikz = prof_at(g, prof, 0, 2, 2, k);
  ex = cax * ex + cex * (dHz_dy * iky + p[2] - dHy_dz * ikz - p[4]);
  ey = cay * ey + cey * (dHx_dz * ikz + p[5] - dHz_dx * ikx - p[0]);
  ez = caz * ez + cez * (dHy_dx * ikx + p[1] - dHx_dy * iky - p[3]);
}

// ── sources ───────────────────────────────────────────────────────────────

constexpr int kMaxSources = 16;

template <typename T>
struct Src {
  int comp, mode, taper;
  int i0, i1, j0, j1, k0, k1;        // unclipped (edge tests)
  int ci0, ci1, cj0, cj1, ck0, ck1;  // clipped to the grid
  T amp, re;
};

template <typename T>
struct SrcList {
  int n;
  Src<T> s[kMaxSources];
};

// Contribution of source s at (i,j,k) to its field component, evaluated the
// way sources.cl:7-49 does: soft adds amp; current adds (-ce*amp)*w.
template <typename T>
__device__ __forceinline__ T src_term(const Src<T> &s, int i, int j, int k, T ce, T amp) {
  if (s.mode == HFDTD_SRC_SOFT) return amp;
  T w = T(1);
  if ((s.taper & 1) && (i == s.i0 || i == s.i1 - 1)) w *= s.re;
  if ((s.taper & 2) && (j == s.j0 || j == s.j1 - 1)) w *= s.re;
  if ((s.taper & 4) && (k == s.k0 || k == s.k1 - 1)) w *= s.re;
  return -ce * amp * w;
}


// Sources of one engine step, passed by value as a kernel argument.
template <typename T>
struct SrcListP {
  const SrcList<T> &l;
  __device__ int count() const { return l.n; }
  __device__ const Src<T> &at(int q) const { return l.s[q]; }
  __device__ T amp(int q) const { return l.s[q].amp; }
};

struct NoHook {
  template <typename T>
  __device__ void operator()(int, int, int, T, T, T) {}
};

// ── fused sweep ─────────────────────────────────────────────────────────────
//
// One block owns a 64(k) x TJ(j) column tile and marches along i over
// `planes` planes.  State is ping-ponged (reads slab 0, writes slab 1), so a
// value a neighbouring block needs is never overwritten while it runs.
// Per plane:
//   0. every global load the plane needs is issued first: E(i+1), H(i), ce,
//      ca, psi of the own cell, the +1 neighbours outside the tile, and the
//      inputs of the halo cells (j0-1 row, k0-1 column);
//   1. E(i) of the tile goes to LDS (E(i) was prefetched as E(i+1) of the
//      previous plane, so every E value is read from HBM once);
//   2. H(i) of the tile is computed and written out;
//   3. H(i) on the j0-1 row (wave 0) and the k0-1 column (lane 0 of each row)
//      is recomputed, not written, so the E update needs no other block's
//      output;
//   4. sources are added to E(i) and E(i) is advanced from H(i) (LDS) and
//      H(i-1) (registers).
// The two barriers per plane wait for LDS traffic only, so global loads of
// one phase stay in flight across them.

__device__ __forceinline__ void lds_barrier() {
  __asm__ volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Body of the fused sweep for the tile (bx, by, bz) = (k block, j block,
// i chunk).  `src` provides count(), at(q) and amp(q); `hook(i, j, k, ex, ey,
// ez)` sees every advanced E value of the tile (used for on-the-fly probes).
template <typename T, int TJ, bool LOSSY, class SrcP, class Hook>
__device__ __forceinline__ void fused_tile(const Geo &g, int planes, T dtm_dl, T dtm, T inv_dl,
                                           const T *__restrict__ s0, T *__restrict__ s1,
                                           const T *__restrict__ coef,
                                           const T *__restrict__ prof,
                                           const T *__restrict__ q0, T *__restrict__ q1,
                                           const SrcP &src, Hook &hook, int bx, int by, int bz,
                                           T (&sE)[3][TJ][64], T (&sH)[3][TJ + 1][65]) {
  const int tk = threadIdx.x, tj = threadIdx.y;
  const int k0 = bx * 64, j0 = by * TJ;
  const int k = k0 + tk, j = j0 + tj;
  const int ib = bz * planes;
  const int ie = min(ib + planes, g.nx);
  const bool act = k < g.nz && j < g.ny;
  const long long vs = g.vs;
  const int nz = g.nz;
  const int sxy = g.ny * g.nz;
  const bool jn = j + 1 < g.ny, kn = k + 1 < g.nz;
  const bool row_halo = tj == 0 && j0 > 0 && k < g.nz;  // cell (i, j0-1, k)
  const bool col_halo = tk == 0 && k0 > 0 && j < g.ny;  // cell (i, j, k0-1)
  const T *Ex0 = s0, *Ey0 = s0 + vs, *Ez0 = s0 + 2 * vs;
  const T *Hx0 = s0 + 3 * vs, *Hy0 = s0 + 4 * vs, *Hz0 = s0 + 5 * vs;

  // sources whose (j, k) footprint contains this column
  unsigned smask = 0;
  if (act)
    for (int q = 0; q < src.count(); ++q) {
      const Src<T> &sq = src.at(q);
      if (j >= sq.cj0 && j < sq.cj1 && k >= sq.ck0 && k < sq.ck1) smask |= 1u << q;
    }

  int idx = ib * sxy + j * nz + k;
  T e_x = 0, e_y = 0, e_z = 0;  // E(i) of the own cell (pre-source)
  T hyp = 0, hzp = 0;           // H(i-1) of the own cell, advanced
  if (act) {
    e_x = Ex0[idx];
    e_y = Ey0[idx];
    e_z = Ez0[idx];
    if (ib > 0) {  // H(ib-1) of the own cell, recomputed (not written)
      const int q = idx - sxy;
      const StripeIdx s = stripe_idx(g, ib - 1, j, k);
      T p[6];
      psi_load(g, q0, 0, s, p);
      T hx = Hx0[q], hy = Hy0[q], hz = Hz0[q];
      h_math(g, prof, ib - 1, j, k, s, Ex0[q], Ey0[q], Ez0[q], jn ? Ex0[q + nz] : T(0),
             jn ? Ez0[q + nz] : T(0), kn ? Ex0[q + 1] : T(0), kn ? Ey0[q + 1] : T(0), e_y, e_z,
             hx, hy, hz, p, dtm_dl, dtm);
      hyp = hy;
      hzp = hz;
    }
  }

  for (int i = ib; i < ie; ++i, idx += sxy) {
    const bool in = i + 1 < g.nx;
    // ── 0. loads ──
    T n_x = 0, n_y = 0, n_z = 0, hx = 0, hy = 0, hz = 0, cx = 0, cy = 0, cz = 0;
    T ax = 1, ay = 1, az = 1;
    T ph[6] = {0, 0, 0, 0, 0, 0}, pe[6] = {0, 0, 0, 0, 0, 0};
    T xj_x = 0, xj_z = 0, xk_x = 0, xk_y = 0;  // E(i) at j+1 / k+1 outside the tile
    StripeIdx so = {false, false, false, 0, 0, 0};
    if (act) {
      if (in) {
        n_x = Ex0[idx + sxy];
        n_y = Ey0[idx + sxy];
        n_z = Ez0[idx + sxy];
      }
      hx = Hx0[idx];
      hy = Hy0[idx];
      hz = Hz0[idx];
      cx = coef[idx];
      cy = coef[vs + idx];
      cz = coef[2 * vs + idx];
      if (LOSSY) {
        ax = coef[3 * vs + idx];
        ay = coef[4 * vs + idx];
        az = coef[5 * vs + idx];
      }
      so = stripe_idx(g, i, j, k);
      if (!is_interior(g, i, j, k)) {
        psi_load(g, q0, 0, so, ph);
        psi_load(g, q0, 6, so, pe);
      }
      if (tj == TJ - 1 && jn) {
        xj_x = Ex0[idx + nz];
        xj_z = Ez0[idx + nz];
      }
      if (tk == 63 && kn) {
        xk_x = Ex0[idx + 1];
        xk_y = Ey0[idx + 1];
      }
    }
    // halo-row inputs: E, H at (i, j0-1, k), E at k+1 and at i+1
    T rx = 0, ry = 0, rz = 0, rhx = 0, rhy = 0, rhz = 0, rkx = 0, rky = 0, riy = 0, riz = 0;
    T rp[6] = {0, 0, 0, 0, 0, 0};
    StripeIdx rs = {false, false, false, 0, 0, 0};
    if (row_halo) {
      const int q = idx - nz;
      rx = Ex0[q];
      ry = Ey0[q];
      rz = Ez0[q];
      rhx = Hx0[q];
      rhy = Hy0[q];
      rhz = Hz0[q];
      if (kn) {
        rkx = Ex0[q + 1];
        rky = Ey0[q + 1];
      }
      if (in) {
        riy = Ey0[q + sxy];
        riz = Ez0[q + sxy];
      }
      rs = stripe_idx(g, i, j0 - 1, k);
      if (!is_interior(g, i, j0 - 1, k)) psi_load(g, q0, 0, rs, rp);
    }
    // halo-column inputs: E, H at (i, j, k0-1), E at j+1 and at i+1
    T cx_ = 0, cy_ = 0, cz_ = 0, chx = 0, chy = 0, chz = 0, cjx = 0, cjz = 0, ciy = 0, ciz = 0;
    T cp[6] = {0, 0, 0, 0, 0, 0};
    StripeIdx cs = {false, false, false, 0, 0, 0};
    if (col_halo) {
      const int q = idx - 1;
      cx_ = Ex0[q];
      cy_ = Ey0[q];
      cz_ = Ez0[q];
      chx = Hx0[q];
      chy = Hy0[q];
      chz = Hz0[q];
      if (jn) {
        cjx = Ex0[q + nz];
        cjz = Ez0[q + nz];
      }
      if (in) {
        ciy = Ey0[q + sxy];
        ciz = Ez0[q + sxy];
      }
      cs = stripe_idx(g, i, j, k0 - 1);
      if (!is_interior(g, i, j, k0 - 1)) psi_load(g, q0, 0, cs, cp);
    }

    // ── 1. E(i) tile to LDS ──
    sE[0][tj][tk] = e_x;
    sE[1][tj][tk] = e_y;
    sE[2][tj][tk] = e_z;
    lds_barrier();

    // ── 2. own H(i) ──
    if (act) {
      const T ex_j1 = tj + 1 < TJ ? sE[0][tj + 1][tk] : xj_x;
      const T ez_j1 = tj + 1 < TJ ? sE[2][tj + 1][tk] : xj_z;
      const T ex_k1 = tk + 1 < 64 ? sE[0][tj][tk + 1] : xk_x;
      const T ey_k1 = tk + 1 < 64 ? sE[1][tj][tk + 1] : xk_y;
      h_math(g, prof, i, j, k, so, e_x, e_y, e_z, ex_j1, ez_j1, ex_k1, ey_k1, n_y, n_z, hx, hy,
             hz, ph, dtm_dl, dtm);
      s1[3 * vs + idx] = hx;
      s1[4 * vs + idx] = hy;
      s1[5 * vs + idx] = hz;
      psi_store(g, q1, 0, so, ph);
    }
    sH[0][tj + 1][tk + 1] = hx;
    sH[1][tj + 1][tk + 1] = hy;
    sH[2][tj + 1][tk + 1] = hz;

    // ── 3. halo H ──
    if (row_halo) {  // E at (i, j0, k) is the own E(i) of this lane
      h_math(g, prof, i, j0 - 1, k, rs, rx, ry, rz, e_x, e_z, rkx, rky, riy, riz, rhx, rhy, rhz,
             rp, dtm_dl, dtm);
      sH[0][0][tk + 1] = rhx;
      sH[1][0][tk + 1] = rhy;
      sH[2][0][tk + 1] = rhz;
    }
    if (col_halo) {  // E at (i, j, k0) is the own E(i) of this lane
      h_math(g, prof, i, j, k0 - 1, cs, cx_, cy_, cz_, cjx, cjz, e_x, e_y, ciy, ciz, chx, chy,
             chz, cp, dtm_dl, dtm);
      sH[0][tj + 1][0] = chx;
      sH[1][tj + 1][0] = chy;
      sH[2][tj + 1][0] = chz;
    }
    lds_barrier();

    // ── 4. sources, then E(i) ──
    if (act) {
      if (smask) {
        for (int q = 0; q < src.count(); ++q) {
          const Src<T> &sq = src.at(q);
          if (!((smask >> q) & 1u) || i < sq.ci0 || i >= sq.ci1) continue;
          const int c = sq.comp;
          const T ce = c == 0 ? cx : (c == 1 ? cy : cz);
          const T v = src_term<T>(sq, i, j, k, ce, src.amp(q));
          if (c == 0) e_x += v;
          else if (c == 1) e_y += v;
          else e_z += v;
        }
      }
      e_math(g, prof, i, j, k, so, hx, hy, hz, sH[0][tj][tk + 1], sH[2][tj][tk + 1],
             sH[0][tj + 1][tk], sH[1][tj + 1][tk], hyp, hzp, e_x, e_y, e_z, ax, ay, az, cx, cy,
             cz, pe, inv_dl);
      s1[idx] = e_x;
      s1[vs + idx] = e_y;
      s1[2 * vs + idx] = e_z;
      psi_store(g, q1, 6, so, pe);
      hook(i, j, k, e_x, e_y, e_z);
    }
    hyp = hy;
    hzp = hz;
    e_x = n_x;
    e_y = n_y;
    e_z = n_z;
  }
}

}  // namespace hfdtd
