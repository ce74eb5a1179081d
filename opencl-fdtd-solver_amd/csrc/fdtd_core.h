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

// Load the old psi values of a cell for the H (base 0) or E (base 6) slots,
// in slot order x0 x1 y0 y1 z0 z1, for the stripes the cell lies in.  The
// caller zero-fills p[] first: values outside the stripes stay 0.  (Plain
// conditional loads, no "else" writes, so an in-flight load is never the
// target of a masked register write.)
template <typename T>
__device__ __forceinline__ void psi_load(const Geo &g, const T *__restrict__ psi, int base,
                                         const StripeIdx &s, T p[6]) {
  if (s.inx) {
    p[0] = psi[g.psi_off[base + 0] + s.qx];
    p[1] = psi[g.psi_off[base + 1] + s.qx];
  }
  if (s.iny) {
    p[2] = psi[g.psi_off[base + 2] + s.qy];
    p[3] = psi[g.psi_off[base + 3] + s.qy];
  }
  if (s.inz) {
    p[4] = psi[g.psi_off[base + 4] + s.qz];
    p[5] = psi[g.psi_off[base + 5] + s.qz];
  }
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

// CPML profile values of one stagger along one axis at one index: b, c and
// 1/(kappa dl).  Loaded once per thread (y, z) or per plane (x) by the
// callers, so the cell math below issues no memory traffic of its own.
template <typename T>
struct Prof1 {
  T b, c, ik;
};

template <typename T>
__device__ __forceinline__ Prof1<T> prof1(const Geo &g, const T *__restrict__ prof, int s, int a,
                                          int idx) {
  return Prof1<T>{prof_at(g, prof, s, 0, a, idx), prof_at(g, prof, s, 1, a, idx),
                  prof_at(g, prof, s, 2, a, idx)};
}

// H at (i,j,k) from E.  E at the cell, ex_j1/ez_j1 at j+1, ex_k1/ey_k1 at k+1,
// ey_i1/ez_i1 at i+1 (values past the last index are ignored: forward
// difference 0, yee_update.cl:136-141).  px/py/pz: H-node profiles at i/j/k.
// p[] holds the old psi of the H slots (Hy_x Hz_x Hx_y Hz_y Hx_z Hy_z) for
// the stripes the cell lies in (other entries are not read) and receives the
// new ones (0 outside the stripes).
template <typename T>
__device__ __forceinline__ void h_math(const Geo &g, int i, int j, int k, const StripeIdx &s,
                                       const Prof1<T> &px, const Prof1<T> &py,
                                       const Prof1<T> &pz, T ex, T ey, T ez, T ex_j1, T ez_j1,
                                       T ex_k1, T ey_k1, T ey_i1, T ez_i1, T &hx, T &hy, T &hz,
                                       T p[6], T dtm_dl, T dtm) {
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
  T n0 = 0, n1 = 0, n2 = 0, n3 = 0, n4 = 0, n5 = 0;
  if (s.inx) {
    n0 = px.b * p[0] + px.c * dEz_dx;  // psi_Hy_x
    n1 = px.b * p[1] + px.c * dEy_dx;  // psi_Hz_x
  }
  if (s.iny) {
    n2 = py.b * p[2] + py.c * dEz_dy;  // psi_Hx_y
    n3 = py.b * p[3] + py.c * dEx_dy;  // psi_Hz_y
  }
  if (s.inz) {
    n4 = pz.b * p[4] + pz.c * dEy_dz;  // psi_Hx_z
    n5 = pz.b * p[5] + pz.c * dEx_dz;  // psi_Hy_z
  }
  p[0] = n0; p[1] = n1; p[2] = n2; p[3] = n3; p[4] = n4; p[5] = n5;
  hx -= dtm * (dEz_dy * py.ik + n2 - dEy_dz * pz.ik - n4);
  hy -= dtm * (dEx_dz * pz.ik + n5 - dEz_dx * px.ik - n0);
  hz -= dtm * (dEy_dx * px.ik + n1 - dEx_dy * py.ik - n3);
}

// E at (i,j,k) from the advanced H.  H at the cell, hx_jm/hz_jm at j-1,
// hx_km/hy_km at k-1, hy_im/hz_im at i-1 (backward difference 0 on index 0,
// yee_update.cl:256-261).  px/py/pz: E-node profiles.  p[] as in h_math for
// the E slots (Ey_x Ez_x Ex_y Ez_y Ex_z Ey_z).
template <typename T>
__device__ __forceinline__ void e_math(const Geo &g, int i, int j, int k, const StripeIdx &s,
                                       const Prof1<T> &px, const Prof1<T> &py,
                                       const Prof1<T> &pz, T hx, T hy, T hz, T hx_jm, T hz_jm,
                                       T hx_km, T hy_km, T hy_im, T hz_im, T &ex, T &ey, T &ez,
                                       T cax, T cay, T caz, T cex, T cey, T cez, T p[6],
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
  T n0 = 0, n1 = 0, n2 = 0, n3 = 0, n4 = 0, n5 = 0;
  if (s.inx) {
    n0 = px.b * p[0] + px.c * dHz_dx;  // psi_Ey_x
    n1 = px.b * p[1] + px.c * dHy_dx;  // psi_Ez_x
  }
  if (s.iny) {
    n2 = py.b * p[2] + py.c * dHz_dy;  // psi_Ex_y
    n3 = py.b * p[3] + py.c * dHx_dy;  // psi_Ez_y
  }
  if (s.inz) {
    n4 = pz.b * p[4] + pz.c * dHy_dz;  // psi_Ex_z
    n5 = pz.b * p[5] + pz.c * dHx_dz;  // psi_Ey_z
  }
  p[0] = n0; p[1] = n1; p[2] = n2; p[3] = n3; p[4] = n4; p[5] = n5;
  ex = cax * ex + cex * (dHz_dy * py.ik + n2 - dHy_dz * pz.ik - n4);
  ey = cay * ey + cey * (dHx_dz * pz.ik + n5 - dHz_dx * px.ik - n0);
  ez = caz * ez + cez * (dHy_dx * px.ik + n1 - dHx_dy * py.ik - n3);
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

// ── fused sweep ─────────────────────────────────────────────────────────────
//
// One block owns a 64(k) x TJ(j) column tile and marches along i over
// `planes` planes.  State is ping-ponged (reads slab 0, writes slab 1), so a
// value a neighbouring block needs is never overwritten while it runs.  The
// block has TJ "row" wavefronts (one per j row of the tile) and one "halo"
// wavefront.  Per plane i:
//   0. every wavefront issues the global loads of plane i+1 (software
//      pipelining: those loads are in flight while plane i is computed);
//      row waves load E(i+2), H(i+1), ce, ca, psi of their cell and the +1
//      neighbours outside the tile, the halo wave loads the inputs of the
//      halo cells (j0-1 row, k0-1 column);
//   1. E(i) of the tile goes to LDS (E(i) was loaded as E(i+1) of the
//      previous plane, so every E value comes from HBM once);
//   2. row waves compute H(i) of the tile and write it out; the halo wave
//      recomputes H(i) on the j0-1 row and the k0-1 column (not written), so
//      the E update needs no other block's output;
//   3. E(i) is advanced from H(i) (LDS) and H(i-1) (registers).  Sources
//      are left out: source_fixup_cell redoes E on the (few) source cells
//      once the sweep is complete, so the sweep carries no per-cell branch.
// The two barriers per plane wait for LDS traffic only, so global loads stay
// in flight across them.

// XCD-aware work order.  Workgroup n runs on XCD n % X (round-robin
// dispatch), and each XCD has its own L2.  Map n to a logical tile so that
// every XCD gets a contiguous run of logical tiles: neighbouring tiles, whose
// halos overlap, then share an L2.  A bijection on [0, N) for any N.
__device__ __forceinline__ int xcd_remap(int n, int N, int X) {
  if (X <= 1) return n;
  const int q = N / X, r = N % X;
  const int x = n % X, m = n / X;
  return x * q + min(x, r) + m;
}

__device__ __forceinline__ void lds_barrier() {
  __asm__ volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Everything a row wave loads for one plane of its cell.  All loads are
// unconditional within a wave-uniform branch (addresses clamped to a valid
// cell), so no masked default write ever targets an in-flight load.
template <typename T>
struct RowPlane {
  T nx, ny, nz;    // E(i+1)
  T hx, hy, hz;    // H(i)
  T cx, cy, cz;    // ce(i)
  T ax, ay, az;    // ca(i) (lossy media)
  T ph[6], pe[6];  // old psi of the H and E slots (valid in the cell's stripes)
};

// Index of a cell inside the z-stripe, or 0 when k is not in it (a valid
// dummy address for the unconditional load).
__device__ __forceinline__ int zstripe_or0(int k, int n, int nz) {
  return k < n ? k : (k >= nz - n ? n + k - (nz - n) : 0);
}

// Loads of plane i for the cell (i, jc, kc) (coordinates already clamped to
// the grid).  zwave: some lane of this wave lies in the z-stripe.
template <typename T, bool LOSSY>
__device__ __forceinline__ void row_load(RowPlane<T> &o, const Geo &g, const T *__restrict__ s0,
                                         const T *__restrict__ coef, const T *__restrict__ q0,
                                         int i, int jc, int kc, bool zwave) {
  const long long vs = g.vs;
  const int sxy = g.ny * g.nz;
  const int idx = (i * g.ny + jc) * g.nz + kc;
  const int in1 = i + 1 < g.nx ? idx + sxy : idx;
  o.nx = s0[in1];
  o.ny = s0[vs + in1];
  o.nz = s0[2 * vs + in1];
  o.hx = s0[3 * vs + idx];
  o.hy = s0[4 * vs + idx];
  o.hz = s0[5 * vs + idx];
  o.cx = coef[idx];
  o.cy = coef[vs + idx];
  o.cz = coef[2 * vs + idx];
  if (LOSSY) {
    o.ax = coef[3 * vs + idx];
    o.ay = coef[4 * vs + idx];
    o.az = coef[5 * vs + idx];
  } else {
    o.ax = o.ay = o.az = T(1);
  }
  const int n = g.npml;
  if (n <= 0) return;
  if (i < n || i >= g.nx - n) {  // uniform per plane
    const int q = (stripe_local(i, n, g.nx) * g.ny + jc) * g.nz + kc;
    o.ph[0] = q0[g.psi_off[PSI_HY_X] + q];
    o.ph[1] = q0[g.psi_off[PSI_HZ_X] + q];
    o.pe[0] = q0[g.psi_off[PSI_EY_X] + q];
    o.pe[1] = q0[g.psi_off[PSI_EZ_X] + q];
  }
  if (jc < n || jc >= g.ny - n) {  // uniform per row wave
    const int q = (i * 2 * n + stripe_local(jc, n, g.ny)) * g.nz + kc;
    o.ph[2] = q0[g.psi_off[PSI_HX_Y] + q];
    o.ph[3] = q0[g.psi_off[PSI_HZ_Y] + q];
    o.pe[2] = q0[g.psi_off[PSI_EX_Y] + q];
    o.pe[3] = q0[g.psi_off[PSI_EZ_Y] + q];
  }
  if (zwave) {  // uniform per wave; lanes outside the stripe read a dummy
    const int q = (i * g.ny + jc) * (2 * n) + zstripe_or0(kc, n, g.nz);
    o.ph[4] = q0[g.psi_off[PSI_HX_Z] + q];
    o.ph[5] = q0[g.psi_off[PSI_HY_Z] + q];
    o.pe[4] = q0[g.psi_off[PSI_EX_Z] + q];
    o.pe[5] = q0[g.psi_off[PSI_EY_Z] + q];
  }
}

// Inputs of one recomputed halo H cell: E and H at the cell, the two E values
// at its in-plane forward neighbour outside the tile, E at i+1, old psi.
template <typename T>
struct HaloCell {
  T ex, ey, ez;
  T fa, fb;  // row halo: Ex, Ey at k+1; column halo: Ex, Ez at j+1
  T iy, iz;  // Ey, Ez at i+1
  T hx, hy, hz;
  T p[6];
};

// Everything the halo wave loads for one plane: the j0-1 row cell and the
// k0-1 column cell of this lane, and the E values of the tile's forward
// neighbours: row j0+TJ (Ex, Ez at k0+lane) and column k0+64 (Ex, Ey at
// j0+lane).
template <typename T>
struct HaloPlane {
  HaloCell<T> r, c;
  T fjx, fjz, fkx, fky;
};

// Halo cell (i, jc, kc), coordinates clamped to the grid; KIND 0 loads the
// k+1 neighbour, KIND 1 the j+1 neighbour.
template <typename T, int KIND>
__device__ __forceinline__ void halo_load(HaloCell<T> &c, const Geo &g, const T *__restrict__ s0,
                                          const T *__restrict__ q0, int i, int jc, int kc,
                                          bool zwave) {
  const long long vs = g.vs;
  const int sxy = g.ny * g.nz;
  const int q = (i * g.ny + jc) * g.nz + kc;
  c.ex = s0[q];
  c.ey = s0[vs + q];
  c.ez = s0[2 * vs + q];
  c.hx = s0[3 * vs + q];
  c.hy = s0[4 * vs + q];
  c.hz = s0[5 * vs + q];
  if (KIND == 0) {
    const int qk = kc + 1 < g.nz ? q + 1 : q;
    c.fa = s0[qk];
    c.fb = s0[vs + qk];
  } else {
    const int qj = jc + 1 < g.ny ? q + g.nz : q;
    c.fa = s0[qj];
    c.fb = s0[2 * vs + qj];
  }
  const int qi = i + 1 < g.nx ? q + sxy : q;
  c.iy = s0[vs + qi];
  c.iz = s0[2 * vs + qi];
  const int n = g.npml;
  if (n <= 0) return;
  if (i < n || i >= g.nx - n) {
    const int t = (stripe_local(i, n, g.nx) * g.ny + jc) * g.nz + kc;
    c.p[0] = q0[g.psi_off[PSI_HY_X] + t];
    c.p[1] = q0[g.psi_off[PSI_HZ_X] + t];
  }
  // y- and z-stripes: load for every lane with dummy indices outside them
  {
    const int jl = jc < n ? jc : (jc >= g.ny - n ? n + jc - (g.ny - n) : 0);
    const int t = (i * 2 * n + jl) * g.nz + kc;
    c.p[2] = q0[g.psi_off[PSI_HX_Y] + t];
    c.p[3] = q0[g.psi_off[PSI_HZ_Y] + t];
  }
  if (zwave) {
    const int t = (i * g.ny + jc) * (2 * n) + zstripe_or0(kc, n, g.nz);
    c.p[4] = q0[g.psi_off[PSI_HX_Z] + t];
    c.p[5] = q0[g.psi_off[PSI_HY_Z] + t];
  }
}

// LDS of one tile.  sE: E(i) of the TJ rows, with row TJ = row j0+TJ and
// column 64 = column k0+64 (the forward neighbours outside the tile).  sH:
// H(i) with row 0 = the j0-1 halo row and column 0 = the k0-1 halo column.
template <typename T, int TJ>
struct TileLds {
  T sE[3][TJ + 1][65];
  T sH[3][TJ + 1][65];
};

// Per-wave state of a row wave that persists from plane to plane.
template <typename T>
struct RowState {
  T ex, ey, ez;  // E(i) of the own cell (pre-source)
  T hyp, hzp;    // H(i-1) of the own cell, advanced
};

// One plane of the halo wave: prefetch plane i+1 (PF), publish the forward
// neighbours of plane i, recompute H(i) on the j0-1 row and k0-1 column.
// PF is a template argument, not a runtime test: with a load-bearing branch
// the wait-count bookkeeping at the join would have to assume no prefetch
// was issued and drain every load before the current plane could be used.
template <typename T, int TJ, bool PF>
__device__ __forceinline__ HaloPlane<T> halo_plane(
    const Geo &g, const T *__restrict__ s0, const T *__restrict__ prof, const T *__restrict__ q0,
    TileLds<T, TJ> &L, T dtm_dl, T dtm, int i, int j0, int k0, int tk, bool row_on,
    bool col_on, bool zrow, bool zcol, const Prof1<T> (&rp)[2], const Prof1<T> (&cp)[2],
    HaloPlane<T> cur) {
  const int kr = min(k0 + tk, g.nz - 1);       // row cell k (clamped)
  const int jr = max(j0 - 1, 0);               // row cell j
  const int jc = min(j0 + tk, g.ny - 1);       // column cell j (clamped)
  const int kcol = max(k0 - 1, 0);             // column cell k
  const int jf = min(j0 + TJ, g.ny - 1);       // far row
  const int kf = min(k0 + 64, g.nz - 1);       // far column
  HaloPlane<T> nxt;
  if (PF) {
    const int i1 = i + 1;
    halo_load<T, 0>(nxt.r, g, s0, q0, i1, jr, kr, zrow);
    halo_load<T, 1>(nxt.c, g, s0, q0, i1, jc, kcol, zcol);
    const long long vs = g.vs;
    const int qf = (i1 * g.ny + jf) * g.nz + kr;
    nxt.fjx = s0[qf];
    nxt.fjz = s0[2 * vs + qf];
    const int qc = (i1 * g.ny + jc) * g.nz + kf;
    nxt.fkx = s0[qc];
    nxt.fky = s0[vs + qc];
  }
  L.sE[0][TJ][tk] = cur.fjx;
  L.sE[2][TJ][tk] = cur.fjz;
  if (tk < TJ) {
    L.sE[0][tk][64] = cur.fkx;
    L.sE[1][tk][64] = cur.fky;
  }
  lds_barrier();  // E(i) tile in LDS
  const Prof1<T> px = prof1(g, prof, 1, 0, i);
  if (row_on) {  // E at (i, j0, k) comes from the tile
    HaloCell<T> &rc = cur.r;
    const StripeIdx s = stripe_idx(g, i, j0 - 1, k0 + tk);
    h_math(g, i, j0 - 1, k0 + tk, s, px, rp[0], rp[1], rc.ex, rc.ey, rc.ez, L.sE[0][0][tk],
           L.sE[2][0][tk], rc.fa, rc.fb, rc.iy, rc.iz, rc.hx, rc.hy, rc.hz, rc.p, dtm_dl, dtm);
    L.sH[0][0][tk + 1] = rc.hx;
    L.sH[1][0][tk + 1] = rc.hy;
    L.sH[2][0][tk + 1] = rc.hz;
  }
  if (col_on) {  // E at (i, j, k0) comes from the tile
    HaloCell<T> &cc = cur.c;
    const StripeIdx s = stripe_idx(g, i, j0 + tk, k0 - 1);
    h_math(g, i, j0 + tk, k0 - 1, s, px, cp[0], cp[1], cc.ex, cc.ey, cc.ez, cc.fa, cc.fb,
           L.sE[0][tk][0], L.sE[1][tk][0], cc.iy, cc.iz, cc.hx, cc.hy, cc.hz, cc.p, dtm_dl, dtm);
    L.sH[0][tk + 1][0] = cc.hx;
    L.sH[1][tk + 1][0] = cc.hy;
    L.sH[2][tk + 1][0] = cc.hz;
  }
  lds_barrier();  // halo H in LDS
  return nxt;
}

// One plane of a row wave: prefetch plane i+1 (PF, a template argument for
// the reason given at halo_plane), advance H(i) and E(i) of the own cell from
// cur.  pyz = {H y, H z, E y, E z} profiles of the own column.
template <typename T, int TJ, bool LOSSY, bool PF>
__device__ __forceinline__ RowPlane<T> row_plane(
    const Geo &g, T dtm_dl, T dtm, T inv_dl, const T *__restrict__ s0, T *__restrict__ s1,
    const T *__restrict__ coef, const T *__restrict__ prof, const T *__restrict__ q0,
    T *__restrict__ q1, TileLds<T, TJ> &L, RowState<T> &st, RowPlane<T> cur, int i, int j,
    int k, int jc, int kc, int w, int tk, bool act, bool zwave, const Prof1<T> (&pyz)[4]) {
  const long long vs = g.vs;
  const int idx = (i * g.ny + j) * g.nz + k;
  // ── 0. loads of plane i+1 ──
  RowPlane<T> nxt;
  if (PF) row_load<T, LOSSY>(nxt, g, s0, coef, q0, i + 1, jc, kc, zwave);
  // ── 1. E(i) tile to LDS ──
  L.sE[0][w][tk] = st.ex;
  L.sE[1][w][tk] = st.ey;
  L.sE[2][w][tk] = st.ez;
  lds_barrier();

  // ── 2. own H(i) ──
  const StripeIdx so = stripe_idx(g, i, j, k);
  T hx = cur.hx, hy = cur.hy, hz = cur.hz;
  if (act) {
    const Prof1<T> px = prof1(g, prof, 1, 0, i);
    h_math(g, i, j, k, so, px, pyz[0], pyz[1], st.ex, st.ey, st.ez, L.sE[0][w + 1][tk],
           L.sE[2][w + 1][tk], L.sE[0][w][tk + 1], L.sE[1][w][tk + 1], cur.ny, cur.nz, hx, hy,
           hz, cur.ph, dtm_dl, dtm);
    s1[3 * vs + idx] = hx;
    s1[4 * vs + idx] = hy;
    s1[5 * vs + idx] = hz;
    psi_store(g, q1, 0, so, cur.ph);
  }
  L.sH[0][w + 1][tk + 1] = hx;
  L.sH[1][w + 1][tk + 1] = hy;
  L.sH[2][w + 1][tk + 1] = hz;
  lds_barrier();

  // ── 3. E(i) (source cells are redone by source_fixup afterwards) ──
  if (act) {
    T ex = st.ex, ey = st.ey, ez = st.ez;
    const Prof1<T> px = prof1(g, prof, 0, 0, i);
    e_math(g, i, j, k, so, px, pyz[2], pyz[3], hx, hy, hz, L.sH[0][w][tk + 1],
           L.sH[2][w][tk + 1], L.sH[0][w + 1][tk], L.sH[1][w + 1][tk], st.hyp, st.hzp, ex, ey,
           ez, cur.ax, cur.ay, cur.az, cur.cx, cur.cy, cur.cz, cur.pe, inv_dl);
    s1[idx] = ex;
    s1[vs + idx] = ey;
    s1[2 * vs + idx] = ez;
    psi_store(g, q1, 6, so, cur.pe);
  }
  st.hyp = hy;
  st.hzp = hz;
  st.ex = cur.nx;
  st.ey = cur.ny;
  st.ez = cur.nz;
  return nxt;
}

// Body of the fused sweep for the tile (bx, by, bz) = (k block, j block,
// i chunk); the block is 64 x (TJ + 1) threads.  Sources are not applied
// here: source_fixup recomputes E on the source cells after the sweep.
// Callers that run several tiles per block separate them with lds_barrier().
template <typename T, int TJ, bool LOSSY>
__device__ __forceinline__ void fused_tile(const Geo &g, int planes, T dtm_dl, T dtm, T inv_dl,
                                           const T *__restrict__ s0, T *__restrict__ s1,
                                           const T *__restrict__ coef,
                                           const T *__restrict__ prof,
                                           const T *__restrict__ q0, T *__restrict__ q1, int bx,
                                           int by, int bz) {
  __shared__ TileLds<T, TJ> L;
  const int tk = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.y);  // one wave per row
  const int k0 = bx * 64, j0 = by * TJ;
  const int ib = bz * planes;
  const int ie = min(ib + planes, g.nx);
  const int n = g.npml;
  // does a 64-lane k range starting at kb touch the z-stripes?
  auto zwave_of = [&](int kb, int kl) { return n > 0 && (kb < n || kl >= g.nz - n); };

  if (w == TJ) {
    // ── halo wave: H on the j0-1 row (lane = k - k0) and on the k0-1 column
    // (lane = j - j0, lanes < TJ), plus the tile's forward neighbours. ──
    const bool row_on = j0 > 0 && k0 + tk < g.nz;
    const bool col_on = tk < TJ && k0 > 0 && j0 + tk < g.ny;
    const bool zrow = zwave_of(k0, min(k0 + 63, g.nz - 1));
    const bool zcol = n > 0 && (k0 - 1 < n || k0 - 1 >= g.nz - n);
    const int kr = min(k0 + tk, g.nz - 1), jr = max(j0 - 1, 0);
    const int jc = min(j0 + tk, g.ny - 1), kcol = max(k0 - 1, 0);
    const Prof1<T> rp[2] = {prof1(g, prof, 1, 1, jr), prof1(g, prof, 1, 2, kr)};
    const Prof1<T> cp[2] = {prof1(g, prof, 1, 1, jc), prof1(g, prof, 1, 2, kcol)};
    HaloPlane<T> a;
    {
      halo_load<T, 0>(a.r, g, s0, q0, ib, jr, kr, zrow);
      halo_load<T, 1>(a.c, g, s0, q0, ib, jc, kcol, zcol);
      const long long vs = g.vs;
      const int qf = (ib * g.ny + min(j0 + TJ, g.ny - 1)) * g.nz + kr;
      a.fjx = s0[qf];
      a.fjz = s0[2 * vs + qf];
      const int qc = (ib * g.ny + jc) * g.nz + min(k0 + 64, g.nz - 1);
      a.fkx = s0[qc];
      a.fky = s0[vs + qc];
    }
    // planes ib .. ie-2 prefetch their successor, plane ie-1 does not
    int i = ib;
    if ((ie - 1 - ib) & 1) {  // odd count of prefetching planes: peel one
      a = halo_plane<T, TJ, true>(g, s0, prof, q0, L, dtm_dl, dtm, i, j0, k0, tk, row_on, col_on,
                                  zrow, zcol, rp, cp, a);
      ++i;
    }
    for (; i < ie - 1; i += 2) {
      const HaloPlane<T> b = halo_plane<T, TJ, true>(g, s0, prof, q0, L, dtm_dl, dtm, i, j0, k0,
                                                     tk, row_on, col_on, zrow, zcol, rp, cp, a);
      a = halo_plane<T, TJ, true>(g, s0, prof, q0, L, dtm_dl, dtm, i + 1, j0, k0, tk, row_on,
                                  col_on, zrow, zcol, rp, cp, b);
    }
    halo_plane<T, TJ, false>(g, s0, prof, q0, L, dtm_dl, dtm, i, j0, k0, tk, row_on, col_on,
                             zrow, zcol, rp, cp, a);
    return;
  }

  // ── row waves ──
  const long long vs = g.vs;
  const int sxy = g.ny * g.nz;
  const int k = k0 + tk, j = j0 + w;
  const bool act = k < g.nz && j < g.ny;
  const int kc = min(k, g.nz - 1), jc = min(j, g.ny - 1);
  const bool zwave = zwave_of(k0, min(k0 + 63, g.nz - 1));
  const Prof1<T> pyz[4] = {prof1(g, prof, 1, 1, jc), prof1(g, prof, 1, 2, kc),
                           prof1(g, prof, 0, 1, jc), prof1(g, prof, 0, 2, kc)};

  RowState<T> st;
  {
    const int idx = (ib * g.ny + jc) * g.nz + kc;
    st.ex = s0[idx];
    st.ey = s0[vs + idx];
    st.ez = s0[2 * vs + idx];
    st.hyp = st.hzp = T(0);
    if (ib > 0 && act) {  // H(ib-1) of the own cell, recomputed (not written)
      const int q = idx - sxy;
      const bool jn = j + 1 < g.ny, kn = k + 1 < g.nz;
      const StripeIdx s = stripe_idx(g, ib - 1, j, k);
      T p[6] = {0, 0, 0, 0, 0, 0};
      if (!is_interior(g, ib - 1, j, k)) psi_load(g, q0, 0, s, p);
      T hx = s0[3 * vs + q], hy = s0[4 * vs + q], hz = s0[5 * vs + q];
      h_math(g, ib - 1, j, k, s, prof1(g, prof, 1, 0, ib - 1), pyz[0], pyz[1], s0[q],
             s0[vs + q], s0[2 * vs + q], jn ? s0[q + g.nz] : T(0), jn ? s0[2 * vs + q + g.nz] : T(0),
             kn ? s0[q + 1] : T(0), kn ? s0[vs + q + 1] : T(0), st.ey, st.ez, hx, hy, hz, p,
             dtm_dl, dtm);
      st.hyp = hy;
      st.hzp = hz;
    }
  }
  // Plane data flows by value through an unrolled pair of planes, so each
  // prefetch lands in registers that are not touched again until the next
  // plane uses them (no register copies that would wait on the loads).
  RowPlane<T> a;
  row_load<T, LOSSY>(a, g, s0, coef, q0, ib, jc, kc, zwave);
  int i = ib;
  if ((ie - 1 - ib) & 1) {  // odd count of prefetching planes: peel one
    a = row_plane<T, TJ, LOSSY, true>(g, dtm_dl, dtm, inv_dl, s0, s1, coef, prof, q0, q1, L, st,
                                      a, i, j, k, jc, kc, w, tk, act, zwave, pyz);
    ++i;
  }
  for (; i < ie - 1; i += 2) {
    const RowPlane<T> b = row_plane<T, TJ, LOSSY, true>(g, dtm_dl, dtm, inv_dl, s0, s1, coef,
                                                        prof, q0, q1, L, st, a, i, j, k, jc, kc,
                                                        w, tk, act, zwave, pyz);
    a = row_plane<T, TJ, LOSSY, true>(g, dtm_dl, dtm, inv_dl, s0, s1, coef, prof, q0, q1, L, st,
                                      b, i + 1, j, k, jc, kc, w, tk, act, zwave, pyz);
  }
  row_plane<T, TJ, LOSSY, false>(g, dtm_dl, dtm, inv_dl, s0, s1, coef, prof, q0, q1, L, st, a, i,
                                 j, k, jc, kc, w, tk, act, zwave, pyz);
}

// Source fix-up for one cell after a fused sweep.  The sweep advanced E as if
// there were no sources; here E(i,j,k) is recomputed from the untouched inputs
// (slab 0: E before the step, psi before the step) with every source of the
// list added to E first, in list order -- exactly what the reference does
// (sources between the H and E updates, engine.py:724-737) -- using the
// advanced H the sweep wrote (slab 1).  Writes E and the E psi of the cell.
template <typename T, bool LOSSY, class SrcP>
__device__ __forceinline__ void source_fixup_cell(const Geo &g, T inv_dl,
                                                  const T *__restrict__ s0, T *__restrict__ s1,
                                                  const T *__restrict__ coef,
                                                  const T *__restrict__ prof,
                                                  const T *__restrict__ q0, T *__restrict__ q1,
                                                  const SrcP &src, int i, int j, int k) {
  const long long vs = g.vs;
  const int sxy = g.ny * g.nz;
  const int idx = (i * g.ny + j) * g.nz + k;
  T ex = s0[idx], ey = s0[vs + idx], ez = s0[2 * vs + idx];
  const T cx = coef[idx], cy = coef[vs + idx], cz = coef[2 * vs + idx];
  for (int q = 0; q < src.count(); ++q) {
    const Src<T> &sq = src.at(q);
    if (i < sq.ci0 || i >= sq.ci1 || j < sq.cj0 || j >= sq.cj1 || k < sq.ck0 || k >= sq.ck1)
      continue;
    const int c = sq.comp;
    const T v = src_term<T>(sq, i, j, k, c == 0 ? cx : (c == 1 ? cy : cz), src.amp(q));
    if (c == 0) ex += v;
    else if (c == 1) ey += v;
    else ez += v;
  }
  const T *H = s1 + 3 * vs;
  const StripeIdx so = stripe_idx(g, i, j, k);
  T p[6] = {0, 0, 0, 0, 0, 0};
  psi_load(g, q0, 6, so, p);
  const T ax = LOSSY ? coef[3 * vs + idx] : T(1);
  const T ay = LOSSY ? coef[4 * vs + idx] : T(1);
  const T az = LOSSY ? coef[5 * vs + idx] : T(1);
  e_math(g, i, j, k, so, prof1(g, prof, 0, 0, i), prof1(g, prof, 0, 1, j),
         prof1(g, prof, 0, 2, k), H[idx], H[vs + idx], H[2 * vs + idx],
         j > 0 ? H[idx - g.nz] : T(0), j > 0 ? H[2 * vs + idx - g.nz] : T(0),
         k > 0 ? H[idx - 1] : T(0), k > 0 ? H[vs + idx - 1] : T(0),
         i > 0 ? H[vs + idx - sxy] : T(0), i > 0 ? H[2 * vs + idx - sxy] : T(0), ex, ey, ez, ax,
         ay, az, cx, cy, cz, p, inv_dl);
  s1[idx] = ex;
  s1[vs + idx] = ey;
  s1[2 * vs + idx] = ez;
  psi_store(g, q1, 6, so, p);
}

}  // namespace hfdtd
