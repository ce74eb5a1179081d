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
  // elements between consecutive j rows of the volumes the march sweep
  // addresses (and of its x- and y-stripe psi): nz for an engine's own
  // slabs, nz rounded up to whole 128-byte rows in the batched march's
  // padded copies
  int sk;
  long long vs;                // elements between volumes of a slab
  int plen;                    // nx + ny + nz (profile slab row length)
  int psi_off[PSI_COUNT + 1];  // element offsets of the psi slots
};

inline Geo geo_of(const Engine *e) {
  Geo g;
  g.nx = e->nx;
  g.ny = e->ny;
  g.nz = e->nz;
  g.sk = e->sk;
  g.npml = e->npml;
  g.vs = (long long)e->vstride;
  g.plen = (int)e->prof_len();
  for (int q = 0; q <= PSI_COUNT; ++q) g.psi_off[q] = (int)e->psi_off[q];
  return g;
}

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
// Interior form of h_math (no PML, every neighbour inside the grid).
template <typename T>
__device__ __forceinline__ void h_interior(T ex, T ey, T ez, T ex_j1, T ez_j1, T ex_k1, T ey_k1,
                                           T ey_i1, T ez_i1, T &hx, T &hy, T &hz, T dtm_dl) {
  const T dEz_dy = ez_j1 - ez;  // yee_update.cl:11-40
  const T dEy_dz = ey_k1 - ey;
  const T dEx_dz = ex_k1 - ex;
  const T dEz_dx = ez_i1 - ez;
  const T dEy_dx = ey_i1 - ey;
  const T dEx_dy = ex_j1 - ex;
  hx -= dtm_dl * (dEz_dy - dEy_dz);
  hy -= dtm_dl * (dEx_dz - dEz_dx);
  hz -= dtm_dl * (dEy_dx - dEx_dy);
}

// Interior form of e_math.
template <typename T>
__device__ __forceinline__ void e_interior(T hx, T hy, T hz, T hx_jm, T hz_jm, T hx_km, T hy_km,
                                           T hy_im, T hz_im, T &ex, T &ey, T &ez, T cax, T cay,
                                           T caz, T cex, T cey, T cez, T inv_dl) {
  const T dHz_dy = hz - hz_jm;  // yee_update.cl:120-155
  const T dHy_dz = hy - hy_km;
  const T dHx_dz = hx - hx_km;
  const T dHz_dx = hz - hz_im;
  const T dHy_dx = hy - hy_im;
  const T dHx_dy = hx - hx_jm;
  ex = cax * ex + cex * inv_dl * (dHz_dy - dHy_dz);
  ey = cay * ey + cey * inv_dl * (dHx_dz - dHz_dx);
  ez = caz * ez + cez * inv_dl * (dHy_dx - dHx_dy);
}

template <typename T>
__device__ __forceinline__ void h_math(const Geo &g, int i, int j, int k, const StripeIdx &s,
                                       const Prof1<T> &px, const Prof1<T> &py,
                                       const Prof1<T> &pz, T ex, T ey, T ez, T ex_j1, T ez_j1,
                                       T ex_k1, T ey_k1, T ey_i1, T ez_i1, T &hx, T &hy, T &hz,
                                       T p[6], T dtm_dl, T dtm) {
  if (is_interior(g, i, j, k)) {
    h_interior(ex, ey, ez, ex_j1, ez_j1, ex_k1, ey_k1, ey_i1, ez_i1, hx, hy, hz, dtm_dl);
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
  if (is_interior(g, i, j, k)) {
    e_interior(hx, hy, hz, hx_jm, hz_jm, hx_km, hy_km, hy_im, hz_im, ex, ey, ez, cax, cay, caz,
               cex, cey, cez, inv_dl);
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

// ── split sweeps: one cell of the H or the E half ─────────────────────────

constexpr int kFlatBlock = 256;  // threads per split-sweep block

// H(i,j,k) and its psi from E (reference yee_update.cl:11-118), in place.
template <typename T>
__device__ __forceinline__ void split_h_cell(const Geo &g, T dtm_dl, T dtm, T *__restrict__ st,
                                             const T *__restrict__ prof, T *__restrict__ psi,
                                             int i, int j, int k) {
  const T *Ex = st, *Ey = st + g.vs, *Ez = st + 2 * g.vs;
  T *Hx = st + 3 * g.vs, *Hy = st + 4 * g.vs, *Hz = st + 5 * g.vs;
  const size_t sxy = (size_t)g.ny * g.nz;
  const size_t idx = (size_t)i * sxy + (size_t)j * g.nz + k;
  const bool jn = j + 1 < g.ny, kn = k + 1 < g.nz, in = i + 1 < g.nx;
  const StripeIdx s = stripe_idx(g, i, j, k);
  T p[6] = {0, 0, 0, 0, 0, 0};
  psi_load(g, psi, 0, s, p);
  T hx = Hx[idx], hy = Hy[idx], hz = Hz[idx];
  h_math(g, i, j, k, s, prof1(g, prof, 1, 0, i), prof1(g, prof, 1, 1, j), prof1(g, prof, 1, 2, k),
         Ex[idx], Ey[idx], Ez[idx], jn ? Ex[idx + g.nz] : T(0), jn ? Ez[idx + g.nz] : T(0),
         kn ? Ex[idx + 1] : T(0), kn ? Ey[idx + 1] : T(0), in ? Ey[idx + sxy] : T(0),
         in ? Ez[idx + sxy] : T(0), hx, hy, hz, p, dtm_dl, dtm);
  Hx[idx] = hx;
  Hy[idx] = hy;
  Hz[idx] = hz;
  psi_store(g, psi, 0, s, p);
}

// E(i,j,k) and its psi from H (reference yee_update.cl:120-238), in place,
// with the cell's edge coefficients cx, cy, cz given.
// Per-cell hooks of the split E update: pre() may change E at the cell
// before its update (after it is loaded, with the cell's ce), post() sees
// the updated E.  The plain sweeps use none.
struct NoCellHook {
  template <typename T>
  __device__ __forceinline__ void pre(int, int, int, T, T, T, T &, T &, T &) const {}
  template <typename T>
  __device__ __forceinline__ void post(int, int, int, T, T, T) const {}
};

template <typename T, bool LOSSY, typename Hook = NoCellHook>
__device__ __forceinline__ void split_e_core(const Geo &g, T inv_dl, T *__restrict__ st,
                                             const T *__restrict__ coef,
                                             const T *__restrict__ prof, T *__restrict__ psi,
                                             int i, int j, int k, size_t idx, T cx, T cy, T cz,
                                             const Hook &hook = Hook()) {
  T *Ex = st, *Ey = st + g.vs, *Ez = st + 2 * g.vs;
  const T *Hx = st + 3 * g.vs, *Hy = st + 4 * g.vs, *Hz = st + 5 * g.vs;
  const size_t sxy = (size_t)g.ny * g.nz;
  const StripeIdx s = stripe_idx(g, i, j, k);
  T p[6] = {0, 0, 0, 0, 0, 0};
  psi_load(g, psi, 6, s, p);
  T ex = Ex[idx], ey = Ey[idx], ez = Ez[idx];
  hook.pre(i, j, k, cx, cy, cz, ex, ey, ez);
  const T ax = LOSSY ? coef[3 * g.vs + idx] : T(1);
  const T ay = LOSSY ? coef[4 * g.vs + idx] : T(1);
  const T az = LOSSY ? coef[5 * g.vs + idx] : T(1);
  e_math(g, i, j, k, s, prof1(g, prof, 0, 0, i), prof1(g, prof, 0, 1, j), prof1(g, prof, 0, 2, k),
         Hx[idx], Hy[idx], Hz[idx], j > 0 ? Hx[idx - g.nz] : T(0), j > 0 ? Hz[idx - g.nz] : T(0),
         k > 0 ? Hx[idx - 1] : T(0), k > 0 ? Hy[idx - 1] : T(0), i > 0 ? Hy[idx - sxy] : T(0),
         i > 0 ? Hz[idx - sxy] : T(0), ex, ey, ez, ax, ay, az, cx, cy, cz, p, inv_dl);
  Ex[idx] = ex;
  Ey[idx] = ey;
  Ez[idx] = ez;
  psi_store(g, psi, 6, s, p);
  hook.post(i, j, k, ex, ey, ez);
}

// E(i,j,k) with ce from the coefficient volumes.
template <typename T, bool LOSSY, typename Hook = NoCellHook>
__device__ __forceinline__ void split_e_cell(const Geo &g, T inv_dl, T *__restrict__ st,
                                             const T *__restrict__ coef,
                                             const T *__restrict__ prof, T *__restrict__ psi,
                                             int i, int j, int k, const Hook &hook = Hook()) {
  const size_t idx = ((size_t)i * g.ny + j) * g.nz + k;
  split_e_core<T, LOSSY, Hook>(g, inv_dl, st, coef, prof, psi, i, j, k, idx, coef[idx],
                               coef[g.vs + idx], coef[2 * g.vs + idx], hook);
}

// E(i,j,k) on a lossless FP32 grid with a material index: one byte of
// material per cell, and ce of an edge from the table of its two cells'
// materials (row: the cell's own, column: the cell at i+1, j+1 or k+1; on
// the grid's last plane, row or column the cell itself, as the edge average
// keeps the cell value there).  The table holds the values the ce volumes
// hold, so results are the same.
template <typename Hook = NoCellHook>
__device__ __forceinline__ void split_e_cell_mat(const Geo &g, float inv_dl,
                                                 float *__restrict__ st,
                                                 const float *__restrict__ coef,
                                                 const unsigned char *__restrict__ mat,
                                                 const float *__restrict__ mtab, int nmat,
                                                 const float *__restrict__ prof,
                                                 float *__restrict__ psi, int i, int j, int k,
                                                 const Hook &hook = Hook()) {
  const size_t sxy = (size_t)g.ny * g.nz;
  const size_t idx = (size_t)i * sxy + (size_t)j * g.nz + k;
  const float *tb = mtab + (unsigned)mat[idx] * (unsigned)nmat;
  const float cx = tb[mat[i + 1 < g.nx ? idx + sxy : idx]];
  const float cy = tb[mat[j + 1 < g.ny ? idx + g.nz : idx]];
  const float cz = tb[mat[k + 1 < g.nz ? idx + 1 : idx]];
  split_e_core<float, false, Hook>(g, inv_dl, st, coef, prof, psi, i, j, k, idx, cx, cy, cz, hook);
}

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
// wavefront.
//
// Every input reaches the block through LDS, filled by direct global->LDS
// loads (global_load_lds_dword: the data never occupies VGPRs while in
// flight):
//  * a ring of D plane slots, filled D-1 planes ahead of use.  The slot of
//    plane x holds (SlotLayout):
//      E   rows j0-1 .. j0+TJ     3 comps: the tile, the halo row, the far row
//      H   rows j0-1 .. j0+TJ-1   3 comps: old H of the tile and the halo row
//      ce  rows j0 .. j0+TJ-1     (then ca, for lossy media)
//      8 column vectors over rows j0-1+r, r < 16: E at k0-1 (3 comps), Ex and
//      Ey at k0+64, old H at k0-1 (3 comps)
//  * per wave, 12 rows of old psi for the next plane (stripe cells only).
// Per plane i:
//   B1  wait until psi(i) and the slots of planes i, i+1 have landed;
//       barrier; copy the own psi(i) rows to registers; issue the psi loads
//       of plane i+1, then the slot loads of plane i+D-1 (into the slot that
//       plane i-1 used);
//       row waves: H(i) of the own cell -> HBM and sH (LDS); halo wave: H(i)
//       of the j0-1 row and the k0-1 column, recomputed, not written -> sH;
//   B2  barrier; row waves: E(i) of the own cell -> HBM.
// The compiler neither waits for LDS-DMA data nor counts LDS-DMA loads in
// its own vmcnt bookkeeping, so the loop holds no compiler-tracked global
// load at all, and B1 waits explicitly: s_waitcnt vmcnt(N), N = vector
// memory operations this wave issued after the newest awaited load.  The
// wave counts its operations exactly: slot and psi loads are issued in
// wave-uniform branches, and stores are unconditional inside wave-uniform
// branches (lanes outside the grid or the stripe write into the padding
// that follows every volume and the psi slab).

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

// Hardware block n runs on XCD n % X.  xcd_runs hands XCD x runs of R
// consecutive logical blocks: within each group of X*R logical blocks, XCD x
// takes blocks x*R .. x*R+R-1, so neighbouring rows stay in one XCD's L2
// while all XCDs work in the same address region.  A bijection on [0, N):
// a last partial group keeps the identity.  R = 1 is the identity.
__device__ __forceinline__ int xcd_runs(int n, int N, int X, int R) {
  if (X <= 1 || R <= 1) return n;
  const int G = X * R;
  if (n >= N / G * G) return n;
  const int x = n % X, m = n / X;
  return (m / R) * G + x * R + m % R;
}

__device__ __forceinline__ void lds_barrier() {
  __asm__ volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

__device__ __forceinline__ void lgkm_wait() {
  __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
}

// Wait until at most n vector memory operations of this wave are pending;
// n is rounded down to one of the immediates below.
__device__ __forceinline__ void vm_wait(int n) {
  n = __builtin_amdgcn_readfirstlane(n);
#define HF_VMW(N) __asm__ volatile("s_waitcnt vmcnt(" #N ")" ::: "memory")
  if (n >= 16) {
    if (n >= 32) {
      if (n >= 48) HF_VMW(48);
      else if (n >= 40) HF_VMW(40);
      else HF_VMW(32);
    } else if (n >= 24) {
      if (n >= 28) HF_VMW(28);
      else HF_VMW(24);
    } else if (n >= 20) {
      HF_VMW(20);
    } else {
      HF_VMW(16);
    }
  } else if (n >= 8) {
    if (n >= 12) {
      if (n >= 14) HF_VMW(14);
      else HF_VMW(12);
    } else if (n >= 10) {
      HF_VMW(10);
    } else {
      HF_VMW(8);
    }
  } else if (n >= 4) {
    if (n >= 6) HF_VMW(6);
    else HF_VMW(4);
  } else if (n >= 2) {
    HF_VMW(2);
  } else {
    HF_VMW(0);
  }
#undef HF_VMW
}

// LDS-DMA through buffer resources: the per-lane part of an address is a
// 32-bit byte offset fixed for the whole tile (VGPR), the per-plane part a
// wave-uniform byte offset (SGPR), so issuing a load costs no vector ALU
// work.  Offsets are 32-bit: a slab must stay below 4 GB (checked on the
// host before a fused sweep is launched).
typedef __amdgpu_buffer_rsrc_t Rsrc;

__device__ __forceinline__ Rsrc rsrc_of(const void *base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, -1, 0x00020000);
}

// Per-lane byte offsets of the dwords a 64-element row is made of: dword
// lane + 64 p of the row (p < sizeof(T)/4) is dword (lane + 64 p) % W of
// element (lane + 64 p) / W.
template <typename T>
struct LaneOff {
  static constexpr int W = (int)(sizeof(T) / 4);
  unsigned v[W];
};

// element e of the row -> element offset, for the kinds of rows in use
enum RowKind { kRowK, kRowZ, kColJ, kColY, kColJ2 };

template <typename T, int KIND>
__device__ __forceinline__ LaneOff<T> lane_off(const Geo &g, int lane, int k0, int j0) {
  LaneOff<T> o;
  const int n = g.npml;
#pragma unroll
  for (int p = 0; p < LaneOff<T>::W; ++p) {
    const int d = lane + 64 * p;
    const int e = d / LaneOff<T>::W;
    int el;
    if (KIND == kRowK) {
      el = min(k0 + e, g.nz - 1);
    } else if (KIND == kRowZ) {
      const int k = min(k0 + e, g.nz - 1);
      el = k < n ? k : (k >= g.nz - n ? n + k - (g.nz - n) : 0);
    } else {
      const int j = min(j0 + e, g.ny - 1);
      if (KIND == kColJ) el = j * g.nz;
      else if (KIND == kColJ2) el = j * 2 * n;
      else el = (j < n ? j : (j >= g.ny - n ? n + j - (g.ny - n) : 0)) * g.nz;
    }
    o.v[p] = (unsigned)el * (unsigned)sizeof(T) + 4u * (unsigned)(d % LaneOff<T>::W);
  }
  return o;
}

// Column-vector rows t = 0, 1 of a slot: element c = 64 t + e is vector
// c / 16 (E x,y,z at k0-1 | Ex, Ey at k0+64 | H x,y,z at k0-1), grid row
// j0 - 1 + c % 16; offsets relative to the start of the plane.
template <typename T>
__device__ __forceinline__ LaneOff<T> col_off(const Geo &g, int lane, int k0, int j0, int t) {
  LaneOff<T> o;
#pragma unroll
  for (int p = 0; p < LaneOff<T>::W; ++p) {
    const int d = lane + 64 * p;
    const int c = 64 * t + d / LaneOff<T>::W;
    const int v = c >> 4;
    const int jr = min(max(j0 - 1 + (c & 15), 0), g.ny - 1);
    const int comp = v < 3 ? v : (v < 5 ? v - 3 : v - 2);
    const int kk = (v == 3 || v == 4) ? min(k0 + 64, g.nz - 1) : max(k0 - 1, 0);
    const unsigned el = (unsigned)comp * (unsigned)g.vs + (unsigned)(jr * g.nz + kk);
    o.v[p] = el * (unsigned)sizeof(T) + 4u * (unsigned)(d % LaneOff<T>::W);
  }
  return o;
}

// One 64-element row into LDS (sizeof(T)/4 loads); soff in elements.
template <typename T>
__device__ __forceinline__ int dma(Rsrc rs, T *dst, const LaneOff<T> &vo, unsigned soff) {
  const unsigned sb = soff * (unsigned)sizeof(T);
#pragma unroll
  for (int p = 0; p < LaneOff<T>::W; ++p)
    __builtin_amdgcn_raw_ptr_buffer_load_lds(
        rs, (__attribute__((address_space(3))) void *)((char *)dst + 256 * p), 4, vo.v[p], sb, 0,
        0);
  return LaneOff<T>::W;
}

// Element offset of psi slot s (slots are stored back to back, in the
// order H: Hy_x Hz_x Hx_y Hz_y Hx_z Hy_z, E: Ey_x Ez_x Ex_y Ez_y Ex_z Ey_z;
// a slot spans its stripe pair).  Cheaper to rebuild from the grid than to
// keep twelve offsets live in scalar registers.
__host__ __device__ __forceinline__ unsigned psi_slot(const Geo &g, int s) {
  const unsigned n2 = 2u * (unsigned)g.npml;
  const unsigned sx = n2 * (unsigned)(g.ny * g.sk), sy = (unsigned)g.nx * n2 * (unsigned)g.sk;
  const unsigned sz = (unsigned)(g.nx * g.ny) * n2;
  const int t = s % 6;
  const unsigned in_group = t == 0 ? 0u : t == 1 ? sx : t == 2 ? 2 * sx : t == 3 ? 2 * sx + sy
                            : t == 4 ? 2 * (sx + sy) : 2 * (sx + sy) + sz;
  return (unsigned)(s / 6) * 2 * (sx + sy + sz) + in_group;  // s == 12: the slab end
}

template <typename T>
__device__ __forceinline__ void buf_store(Rsrc rs, T v, unsigned voff, unsigned soff) {
  if constexpr (sizeof(T) == 4) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), rs, voff, soff, 0);
  } else {
    typedef unsigned u2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), rs, voff, soff, 0);
  }
}

// Resources and per-lane offsets of one tile.
template <typename T>
struct TileDma {
  Rsrc st, cf, ps;        // state slab 0, coefficients, old psi
  Rsrc wst, wps;          // state slab 1, new psi (stores)
  LaneOff<T> ok, oz;      // k runs: plain, z-stripe local
  LaneOff<T> oc[2];       // column-vector rows
  LaneOff<T> ojn, ojy, oj2;  // halo columns over j: x-, y-, z-stripe psi
};

// LDS layout of one plane slot, in elements (rows of 64).
template <typename T, int TJ, bool LOSSY>
struct SlotLayout {
  static_assert(TJ >= 1 && TJ + 2 <= 16, "column vectors hold 16 rows");
  static constexpr int kRowsE = 3 * (TJ + 2);
  static constexpr int kRowsH = 3 * (TJ + 1);
  static constexpr int kRowsC = (LOSSY ? 6 : 3) * TJ;
  static constexpr int kRows = kRowsE + kRowsH + kRowsC + 2;
  static constexpr int E = 0;
  static constexpr int H = 64 * kRowsE;
  static constexpr int C = H + 64 * kRowsH;
  static constexpr int COL = C + 64 * kRowsC;
  static constexpr int kSize = 64 * kRows;
  static constexpr int kWaves = TJ + 1;
  static constexpr unsigned long bytes(int d) {
    return ((unsigned long)d * kSize + kWaves * 12 * 64 + 3 * (TJ + 1) * 65) * sizeof(T);
  }
  // ring depth: 4 while two blocks still fit the 160 KB of a CU, else 3
  static constexpr int kDepth = bytes(4) <= 80 * 1024 ? 4 : 3;
};

template <typename T, int TJ, bool LOSSY>
struct FusedLds {
  using S = SlotLayout<T, TJ, LOSSY>;
  T ring[S::kDepth][S::kSize];
  T psi[TJ + 1][12][64];  // per wave: psi rows of the next plane
  T sH[3][TJ + 1][65];    // H(i): row 0 = j0-1 halo, column 0 = k0-1 halo
};

// Slot loads of plane x (clamped to the grid).  Slot row r of the E and H
// blocks is grid row j0 - 1 + r, of the ce / ca block grid row j0 + r.  Row
// wave w loads row r = w of every block (3 E, 3 H, 3 or 6 coefficient
// rows); the halo wave loads E rows TJ and TJ + 1, H row TJ and the two
// column-vector rows.  Rows outside the grid are clamped to its edge (their
// values are never used).  Returns the number of loads issued.
template <typename T, int TJ, bool LOSSY, bool HALO>
__device__ __forceinline__ int issue_slot(const Geo &g, const TileDma<T> &m, T *slot, int x,
                                          int j0, int w) {
  using S = SlotLayout<T, TJ, LOSSY>;
  constexpr int RE = TJ + 2, RH = TJ + 1;
  const unsigned vs = (unsigned)g.vs;
  const unsigned pbase = (unsigned)(min(x, g.nx - 1) * g.ny * g.nz);
  const int r = HALO ? TJ : w;
  const unsigned e0 = pbase + (unsigned)(min(max(j0 - 1 + r, 0), g.ny - 1) * g.nz);
  int n = 0;
#pragma unroll
  for (int c = 0; c < 3; ++c) n += dma(m.st, slot + S::E + (c * RE + r) * 64, m.ok, e0 + c * vs);
#pragma unroll
  for (int c = 0; c < 3; ++c)
    n += dma(m.st, slot + S::H + (c * RH + r) * 64, m.ok, e0 + (3 + c) * vs);
  if (HALO) {
    const unsigned e1 = pbase + (unsigned)(min(j0 + TJ, g.ny - 1) * g.nz);
#pragma unroll
    for (int c = 0; c < 3; ++c)
      n += dma(m.st, slot + S::E + (c * RE + TJ + 1) * 64, m.ok, e1 + c * vs);
    n += dma(m.st, slot + S::COL, m.oc[0], pbase);
    n += dma(m.st, slot + S::COL + 64, m.oc[1], pbase);
  } else {
    const unsigned c0 = pbase + (unsigned)(min(j0 + w, g.ny - 1) * g.nz);
#pragma unroll
    for (int c = 0; c < (LOSSY ? 6 : 3); ++c)
      n += dma(m.cf, slot + S::C + (c * TJ + w) * 64, m.ok, c0 + c * vs);
  }
  return n;
}

// Slot of plane x + the ring bookkeeping of one wave.
template <int D>
struct Ring {
  int issued;       // vector memory operations issued so far
  int mark[D - 1];  // `issued` right after the slot loads of planes i+1 ..
  int mark_psi;     // `issued` right after the psi loads of plane i
  int si;           // slot of plane i
  __device__ __forceinline__ void next() {
    si = si + 1 == D ? 0 : si + 1;
#pragma unroll
    for (int t = 0; t + 1 < D - 1; ++t) mark[t] = mark[t + 1];
  }
  // pending operations allowed while psi(i) and the slots <= i+1 are needed
  __device__ __forceinline__ int slack() const { return issued - max(mark[0], mark_psi); }
};

// Kernel arguments of one tile.
template <typename T>
struct TileArgs {
  Geo g;
  T dtm_dl, dtm, inv_dl;
  const T *s0;
  T *s1;
  const T *coef, *prof, *q0;
  T *q1;
  int j0, k0, w, tk;
  int last;  // last plane whose slot is needed
  int ie;    // end of the i chunk
  TileDma<T> m;
};

// Per-wave constants of a row wave.
template <typename T>
struct RowWave {
  int jc, kc;        // own cell, clamped to the grid
  unsigned own;      // byte offset of the own cell within its plane
  unsigned junk;     // byte offset of this lane's volume padding slot
  bool act;          // own cell inside the grid
  bool ywave, zwave; // the row lies in a y-stripe / the tile meets a z-stripe
  bool zin;          // own cell in a z-stripe
  Prof1<T> pyz[4];   // H y, H z, E y, E z profiles of the own column
  T hyp, hzp;        // H(i-1) of the own cell, advanced
};

// Per-wave constants of the halo wave.
template <typename T>
struct HaloWave {
  int jr, kr;      // row-halo cell (j0-1, k0+lane), clamped
  int kcol;        // column-halo cells (j0+lane, k0-1): k, clamped
  bool row_on, col_on, yrow, zrow, ycol, zcol;
  Prof1<T> rp[2], cp[2];  // H y, z profiles of the two cells
};

// local index inside the stripes of an axis of length N, 0 outside them
__device__ __forceinline__ int stripe_local_or0(int i, int n, int N) {
  return i < n ? i : (i >= N - n ? n + i - (N - n) : 0);
}

// psi rows of plane x for the own cell column of a row wave: rows 0..5 the
// H slots, 6..11 the E slots, stripes the wave touches only.
template <typename T>
__device__ __forceinline__ int issue_psi_row(const Geo &g, const TileDma<T> &m, T (*stage)[64],
                                             int x, const RowWave<T> &r) {
  const int n = g.npml;
  int c = 0;
  if (n <= 0) return 0;
  if (x < n || x >= g.nx - n) {
    const unsigned off = (unsigned)((stripe_local(x, n, g.nx) * g.ny + r.jc) * g.nz);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int s = t < 2 ? t : t + 4;  // Hy_x Hz_x Ey_x Ez_x
      c += dma(m.ps, stage[s], m.ok, psi_slot(g, s) + off);
    }
  }
  if (r.ywave) {
    const unsigned off = (unsigned)((x * 2 * n + stripe_local(r.jc, n, g.ny)) * g.nz);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int s = t < 2 ? 2 + t : 6 + t;  // Hx_y Hz_y Ex_y Ez_y
      c += dma(m.ps, stage[s], m.ok, psi_slot(g, s) + off);
    }
  }
  if (r.zwave) {
    const unsigned off = (unsigned)((x * g.ny + r.jc) * (2 * n));
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int s = t < 2 ? 4 + t : 8 + t;  // Hx_z Hy_z Ex_z Ey_z
      c += dma(m.ps, stage[s], m.oz, psi_slot(g, s) + off);
    }
  }
  return c;
}

// psi rows of plane x for the halo cells: rows 0..5 the H slots of the
// row-halo cell of each lane, rows 6..11 those of the column-halo cells.
template <typename T>
__device__ __forceinline__ int issue_psi_halo(const Geo &g, const TileDma<T> &m, T (*stage)[64],
                                              int x, const HaloWave<T> &h) {
  const int n = g.npml;
  int c = 0;
  if (n <= 0) return 0;
  if (x < n || x >= g.nx - n) {
    const int il = stripe_local(x, n, g.nx);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      c += dma(m.ps, stage[s], m.ok, psi_slot(g, s) + (unsigned)((il * g.ny + h.jr) * g.nz));
      c += dma(m.ps, stage[6 + s], m.ojn, psi_slot(g, s) + (unsigned)(il * g.ny * g.nz + h.kcol));
    }
  }
  if (h.yrow) {
    const int off = (x * 2 * n + stripe_local(h.jr, n, g.ny)) * g.nz;
#pragma unroll
    for (int s = 2; s < 4; ++s) c += dma(m.ps, stage[s], m.ok, psi_slot(g, s) + (unsigned)off);
  }
  if (h.ycol) {
#pragma unroll
    for (int s = 2; s < 4; ++s)
      c += dma(m.ps, stage[6 + s], m.ojy, psi_slot(g, s) + (unsigned)(x * 2 * n * g.nz + h.kcol));
  }
  if (h.zrow) {
    const int off = (x * g.ny + h.jr) * (2 * n);
#pragma unroll
    for (int s = 4; s < 6; ++s) c += dma(m.ps, stage[s], m.oz, psi_slot(g, s) + (unsigned)off);
  }
  if (h.zcol) {
    const int kl = stripe_local_or0(h.kcol, n, g.nz);
#pragma unroll
    for (int s = 4; s < 6; ++s)
      c += dma(m.ps, stage[6 + s], m.oj2, psi_slot(g, s) + (unsigned)(x * g.ny * 2 * n + kl));
  }
  return c;
}

// New psi of a row-wave cell, stored for every lane inside wave-uniform
// branches (lanes outside the grid or the stripe write to the slab padding).
// Returns the number of stores issued.
// Whether a psi slab is small enough (below 2 GB with its padding) for its
// store resource to carry the true size, so out-of-range stores are dropped.
template <typename T>
__device__ __forceinline__ bool psi_drop_ok(const Geo &g) {
  return ((unsigned long long)psi_slot(g, PSI_COUNT) + 64ull) * sizeof(T) < 0x80000000ull;
}

template <typename T>
__device__ __forceinline__ Rsrc psi_store_rsrc(const Geo &g, T *q1) {
  const int bytes = (int)((psi_slot(g, PSI_COUNT) + 64u) * (unsigned)sizeof(T));
  return psi_drop_ok<T>(g) ? __builtin_amdgcn_make_buffer_rsrc(q1, (short)0, bytes, 0x00020000)
                           : rsrc_of(q1);
}

template <typename T>
__device__ __forceinline__ int store_psi_row(const Geo &g, const TileDma<T> &m, int base,
                                             const StripeIdx &so, bool xst, const RowWave<T> &r,
                                             int tk, const T p[6]) {
  const unsigned es = (unsigned)sizeof(T);
  // lanes that must not store: past the end of the psi resource (sized when
  // the slab is below 2 GB: the hardware drops the store), else this lane's
  // slot in the slab padding; one shared padding address for every wave of
  // the grid would pile all their stores onto one line
  const unsigned junk = psi_drop_ok<T>(g) ? 0x80000000u : (psi_slot(g, PSI_COUNT) + (unsigned)tk) * es;
  int c = 0;
  if (xst) {
    buf_store(m.wps, p[0], r.act ? (psi_slot(g, base) + so.qx) * es : junk, 0u);
    buf_store(m.wps, p[1], r.act ? (psi_slot(g, base + 1) + so.qx) * es : junk, 0u);
    c += 2;
  }
  if (r.ywave) {
    buf_store(m.wps, p[2], r.act ? (psi_slot(g, base + 2) + so.qy) * es : junk, 0u);
    buf_store(m.wps, p[3], r.act ? (psi_slot(g, base + 3) + so.qy) * es : junk, 0u);
    c += 2;
  }
  if (r.zwave) {
    const bool ok = r.act && r.zin;
    buf_store(m.wps, p[4], ok ? (psi_slot(g, base + 4) + so.qz) * es : junk, 0u);
    buf_store(m.wps, p[5], ok ? (psi_slot(g, base + 5) + so.qz) * es : junk, 0u);
    c += 2;
  }
  return c;
}

// Issue the loads of plane i's successors: psi of plane i+1 (own staging,
// after this wave's reads of it are complete), then the slot of plane i+D-1.
template <typename T, int TJ, bool LOSSY, bool HALO, class W>
__device__ __forceinline__ void plane_loads(const TileArgs<T> &a, FusedLds<T, TJ, LOSSY> &L,
                                            Ring<SlotLayout<T, TJ, LOSSY>::kDepth> &p,
                                            const W &wv, int i) {
  using S = SlotLayout<T, TJ, LOSSY>;
  constexpr int D = S::kDepth;
  lgkm_wait();  // the staging reads of this wave have completed
  if (i + 1 < a.ie) {
    if constexpr (HALO)
      p.issued += issue_psi_halo(a.g, a.m, L.psi[a.w], i + 1, wv);
    else
      p.issued += issue_psi_row(a.g, a.m, L.psi[a.w], i + 1, wv);
  }
  p.mark_psi = p.issued;
  if (i + D - 1 <= a.last) {
    const int s = p.si == 0 ? D - 1 : p.si - 1;
    p.issued += issue_slot<T, TJ, LOSSY, HALO>(a.g, a.m, L.ring[s], i + D - 1, a.j0, a.w);
  }
  p.mark[D - 2] = p.issued;
}

// Prologue after the psi loads of plane ib: the slots of planes ib .. ib+D-2.
template <typename T, int TJ, bool LOSSY, bool HALO>
__device__ __forceinline__ void prologue_slots(const TileArgs<T> &a, FusedLds<T, TJ, LOSSY> &L,
                                               Ring<SlotLayout<T, TJ, LOSSY>::kDepth> &p,
                                               int ib) {
  constexpr int D = SlotLayout<T, TJ, LOSSY>::kDepth;
  p.mark_psi = p.issued;
  p.issued += issue_slot<T, TJ, LOSSY, HALO>(a.g, a.m, L.ring[0], ib, a.j0, a.w);
#pragma unroll
  for (int t = 0; t + 2 < D; ++t) {
    if (ib + 1 + t <= a.last)
      p.issued +=
          issue_slot<T, TJ, LOSSY, HALO>(a.g, a.m, L.ring[t + 1], ib + 1 + t, a.j0, a.w);
    p.mark[t] = p.issued;
  }
}

// One plane of a row wave.
template <typename T, int TJ, bool LOSSY>
__device__ __forceinline__ void row_plane(const TileArgs<T> &a, FusedLds<T, TJ, LOSSY> &L,
                                          Ring<SlotLayout<T, TJ, LOSSY>::kDepth> &p,
                                          RowWave<T> &r, int i) {
  using S = SlotLayout<T, TJ, LOSSY>;
  constexpr int D = S::kDepth;
  constexpr int RE = TJ + 2, RH = TJ + 1;
  const Geo &g = a.g;
  const long long vs = g.vs;
  const int w = a.w, tk = a.tk;
  vm_wait(p.slack());
  lds_barrier();  // B1
  T ph[6], pe[6];
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    ph[s] = L.psi[w][s][tk];
    pe[s] = L.psi[w][6 + s][tk];
  }
  plane_loads<T, TJ, LOSSY, false>(a, L, p, r, i);

  const T *X = L.ring[p.si];
  const T *Y = L.ring[p.si + 1 == D ? 0 : p.si + 1];
  const int rw = w + 1;  // row of the own cell in the E and H blocks
  const T ex = X[S::E + (0 * RE + rw) * 64 + tk];
  const T ey = X[S::E + (1 * RE + rw) * 64 + tk];
  const T ez = X[S::E + (2 * RE + rw) * 64 + tk];
  T hx = X[S::H + (0 * RH + rw) * 64 + tk];
  T hy = X[S::H + (1 * RH + rw) * 64 + tk];
  T hz = X[S::H + (2 * RH + rw) * 64 + tk];
  const StripeIdx so = stripe_idx(g, i, r.jc, r.kc);
  const bool xst = g.npml > 0 && (i < g.npml || i >= g.nx - g.npml);
  h_math(g, i, r.jc, r.kc, so, prof1(g, a.prof, 1, 0, i), r.pyz[0], r.pyz[1], ex, ey, ez,
         X[S::E + (0 * RE + rw + 1) * 64 + tk], X[S::E + (2 * RE + rw + 1) * 64 + tk],
         X[tk < 63 ? S::E + (0 * RE + rw) * 64 + tk + 1 : S::COL + 3 * 16 + rw],
         X[tk < 63 ? S::E + (1 * RE + rw) * 64 + tk + 1 : S::COL + 4 * 16 + rw],
         Y[S::E + (1 * RE + rw) * 64 + tk], Y[S::E + (2 * RE + rw) * 64 + tk], hx, hy, hz, ph,
         a.dtm_dl, a.dtm);
  const unsigned es = (unsigned)sizeof(T), vsb = (unsigned)vs * es;
  const unsigned o = r.act ? r.own + (unsigned)(i * g.ny * g.nz) * es : r.junk;
  buf_store(a.m.wst, hx, o, 3 * vsb);
  buf_store(a.m.wst, hy, o, 4 * vsb);
  buf_store(a.m.wst, hz, o, 5 * vsb);
  p.issued += 3;
  if (g.npml > 0) p.issued += store_psi_row(g, a.m, 0, so, xst, r, tk, ph);
  L.sH[0][w + 1][tk + 1] = hx;
  L.sH[1][w + 1][tk + 1] = hy;
  L.sH[2][w + 1][tk + 1] = hz;
  lds_barrier();  // B2

  T ex2 = ex, ey2 = ey, ez2 = ez;
  const T ax = LOSSY ? X[S::C + (3 * TJ + w) * 64 + tk] : T(1);
  const T ay = LOSSY ? X[S::C + (4 * TJ + w) * 64 + tk] : T(1);
  const T az = LOSSY ? X[S::C + (5 * TJ + w) * 64 + tk] : T(1);
  e_math(g, i, r.jc, r.kc, so, prof1(g, a.prof, 0, 0, i), r.pyz[2], r.pyz[3], hx, hy, hz,
         L.sH[0][w][tk + 1], L.sH[2][w][tk + 1], L.sH[0][w + 1][tk], L.sH[1][w + 1][tk], r.hyp,
         r.hzp, ex2, ey2, ez2, ax, ay, az, X[S::C + (0 * TJ + w) * 64 + tk],
         X[S::C + (1 * TJ + w) * 64 + tk], X[S::C + (2 * TJ + w) * 64 + tk], pe, a.inv_dl);
  buf_store(a.m.wst, ex2, o, 0u);
  buf_store(a.m.wst, ey2, o, vsb);
  buf_store(a.m.wst, ez2, o, 2 * vsb);
  p.issued += 3;
  if (g.npml > 0) p.issued += store_psi_row(g, a.m, 6, so, xst, r, tk, pe);
  r.hyp = hy;
  r.hzp = hz;
  p.next();
}

// One plane of the halo wave.
template <typename T, int TJ, bool LOSSY>
__device__ __forceinline__ void halo_plane(const TileArgs<T> &a, FusedLds<T, TJ, LOSSY> &L,
                                           Ring<SlotLayout<T, TJ, LOSSY>::kDepth> &p,
                                           const HaloWave<T> &h, int i) {
  using S = SlotLayout<T, TJ, LOSSY>;
  constexpr int D = S::kDepth;
  constexpr int RE = TJ + 2, RH = TJ + 1;
  const Geo &g = a.g;
  const int w = a.w, tk = a.tk;
  vm_wait(p.slack());
  lds_barrier();  // B1
  T pr[6], pc[6];
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    pr[s] = L.psi[w][s][tk];
    pc[s] = L.psi[w][6 + s][tk];
  }
  plane_loads<T, TJ, LOSSY, true>(a, L, p, h, i);

  const T *X = L.ring[p.si];
  const T *Y = L.ring[p.si + 1 == D ? 0 : p.si + 1];
  const Prof1<T> px = prof1(g, a.prof, 1, 0, i);
  if (h.row_on) {  // cell (i, j0-1, k0+lane): E and H rows 0, far column row 0
    T hx = X[S::H + (0 * RH) * 64 + tk];
    T hy = X[S::H + (1 * RH) * 64 + tk];
    T hz = X[S::H + (2 * RH) * 64 + tk];
    const StripeIdx s = stripe_idx(g, i, h.jr, h.kr);
    h_math(g, i, h.jr, h.kr, s, px, h.rp[0], h.rp[1], X[S::E + (0 * RE) * 64 + tk],
           X[S::E + (1 * RE) * 64 + tk], X[S::E + (2 * RE) * 64 + tk],
           X[S::E + (0 * RE + 1) * 64 + tk], X[S::E + (2 * RE + 1) * 64 + tk],
           X[tk < 63 ? S::E + (0 * RE) * 64 + tk + 1 : S::COL + 3 * 16],
           X[tk < 63 ? S::E + (1 * RE) * 64 + tk + 1 : S::COL + 4 * 16],
           Y[S::E + (1 * RE) * 64 + tk], Y[S::E + (2 * RE) * 64 + tk], hx, hy, hz, pr, a.dtm_dl,
           a.dtm);
    L.sH[0][0][tk + 1] = hx;
    L.sH[1][0][tk + 1] = hy;
    L.sH[2][0][tk + 1] = hz;
  }
  if (h.col_on) {  // cell (i, j0+lane, k0-1): column vectors, row lane+1
    const int rr = tk + 1;
    const int jc = a.j0 + tk;
    T hx = X[S::COL + 5 * 16 + rr];
    T hy = X[S::COL + 6 * 16 + rr];
    T hz = X[S::COL + 7 * 16 + rr];
    const StripeIdx s = stripe_idx(g, i, jc, h.kcol);
    h_math(g, i, jc, h.kcol, s, px, h.cp[0], h.cp[1], X[S::COL + rr], X[S::COL + 16 + rr],
           X[S::COL + 32 + rr], X[S::COL + rr + 1], X[S::COL + 32 + rr + 1],
           X[S::E + (0 * RE + rr) * 64], X[S::E + (1 * RE + rr) * 64], Y[S::COL + 16 + rr],
           Y[S::COL + 32 + rr], hx, hy, hz, pc, a.dtm_dl, a.dtm);
    L.sH[0][rr][0] = hx;
    L.sH[1][rr][0] = hy;
    L.sH[2][rr][0] = hz;
  }
  lds_barrier();  // B2
  p.next();
}

// Body of the fused sweep for the tile (bx, by, bz) = (k block, j block,
// i chunk); the block is 64 x (TJ + 1) threads.  Sources are not applied
// here: source_fixup_cell redoes E on the source cells after the sweep.
// When a tile ends, every load it issued has landed (its last plane waits
// for all of them), so callers running several tiles per block only need
// an lds_barrier() between tiles.
template <typename T, int TJ, bool LOSSY>
__device__ __forceinline__ void fused_tile(const Geo &g, int planes, T dtm_dl, T dtm, T inv_dl,
                                           const T *__restrict__ s0, T *__restrict__ s1,
                                           const T *__restrict__ coef,
                                           const T *__restrict__ prof,
                                           const T *__restrict__ q0, T *__restrict__ q1, int bx,
                                           int by, int bz) {
  using S = SlotLayout<T, TJ, LOSSY>;
  constexpr int D = S::kDepth;
  __shared__ FusedLds<T, TJ, LOSSY> L;
  TileArgs<T> a;
  a.g = g;
  a.dtm_dl = dtm_dl;
  a.dtm = dtm;
  a.inv_dl = inv_dl;
  a.s0 = s0;
  a.s1 = s1;
  a.coef = coef;
  a.prof = prof;
  a.q0 = q0;
  a.q1 = q1;
  a.tk = threadIdx.x;
  a.w = __builtin_amdgcn_readfirstlane(threadIdx.y);
  a.k0 = bx * 64;
  a.j0 = by * TJ;
  const int tk = a.tk, w = a.w, k0 = a.k0, j0 = a.j0;
  const int ib = bz * planes;
  a.ie = min(ib + planes, g.nx);
  a.last = a.ie;  // E(i+1) of the last plane (clamped to the grid when ie == nx)
  const int n = g.npml;
  const bool zt = n > 0 && (k0 < n || min(k0 + 63, g.nz - 1) >= g.nz - n);  // tile meets z-stripes
  a.m.st = rsrc_of(s0);
  a.m.cf = rsrc_of(coef);
  a.m.ps = rsrc_of(q0);
  a.m.wst = rsrc_of(s1);
  a.m.wps = psi_store_rsrc<T>(g, q1);
  a.m.ok = lane_off<T, kRowK>(g, tk, k0, j0);
  a.m.oz = lane_off<T, kRowZ>(g, tk, k0, j0);
  a.m.oc[0] = col_off<T>(g, tk, k0, j0, 0);
  a.m.oc[1] = col_off<T>(g, tk, k0, j0, 1);

  Ring<D> p;
  p.issued = 0;
  p.si = 0;


  if (w == TJ) {
    // ── halo wave ──
    HaloWave<T> h;
    h.jr = max(j0 - 1, 0);
    h.kr = min(k0 + tk, g.nz - 1);
    h.kcol = max(k0 - 1, 0);
    h.row_on = j0 > 0 && k0 + tk < g.nz;
    h.col_on = tk < TJ && k0 > 0 && j0 + tk < g.ny;
    h.yrow = n > 0 && (h.jr < n || h.jr >= g.ny - n);
    h.zrow = zt;
    h.ycol = n > 0 && (j0 < n || min(j0 + TJ - 1, g.ny - 1) >= g.ny - n);
    h.zcol = n > 0 && (h.kcol < n || h.kcol >= g.nz - n);
    h.rp[0] = prof1(g, prof, 1, 1, h.jr);
    h.rp[1] = prof1(g, prof, 1, 2, h.kr);
    h.cp[0] = prof1(g, prof, 1, 1, min(j0 + tk, g.ny - 1));
    h.cp[1] = prof1(g, prof, 1, 2, h.kcol);
    a.m.ojn = lane_off<T, kColJ>(g, tk, k0, j0);
    a.m.ojy = lane_off<T, kColY>(g, tk, k0, j0);
    a.m.oj2 = lane_off<T, kColJ2>(g, tk, k0, j0);
    p.issued += issue_psi_halo(g, a.m, L.psi[w], ib, h);
    prologue_slots<T, TJ, LOSSY, true>(a, L, p, ib);
    for (int i = ib; i < a.ie; ++i) halo_plane<T, TJ, LOSSY>(a, L, p, h, i);
    return;
  }

  // ── row waves ──
  const long long vs = g.vs;
  const int k = k0 + tk, j = j0 + w;
  RowWave<T> r;
  r.act = k < g.nz && j < g.ny;
  r.kc = min(k, g.nz - 1);
  r.jc = min(j, g.ny - 1);
  r.own = (unsigned)(r.jc * g.nz + r.kc) * (unsigned)sizeof(T);
  r.junk = (unsigned)(g.nx * g.ny * g.nz + tk) * (unsigned)sizeof(T);
  r.ywave = n > 0 && (r.jc < n || r.jc >= g.ny - n);
  r.zwave = zt;
  r.zin = n > 0 && (r.kc < n || r.kc >= g.nz - n);
  r.pyz[0] = prof1(g, prof, 1, 1, r.jc);
  r.pyz[1] = prof1(g, prof, 1, 2, r.kc);
  r.pyz[2] = prof1(g, prof, 0, 1, r.jc);
  r.pyz[3] = prof1(g, prof, 0, 2, r.kc);
  r.hyp = r.hzp = T(0);

  p.issued += issue_psi_row(g, a.m, L.psi[w], ib, r);
  prologue_slots<T, TJ, LOSSY, false>(a, L, p, ib);
  if (ib > 0 && r.act) {
    // H(ib-1) of the own cell, recomputed (not written) for E(ib)
    const int sxy = g.ny * g.nz;
    const int q = ((ib - 1) * g.ny + r.jc) * g.nz + r.kc;
    const bool jn = r.jc + 1 < g.ny, kn = r.kc + 1 < g.nz;
    const StripeIdx s = stripe_idx(g, ib - 1, r.jc, r.kc);
    T pm[6] = {0, 0, 0, 0, 0, 0};
    if (!is_interior(g, ib - 1, r.jc, r.kc)) psi_load(g, q0, 0, s, pm);
    T hx = T(0), hy = s0[4 * vs + q], hz = s0[5 * vs + q];
    h_math(g, ib - 1, r.jc, r.kc, s, prof1(g, prof, 1, 0, ib - 1), r.pyz[0], r.pyz[1], s0[q],
           s0[vs + q], s0[2 * vs + q], jn ? s0[q + g.nz] : T(0),
           jn ? s0[2 * vs + q + g.nz] : T(0), kn ? s0[q + 1] : T(0), kn ? s0[vs + q + 1] : T(0),
           s0[vs + q + sxy], s0[2 * vs + q + sxy], hx, hy, hz, pm, dtm_dl, dtm);
    r.hyp = hy;
    r.hzp = hz;
  }
  for (int i = ib; i < a.ie; ++i) row_plane<T, TJ, LOSSY>(a, L, p, r, i);
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
