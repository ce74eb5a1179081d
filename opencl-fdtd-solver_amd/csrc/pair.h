// Two Yee steps per launch ("pair" sweep): fields, coefficients and psi are
// read and written once for two time steps.
//
// A block of NW wavefronts loads a column tile of 64 (k) x NW*RW (j) cells,
// lane = k - k0 + 2, wave w holding rows j0 - 2 + RW*w ... + RW - 1 in
// registers, and marches along i.  On the whole tile it advances the old
// fields (step 1: H1, then E1) and, behind them, the step-1 fields (step 2:
// H2, then E2).  Each sub-update shrinks the valid region by one cell on the
// side its stencil reaches (H reads E at +1, E reads H at -1), so after the
// four of them the two halo rows and columns on each side are spent and the
// core of 60 (k) x NW*RW-4 (j) cells holds two exact steps; only the core is
// stored.  Along i the same holds with a lag: at front plane p the block
// computes H1(p), E1(p), H2(p-1) and E2(p-2), and a chunk of planes [i0, i1)
// starts its front at i0 - 2.
//
// Neighbours: i from registers (carries of the previous planes); k from the
// next / previous lane (DPP wavefront shifts); j from the same lane's next /
// previous row, or, for a wave's first and last rows, through LDS (two
// barriers per plane).  Every cell of the tile, halo included, goes through
// h_math / e_math (h_interior / e_interior on tiles wholly inside the CPML-
// free interior) with its own indices, its psi and its sources, so a cell
// stored by this sweep equals the one two single steps produce, bit for
// bit.  Values of halo cells outside the grid are never used: h_math and
// e_math ignore neighbours past the grid's faces.  Sub-updates whose results
// no stored cell depends on (outer rows, the first and last fronts of a
// chunk) are skipped.
//
// psi: the old psi of a cell is loaded with its plane, updated by step 1,
// carried in registers to step 2 (H: one plane, E: two planes) and stored
// once.  As in the march sweep a block takes the variant with only the
// stripes its tile (halo included) meets.
//
// Watched cells: the state between the two steps never reaches the field
// slab.  Core cells inside up to kPairWatch boxes also write the boxes'
// components of their step-1 H and E to a shadow slab (same layout), for
// observers of that step.
#pragma once

#include "march.h"

namespace hfdtd {

constexpr int kPairCols = 60;   // core columns of a tile: 64 lanes less two halo lanes a side
constexpr int kPairWatch = 8;   // watched boxes
constexpr int kPairMaxPlanes = 1024;  // i planes per block at most

template <typename T>
struct PairLoad {
  T ex, ey, ez;  // E(p+1), old
  T hx, hy, hz;  // H(p), old
  T cx, cy, cz;  // ce(p)
  T ax, ay, az;  // ca(p) (lossy grids)
  T ph[6];       // old psi of the H slots at plane p (the block's stripes)
  T pe[6];       // old psi of the E slots
};

// Sources of the two steps (list 0: step 1, list 1: step 2), at most
// kMarchSources each; copied to LDS by the block.
template <typename T>
struct PairSources {
  int n[2];
  Src<T> s[2][kMarchSources];
};

struct PairWatch {
  int n;
  int b[kPairWatch][6];  // i0, i1, j0, j1, k0, k1 (half-open)
  int c[kPairWatch];     // components recorded (bit c: Ex Ey Ez Hx Hy Hz)
};

// Rows exchanged between wavefronts; wave w writes slot w + 1, reads its
// neighbours in slots w and w + 2 (slots 0 and NW + 1 are never written:
// their values reach only halo cells whose results are discarded).
template <typename T, int NW>
struct PairLds {
  T e0[NW + 2][2][64];     // old E (Ex, Ez) of the front plane, first row of each wave
  T h1[NW + 2][2][64];     // H1 (Hx, Hz), last row
  T e1[2][NW + 2][2][64];  // E1 (Ex, Ez), first row, by front parity
  T h2[NW + 2][2][64];     // H2 (Hx, Hz), last row
  Src<T> src[2][kMarchSources];
  // per plane of the chunk, byte c: the watched boxes holding the plane
  // that record component c
  unsigned long long pmask[kPairMaxPlanes];
  int wb[kPairWatch][4];  // j0, j1, k0, k1 of the watched boxes
  int wn;
};

// Value of lane + 1 / lane - 1 (DPP wavefront shifts; the end lanes get
// their own value).
template <int CTRL, typename T>
__device__ __forceinline__ T lane_shift(T v) {
  if constexpr (sizeof(T) == 4) {
    const int x = __builtin_bit_cast(int, v);
    return __builtin_bit_cast(T, __builtin_amdgcn_update_dpp(x, x, CTRL, 0xF, 0xF, false));
  } else {
    typedef int i2 __attribute__((ext_vector_type(2)));
    const i2 x = __builtin_bit_cast(i2, v);
    i2 y;
    y.x = __builtin_amdgcn_update_dpp(x.x, x.x, CTRL, 0xF, 0xF, false);
    y.y = __builtin_amdgcn_update_dpp(x.y, x.y, CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(T, y);
  }
}
template <typename T>
__device__ __forceinline__ T lane_next(T v) {
  return lane_shift<0x130>(v);  // wave_shl:1
}
template <typename T>
__device__ __forceinline__ T lane_prev(T v) {
  return lane_shift<0x138>(v);  // wave_shr:1
}

template <typename T, int NW, int RW, bool LOSSY, bool INT, bool X, bool Y, bool Z>
struct Pair {
  static constexpr unsigned E = sizeof(T);
  static constexpr int TJ = NW * RW - 4;  // core rows

  Geo g;
  T dtm_dl, dtm, inv_dl;
  Rsrc r0, r1, rc, rq0, rq1, rs;  // old slab, new slab, coefficients, psi old / new, shadow
  const T *__restrict__ prof;
  int i0, i1;  // core planes
  int lane, w;
  int kc;               // clamped column of the lane
  bool inz;             // column in a z-stripe
  int jr[RW];           // clamped row of each slot
  unsigned o4[RW];      // byte offset of (jr, kc) inside a plane
  unsigned qy4[RW];     // byte offset inside a y-stripe plane (kDropOffset outside)
  unsigned qz4[RW];     // byte offset inside the z-stripe part of a plane (kDropOffset outside)
  bool iny[RW];
  bool own[RW];         // a core cell inside the grid
  unsigned wm[RW];      // watched boxes holding the cell (core cells only), one bit each
  unsigned wmw;         // their union over the wavefront
  Prof1<T> pyh[RW], pye[RW], pzh, pze;
  // carries, per row slot
  T e0[RW][3];    // old E of the front plane
  T h1m[RW][3];   // H1 of the plane before the front
  T e1m[RW][3];   // E1 one plane back
  T e1mm[RW][3];  // E1 two planes back
  T h2m[RW][3];   // H2 two planes back
  T h2y[RW], h2z[RW];       // Hy, Hz of H2 three planes back
  T cm[RW][3], cmm[RW][3];  // ce one and two planes back
  T am[RW][3], amm[RW][3];  // ca one and two planes back (lossy grids)
  T ph1m[RW][6];             // psi_H after step 1, one plane back
  T pe1m[RW][6], pe1mm[RW][6];  // psi_E after step 1, one and two planes back

  __device__ __forceinline__ unsigned vsb() const { return (unsigned)g.vs * E; }
  __device__ __forceinline__ int clampx(int i) const { return min(max(i, 0), g.nx - 1); }
  __device__ __forceinline__ unsigned plane_b(int i) const {
    return (unsigned)clampx(i) * (unsigned)(g.ny * g.nz) * E;
  }
  __device__ __forceinline__ unsigned psib(int s, unsigned elems) const {
    return (psi_slot(g, s) + elems) * E;
  }
  __device__ __forceinline__ PsiPlane psi_plane(int i) const {
    const int n = g.npml, ic = clampx(i);
    PsiPlane pp;
    pp.inx = X && n > 0 && (ic < n || ic >= g.nx - n);
    pp.xs = pp.inx ? (unsigned)stripe_local(ic, n, g.nx) * (unsigned)(g.ny * g.nz) : 0u;
    pp.ys = (unsigned)ic * 2u * (unsigned)n * (unsigned)g.nz;
    pp.zs = (unsigned)ic * (unsigned)g.ny * 2u * (unsigned)n;
    pp.dead = false;
    return pp;
  }
  __device__ __forceinline__ StripeIdx stripes(int q, bool inx) const {
    StripeIdx st;
    st.inx = inx;
    st.iny = iny[q];
    st.inz = inz;
    st.qx = st.qy = st.qz = 0;
    return st;
  }

  // plane p: old H(p), ce(p), ca(p), psi(p) and old E(p+1).  Sets past the
  // last plane the chunk needs (dead) are dropped: their loads return 0.
  __device__ __forceinline__ void load(PairLoad<T> (&s)[RW], int p) const {
    const bool dead = p > i1;
    const unsigned vs = vsb();
    const unsigned pl = dead ? kDropOffset : plane_b(p), pl1 = dead ? kDropOffset : plane_b(p + 1);
    const PsiPlane pp = psi_plane(p);
#pragma unroll
    for (int q = 0; q < RW; ++q) {
      s[q].ex = bld<T, false>(r0, o4[q], pl1);
      s[q].ey = bld<T, false>(r0, o4[q], vs + pl1);
      s[q].ez = bld<T, false>(r0, o4[q], 2 * vs + pl1);
      s[q].hx = bld<T, false>(r0, o4[q], 3 * vs + pl);
      s[q].hy = bld<T, false>(r0, o4[q], 4 * vs + pl);
      s[q].hz = bld<T, false>(r0, o4[q], 5 * vs + pl);
      s[q].cx = bld<T, false>(rc, o4[q], pl);
      s[q].cy = bld<T, false>(rc, o4[q], vs + pl);
      s[q].cz = bld<T, false>(rc, o4[q], 2 * vs + pl);
      if constexpr (LOSSY) {
        s[q].ax = bld<T, false>(rc, o4[q], 3 * vs + pl);
        s[q].ay = bld<T, false>(rc, o4[q], 4 * vs + pl);
        s[q].az = bld<T, false>(rc, o4[q], 5 * vs + pl);
      }
      if constexpr (X) {
        const unsigned vx = pp.inx && !dead ? o4[q] : kDropOffset;
        s[q].ph[0] = bld<T, false>(rq0, vx, psib(0, pp.xs));
        s[q].ph[1] = bld<T, false>(rq0, vx, psib(1, pp.xs));
        s[q].pe[0] = bld<T, false>(rq0, vx, psib(6, pp.xs));
        s[q].pe[1] = bld<T, false>(rq0, vx, psib(7, pp.xs));
      }
      if constexpr (Y) {
        const unsigned vy = dead ? kDropOffset : qy4[q];
        s[q].ph[2] = bld<T, false>(rq0, vy, psib(2, pp.ys));
        s[q].ph[3] = bld<T, false>(rq0, vy, psib(3, pp.ys));
        s[q].pe[2] = bld<T, false>(rq0, vy, psib(8, pp.ys));
        s[q].pe[3] = bld<T, false>(rq0, vy, psib(9, pp.ys));
      }
      if constexpr (Z) {
        const unsigned vz = dead ? kDropOffset : qz4[q];
        s[q].ph[4] = bld<T, false>(rq0, vz, psib(4, pp.zs));
        s[q].ph[5] = bld<T, false>(rq0, vz, psib(5, pp.zs));
        s[q].pe[4] = bld<T, false>(rq0, vz, psib(10, pp.zs));
        s[q].pe[5] = bld<T, false>(rq0, vz, psib(11, pp.zs));
      }
    }
  }

  // psi store of a core cell to slot s at byte offset a (slot base excluded);
  // other lanes get an offset past the psi slab's end and are dropped.
  __device__ __forceinline__ void psi_st(T v, int s, bool ok, unsigned a) const {
    bst<T, false>(rq1, v, ok ? a : kDropOffset, psi_slot(g, s) * E);
  }

  // Sources of list l on the cell (plane i), added to E before its update
  // in list order, as the single-step sweeps add them.
  __device__ __forceinline__ void add_sources(const Src<T> *ls, int n, int ilo, int ihi, int i,
                                              int q, T cx, T cy, T cz, T &ex, T &ey,
                                              T &ez) const {
    for (int t = 0; t < (i >= ilo && i < ihi ? n : 0); ++t) {
      const auto *vp = lds_src_ptr(ls, t);
      if (i < vp->ci0 || i >= vp->ci1 || jr[q] < vp->cj0 || jr[q] >= vp->cj1) continue;
      const Src<T> sq = lds_src(ls, t);
      if (kc >= sq.ck0 && kc < sq.ck1) {
        const int c = sq.comp;
        const bool c0 = c == 0, c1 = c == 1;
        const T ce = pick3(c0, c1, cx, cy, cz);
        const T v = src_term<T>(sq, i, jr[q], kc, ce, sq.amp);
        ex = c0 ? ex + v : ex;
        ey = c1 ? ey + v : ey;
        ez = !c0 && !c1 ? ez + v : ez;
      }
    }
  }


  // Front plane p: loads of plane p+1, H1(p) and E2(p-2), barrier, E1(p)
  // and H2(p-1), barrier.  cur holds the loads of plane p, nxt receives
  // those of plane p+1.  A sub-update runs only on the rows and fronts
  // whose results some stored cell depends on (wave-uniform tests): H1 on
  // rows 0 .. R-2 and fronts up to i1, E1 on rows 1 .. R-2 and fronts
  // i0-1 .. i1, H2 on rows 1 .. R-3 and fronts i0 .. i1, E2 on rows
  // 2 .. R-3 and fronts i0+2 .. i1+1 (R rows, chunk [i0, i1)).
  __device__ __forceinline__ void step(PairLds<T, NW> &L, int par, int p, PairLoad<T> (&cur)[RW],
                                       PairLoad<T> (&nxt)[RW], int nsrc0, int nsrc1, int sa0,
                                       int sa1, int sb0, int sb1, bool wplane) {
    constexpr int R = NW * RW;
    __builtin_amdgcn_sched_barrier(0);
    load(nxt, p + 1);
    __builtin_amdgcn_sched_barrier(0);
    const unsigned vs = vsb();
    const int ia = clampx(p), ib = clampx(p - 1), ie = clampx(p - 2);
    const PsiPlane ppa = psi_plane(p), ppb = psi_plane(p - 1), ppe = psi_plane(p - 2);
    const bool fh1 = p <= i1, fe1 = p >= i0 - 1 && p <= i1, fh2 = p >= i0 && p <= i1;
    const bool fe2 = p >= i0 + 2 && p <= i1 + 1;

    // ── H1(p) ──
    T h1[RW][3], ph1[RW][6];
#pragma unroll
    for (int q = 0; q < RW; ++q) {
      const int r = RW * w + q;
      h1[q][0] = cur[q].hx;
      h1[q][1] = cur[q].hy;
      h1[q][2] = cur[q].hz;
#pragma unroll
      for (int s = 0; s < 6; ++s) ph1[q][s] = (s < 2 ? X : s < 4 ? Y : Z) ? cur[q].ph[s] : T(0);
      T exj, ezj;
      if (q + 1 < RW) {
        exj = e0[q + 1][0];
        ezj = e0[q + 1][2];
      } else {
        exj = L.e0[w + 2][0][lane];
        ezj = L.e0[w + 2][1][lane];
      }
      const T exk = lane_next(e0[q][0]), eyk = lane_next(e0[q][1]);
      if (fh1 && r <= R - 2) {
        if constexpr (INT) {
          h_interior(e0[q][0], e0[q][1], e0[q][2], exj, ezj, exk, eyk, cur[q].ey, cur[q].ez,
                     h1[q][0], h1[q][1], h1[q][2], dtm_dl);
        } else {
          const Prof1<T> pxh = prof1(g, prof, 1, 0, ia);
          h_math(g, ia, jr[q], kc, stripes(q, ppa.inx), pxh, pyh[q], pzh, e0[q][0], e0[q][1],
                 e0[q][2], exj, ezj, exk, eyk, cur[q].ey, cur[q].ez, h1[q][0], h1[q][1],
                 h1[q][2], ph1[q], dtm_dl, dtm);
        }
      }
    }
    // ── E2(p-2) ──
    {
      const bool live = p - 2 >= i0 && p - 2 < i1;
      const unsigned pl = plane_b(p - 2);
#pragma unroll
      for (int q = 0; q < RW; ++q) {
        const int r = RW * w + q;
        T hx_jm, hz_jm;
        if (q > 0) {
          hx_jm = h2m[q - 1][0];
          hz_jm = h2m[q - 1][2];
        } else {
          hx_jm = L.h2[w][0][lane];
          hz_jm = L.h2[w][1][lane];
        }
        const T hx_km = lane_prev(h2m[q][0]), hy_km = lane_prev(h2m[q][1]);
        if (fe2 && r >= 2 && r <= R - 3) {
          T ex = e1mm[q][0], ey = e1mm[q][1], ez = e1mm[q][2];
          add_sources(L.src[1], nsrc1, sb0, sb1, p - 2, q, cmm[q][0], cmm[q][1], cmm[q][2], ex,
                      ey, ez);
          const T ax = LOSSY ? amm[q][0] : T(1), ay = LOSSY ? amm[q][1] : T(1);
          const T az = LOSSY ? amm[q][2] : T(1);
          T ps[6];
#pragma unroll
          for (int s = 0; s < 6; ++s) ps[s] = pe1mm[q][s];
          if constexpr (INT) {
            e_interior(h2m[q][0], h2m[q][1], h2m[q][2], hx_jm, hz_jm, hx_km, hy_km, h2y[q],
                       h2z[q], ex, ey, ez, ax, ay, az, cmm[q][0], cmm[q][1], cmm[q][2], inv_dl);
          } else {
            const Prof1<T> pxe = prof1(g, prof, 0, 0, ie);
            e_math(g, ie, jr[q], kc, stripes(q, ppe.inx), pxe, pye[q], pze, h2m[q][0],
                   h2m[q][1], h2m[q][2], hx_jm, hz_jm, hx_km, hy_km, h2y[q], h2z[q], ex, ey, ez,
                   ax, ay, az, cmm[q][0], cmm[q][1], cmm[q][2], ps, inv_dl);
          }
          const bool ok = own[q] && live;
          const unsigned v = ok ? o4[q] + pl : kDropOffset;
          bst<T, false>(r1, ex, v, 0u);
          bst<T, false>(r1, ey, v, vs);
          bst<T, false>(r1, ez, v, 2 * vs);
          if constexpr (X) {
            psi_st(ps[0], 6, ok && ppe.inx, o4[q] + ppe.xs * E);
            psi_st(ps[1], 7, ok && ppe.inx, o4[q] + ppe.xs * E);
          }
          if constexpr (Y) {
            psi_st(ps[2], 8, ok && iny[q], qy4[q] + ppe.ys * E);
            psi_st(ps[3], 9, ok && iny[q], qy4[q] + ppe.ys * E);
          }
          if constexpr (Z) {
            psi_st(ps[4], 10, ok && inz, qz4[q] + ppe.zs * E);
            psi_st(ps[5], 11, ok && inz, qz4[q] + ppe.zs * E);
          }
        }
      }
    }
    L.h1[w + 1][0][lane] = h1[RW - 1][0];
    L.h1[w + 1][1][lane] = h1[RW - 1][2];
    lds_barrier();

    // ── E1(p) ──
    T e1[RW][3], pe1[RW][6];
#pragma unroll
    for (int q = 0; q < RW; ++q) {
      const int r = RW * w + q;
      e1[q][0] = e0[q][0];
      e1[q][1] = e0[q][1];
      e1[q][2] = e0[q][2];
#pragma unroll
      for (int s = 0; s < 6; ++s) pe1[q][s] = (s < 2 ? X : s < 4 ? Y : Z) ? cur[q].pe[s] : T(0);
      T hx_jm, hz_jm;
      if (q > 0) {
        hx_jm = h1[q - 1][0];
        hz_jm = h1[q - 1][2];
      } else {
        hx_jm = L.h1[w][0][lane];
        hz_jm = L.h1[w][1][lane];
      }
      const T hx_km = lane_prev(h1[q][0]), hy_km = lane_prev(h1[q][1]);
      if (fe1 && r >= 1 && r <= R - 2) {
        add_sources(L.src[0], nsrc0, sa0, sa1, p, q, cur[q].cx, cur[q].cy, cur[q].cz, e1[q][0],
                    e1[q][1], e1[q][2]);
        const T ax = LOSSY ? cur[q].ax : T(1), ay = LOSSY ? cur[q].ay : T(1);
        const T az = LOSSY ? cur[q].az : T(1);
        if constexpr (INT) {
          e_interior(h1[q][0], h1[q][1], h1[q][2], hx_jm, hz_jm, hx_km, hy_km, h1m[q][1],
                     h1m[q][2], e1[q][0], e1[q][1], e1[q][2], ax, ay, az, cur[q].cx, cur[q].cy,
                     cur[q].cz, inv_dl);
        } else {
          const Prof1<T> pxe = prof1(g, prof, 0, 0, ia);
          e_math(g, ia, jr[q], kc, stripes(q, ppa.inx), pxe, pye[q], pze, h1[q][0], h1[q][1],
                 h1[q][2], hx_jm, hz_jm, hx_km, hy_km, h1m[q][1], h1m[q][2], e1[q][0], e1[q][1],
                 e1[q][2], ax, ay, az, cur[q].cx, cur[q].cy, cur[q].cz, pe1[q], inv_dl);
        }
      }
    }
    // the state between the two steps, for the watched cells and
    // components (plane, wave and component tests uniform; rarely taken)
    const unsigned long long pm = wplane ? L.pmask[p - i0] : 0ull;
    if (wplane && pm) {
      const unsigned pl = plane_b(p);
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        const unsigned pc = __builtin_amdgcn_readfirstlane((unsigned)(pm >> (8 * c)) & 0xFFu);
        if (pc & wmw) {
#pragma unroll
          for (int q = 0; q < RW; ++q)
            if (wm[q] & pc) bst<T, false>(rs, c < 3 ? e1[q][c] : h1[q][c - 3], o4[q] + pl, c * vs);
        }
      }
    }
    // ── H2(p-1) ──
    T h2[RW][3];
    {
      const bool live = p - 1 >= i0 && p - 1 < i1;
      const unsigned pl = plane_b(p - 1);
#pragma unroll
      for (int q = 0; q < RW; ++q) {
        const int r = RW * w + q;
        h2[q][0] = h1m[q][0];
        h2[q][1] = h1m[q][1];
        h2[q][2] = h1m[q][2];
        T exj, ezj;
        if (q + 1 < RW) {
          exj = e1m[q + 1][0];
          ezj = e1m[q + 1][2];
        } else {
          exj = L.e1[par ^ 1][w + 2][0][lane];
          ezj = L.e1[par ^ 1][w + 2][1][lane];
        }
        const T exk = lane_next(e1m[q][0]), eyk = lane_next(e1m[q][1]);
        if (fh2 && r >= 1 && r <= R - 3) {
          T ps[6];
#pragma unroll
          for (int s = 0; s < 6; ++s) ps[s] = ph1m[q][s];
          if constexpr (INT) {
            h_interior(e1m[q][0], e1m[q][1], e1m[q][2], exj, ezj, exk, eyk, e1[q][1], e1[q][2],
                       h2[q][0], h2[q][1], h2[q][2], dtm_dl);
          } else {
            const Prof1<T> pxh = prof1(g, prof, 1, 0, ib);
            h_math(g, ib, jr[q], kc, stripes(q, ppb.inx), pxh, pyh[q], pzh, e1m[q][0], e1m[q][1],
                   e1m[q][2], exj, ezj, exk, eyk, e1[q][1], e1[q][2], h2[q][0], h2[q][1],
                   h2[q][2], ps, dtm_dl, dtm);
          }
          const bool ok = own[q] && live;
          const unsigned v = ok ? o4[q] + pl : kDropOffset;
          bst<T, false>(r1, h2[q][0], v, 3 * vs);
          bst<T, false>(r1, h2[q][1], v, 4 * vs);
          bst<T, false>(r1, h2[q][2], v, 5 * vs);
          if constexpr (X) {
            psi_st(ps[0], 0, ok && ppb.inx, o4[q] + ppb.xs * E);
            psi_st(ps[1], 1, ok && ppb.inx, o4[q] + ppb.xs * E);
          }
          if constexpr (Y) {
            psi_st(ps[2], 2, ok && iny[q], qy4[q] + ppb.ys * E);
            psi_st(ps[3], 3, ok && iny[q], qy4[q] + ppb.ys * E);
          }
          if constexpr (Z) {
            psi_st(ps[4], 4, ok && inz, qz4[q] + ppb.zs * E);
            psi_st(ps[5], 5, ok && inz, qz4[q] + ppb.zs * E);
          }
        }
      }
    }
    L.e1[par][w + 1][0][lane] = e1[0][0];
    L.e1[par][w + 1][1][lane] = e1[0][2];
    L.h2[w + 1][0][lane] = h2[RW - 1][0];
    L.h2[w + 1][1][lane] = h2[RW - 1][2];
    L.e0[w + 1][0][lane] = cur[0].ex;
    L.e0[w + 1][1][lane] = cur[0].ez;
#pragma unroll
    for (int q = 0; q < RW; ++q) {
      h2y[q] = h2m[q][1];
      h2z[q] = h2m[q][2];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        h2m[q][c] = h2[q][c];
        e1mm[q][c] = e1m[q][c];
        e1m[q][c] = e1[q][c];
        h1m[q][c] = h1[q][c];
        cmm[q][c] = cm[q][c];
        if constexpr (LOSSY) amm[q][c] = am[q][c];
      }
#pragma unroll
      for (int s = 0; s < 6; ++s) {
        ph1m[q][s] = ph1[q][s];
        pe1mm[q][s] = pe1m[q][s];
        pe1m[q][s] = pe1[q][s];
      }
      cm[q][0] = cur[q].cx;
      cm[q][1] = cur[q].cy;
      cm[q][2] = cur[q].cz;
      if constexpr (LOSSY) {
        am[q][0] = cur[q].ax;
        am[q][1] = cur[q].ay;
        am[q][2] = cur[q].az;
      }
      e0[q][0] = cur[q].ex;
      e0[q][1] = cur[q].ey;
      e0[q][2] = cur[q].ez;
    }
    lds_barrier();
  }

  __device__ __forceinline__ void run(PairLds<T, NW> &L, int k0, int j0, int ci0, int ci1,
                                      const PairSources<T> &src, const PairWatch &watch) {
    i0 = ci0;
    i1 = ci1;
    lane = threadIdx.x & 63;
    w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int n = g.npml;
    const int kk = k0 - 2 + lane;
    kc = min(max(kk, 0), g.nz - 1);
    inz = Z && n > 0 && (kc < n || kc >= g.nz - n);
    const bool kown = kk >= k0 && kk < k0 + kPairCols && kk < g.nz;
#pragma unroll
    for (int q = 0; q < RW; ++q) {
      const int jj = j0 - 2 + RW * w + q;
      jr[q] = min(max(jj, 0), g.ny - 1);
      o4[q] = (unsigned)(jr[q] * g.nz + kc) * E;
      own[q] = kown && jj >= j0 && jj < j0 + TJ && jj < g.ny;
      iny[q] = Y && n > 0 && (jr[q] < n || jr[q] >= g.ny - n);
      qy4[q] = iny[q] ? (unsigned)(stripe_local(jr[q], n, g.ny) * g.nz + kc) * E : kDropOffset;
      qz4[q] = inz ? (unsigned)(jr[q] * 2 * n + stripe_local(kc, n, g.nz)) * E : kDropOffset;
      pyh[q] = prof1(g, prof, 1, 1, jr[q]);
      pye[q] = prof1(g, prof, 0, 1, jr[q]);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        h1m[q][c] = e1m[q][c] = e1mm[q][c] = h2m[q][c] = T(0);
        cm[q][c] = cmm[q][c] = am[q][c] = amm[q][c] = T(0);
      }
#pragma unroll
      for (int s = 0; s < 6; ++s) ph1m[q][s] = pe1m[q][s] = pe1mm[q][s] = T(0);
      h2y[q] = h2z[q] = T(0);
    }
    pzh = prof1(g, prof, 1, 2, kc);
    pze = prof1(g, prof, 0, 2, kc);
    const unsigned vs = vsb(), pl = plane_b(i0 - 2);
#pragma unroll
    for (int q = 0; q < RW; ++q) {
      e0[q][0] = bld<T, false>(r0, o4[q], pl);
      e0[q][1] = bld<T, false>(r0, o4[q], vs + pl);
      e0[q][2] = bld<T, false>(r0, o4[q], 2 * vs + pl);
    }
    PairLoad<T> A[RW], B[RW];
    load(A, i0 - 2);
    if (threadIdx.x == 0) {  // constant indices: no indexed copy of the arguments
#pragma unroll
      for (int t = 0; t < kMarchSources; ++t) {
        L.src[0][t] = src.s[0][t];
        L.src[1][t] = src.s[1][t];
      }
#pragma unroll
      for (int b = 0; b < kPairWatch; ++b)
#pragma unroll
        for (int c = 0; c < 4; ++c) L.wb[b][c] = watch.b[b][2 + c];
      L.wn = watch.n;
    }
    // per plane of the chunk and component, the watched boxes holding it
    for (int t = threadIdx.x; t < i1 - i0; t += NW * 64) {
      unsigned long long m = 0;
#pragma unroll
      for (int b = 0; b < kPairWatch; ++b)
        if (b < watch.n && i0 + t >= watch.b[b][0] && i0 + t < watch.b[b][1])
#pragma unroll
          for (int c = 0; c < 6; ++c)
            if ((watch.c[b] >> c) & 1) m |= 1ull << (8 * c + b);
      L.pmask[t] = m;
    }
    L.e0[w + 1][0][lane] = e0[0][0];
    L.e0[w + 1][1][lane] = e0[0][2];
    lds_barrier();
    const int ns0 = src.n[0], ns1 = src.n[1];
    // plane windows of the sources of each list that meet this tile's rows
    // j0-2 ... j0+TJ+1 and columns k0-2 ... k0+61 (every slot's clamped cell
    // lies there); the per-cell tests still decide, so a tile no source
    // touches skips the source loops
    int sa0 = g.nx, sa1 = 0, sb0 = g.nx, sb1 = 0;
#pragma unroll
    for (int t = 0; t < kMarchSources; ++t) {
      const Src<T> &a = src.s[0][t], &b = src.s[1][t];
      if (t < ns0 && a.cj0 <= j0 + TJ + 1 && a.cj1 > j0 - 2 && a.ck0 <= k0 + 61 && a.ck1 > k0 - 2) {
        sa0 = min(sa0, a.ci0);
        sa1 = max(sa1, a.ci1);
      }
      if (t < ns1 && b.cj0 <= j0 + TJ + 1 && b.cj1 > j0 - 2 && b.ck0 <= k0 + 61 && b.ck1 > k0 - 2) {
        sb0 = min(sb0, b.ci0);
        sb1 = max(sb1, b.ci1);
      }
    }
    // watched planes: those of the chunk only (a halo plane's step-1 values
    // are not all exact, and another block owns the plane)
    unsigned any = 0;
#pragma unroll
    for (int q = 0; q < RW; ++q) {
      wm[q] = 0;
      for (int b = 0; b < L.wn; ++b) {
        const int *x = L.wb[b];
        if (jr[q] >= x[0] && jr[q] < x[1] && kc >= x[2] && kc < x[3]) wm[q] |= 1u << b;
      }
      if (!own[q]) wm[q] = 0;
      any |= wm[q];
    }

    wmw = 0;
    for (int b = 0; b < kPairWatch; ++b)
      if (__ballot((any >> b) & 1u)) wmw |= 1u << b;
    wmw = __builtin_amdgcn_readfirstlane(wmw);
    // fronts i0-2 ... i1+1, an even count (the extra front is discarded)
    const int nf = (i1 - i0 + 5) / 2 * 2;
    int t = 0;
#pragma unroll 1
    do {
      const int p = i0 - 2 + t;
      step(L, 0, p, A, B, ns0, ns1, sa0, sa1, sb0, sb1, wmw && p >= i0 && p < i1);
      step(L, 1, p + 1, B, A, ns0, ns1, sa0, sa1, sb0, sb1, wmw && p + 1 >= i0 && p + 1 < i1);
      t += 2;
    } while (t < nf);
  }
};

// One block per (k tile, j tile, i chunk) with the block order of the march
// sweep (XCD runs, j tiles rotated from chunk to chunk).  A block takes the
// variant with the psi its tile meets, halo included.
template <typename T, int NW, int RW, bool LOSSY>
__global__ __launch_bounds__(NW * 64) void k_pair(Geo g, int planes, T dtm_dl, T dtm, T inv_dl,
                                                  const T *__restrict__ s0, T *__restrict__ s1,
                                                  const T *__restrict__ coef,
                                                  const T *__restrict__ prof,
                                                  const T *__restrict__ q0, T *__restrict__ q1,
                                                  T *__restrict__ shadow, int tiles_k,
                                                  int tiles_j, int xcd, int run,
                                                  const PairSources<T> src, const PairWatch watch) {
  __shared__ PairLds<T, NW> L;
  constexpr int TJ = NW * RW - 4;
  const int B = xcd_runs(blockIdx.x, gridDim.x, xcd, run);
  const int bk = B % tiles_k, bc = B / (tiles_k * tiles_j);
  const int bj = ((B / tiles_k) % tiles_j + bc * max(1, run / tiles_k)) % tiles_j;
  const int k0 = bk * kPairCols, j0 = bj * TJ;
  const int i0 = bc * planes, i1 = min(i0 + planes, g.nx);
  const int n = g.npml;
  const bool x = n > 0 && (i0 - 2 < n || i1 + 2 > g.nx - n);
  const bool y = n > 0 && (j0 - 2 < n || min(j0 + TJ + 2, g.ny) > g.ny - n);
  const bool z = n > 0 && (k0 - 2 < n || min(k0 + kPairCols + 2, g.nz) > g.nz - n);
  const int sbr = (int)slab_bytes<T>(g);  // below 2 GB (checked on the host)
  const Rsrc r0 = __builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(s0), (short)0, sbr, 0x00020000);
  const Rsrc rc = __builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(coef), (short)0, sbr,
                                                    0x00020000);
  const Rsrc r1 = __builtin_amdgcn_make_buffer_rsrc(s1, (short)0, sbr, 0x00020000);
  const Rsrc rs = __builtin_amdgcn_make_buffer_rsrc(shadow, (short)0, shadow ? sbr : 0, 0x00020000);
  const int psi_bytes = (int)((psi_slot(g, PSI_COUNT) + 64u) * (unsigned)sizeof(T));
  const Rsrc rq0 = __builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(q0), (short)0, psi_bytes,
                                                     0x00020000);
  const Rsrc rq1 = __builtin_amdgcn_make_buffer_rsrc(q1, (short)0, psi_bytes, 0x00020000);
#define HF_PAIR_RUN(I, X, Y, Z)                                                                \
  {                                                                                            \
    Pair<T, NW, RW, LOSSY, I, X, Y, Z> m{g,   dtm_dl, dtm, inv_dl, r0,  r1,                      \
                                         rc,  rq0,    rq1, rs,     prof};                        \
    m.run(L, k0, j0, i0, i1, src, watch);                                                      \
  }
  if (n > 0 && !x && !y && !z) {
    HF_PAIR_RUN(true, false, false, false)  // every cell of the tile interior
  } else if (x) {
    if (y) {
      if (z) HF_PAIR_RUN(false, true, true, true) else HF_PAIR_RUN(false, true, true, false)
    } else {
      if (z) HF_PAIR_RUN(false, true, false, true) else HF_PAIR_RUN(false, true, false, false)
    }
  } else {
    if (y) {
      if (z) HF_PAIR_RUN(false, false, true, true) else HF_PAIR_RUN(false, false, true, false)
    } else {
      if (z) HF_PAIR_RUN(false, false, false, true) else HF_PAIR_RUN(false, false, false, false)
    }
  }
#undef HF_PAIR_RUN
}

}  // namespace hfdtd
