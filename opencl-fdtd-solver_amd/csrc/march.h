// Register-pipelined fused sweep ("march"): H and E of one Yee step in one
// launch, every value of the grid read and written once.
//
// A block of four wavefronts owns a 64(k) x 4*RW(j) column tile and marches
// along i over `planes` planes, H then E on each.  Wave w owns rows
// j0 + RW*w ... j0 + RW*w + RW - 1 (lane = k - k0), one "slot" per row.  A
// further halo slot per wave recomputes the H values the tile's E needs from
// outside the tile: lanes 0-15 of wave w take row j0 - 1 at columns
// k0 + 16w ... k0 + 16w + 15, lanes 16 ... 16 + RW - 1 take column k0 - 1 of
// the wave's own rows.  New H goes to LDS (double-buffered by plane parity,
// one barrier per plane); E then takes its j-1 and k-1 H neighbours from
// LDS and its i-1 neighbour from registers (H of the previous plane).
//
// Loads are software-pipelined through registers: while plane i is
// computed, the loads of plane i+1 are in flight (the own cell's old H, ce,
// ca and psi_z, E at i+2, and the E neighbours at j+1 and k+1).  E of a
// plane is loaded once, as the i+1 neighbour of the plane before, and kept.
// Every field load and store of the loop is unconditional and goes through
// a buffer resource: the per-lane byte offset of a slot is fixed for the
// whole block, the plane and volume parts are wave-uniform, so issuing one
// costs no vector ALU work and the compiler's counted waits stay exact from
// one plane to the next.  Lanes and planes outside the grid load clamped
// addresses and store into the volume padding (psi into the stripe slab's
// padding).  psi of the x and y stripes, uniform per plane and per row, is
// loaded at the top of the plane that uses it, before the prefetch, so
// waiting for it never waits for the prefetch.
//
// The per-cell arithmetic is h_math / e_math, shared with every other
// sweep.  State is ping-ponged like the LDS-DMA fused sweep (reads slab 0,
// writes slab 1); up to four sources of the step are added to E on their
// cells right before its update, in the order the fix-up kernel of the
// LDS-DMA sweep adds them (steps with more sources use that kernel).  Offsets are 32-bit: slabs stay below 4 GB
// (checked on the host).
#pragma once

#include <type_traits>

#include "fdtd_core.h"
#include "n2f_gather.h"

namespace hfdtd {

constexpr int kMarchThreads = 256;  // four wavefronts
// kMarchSources (common.h): sources applied inside the sweep (more: fix-up kernel)

// Sources of one step, passed to the sweep by value.  The block copies them
// to LDS once and reads them there with volatile loads, so they occupy no
// registers across the plane loop.
template <typename T>
struct MarchSources {
  int n;
  Src<T> s[kMarchSources];
};

// a if c0, else b if c1, else c, by masking the bit patterns.
template <typename T>
__device__ __forceinline__ T pick3(bool c0, bool c1, T a, T b, T c) {
  typedef typename std::conditional<sizeof(T) == 4, unsigned, unsigned long long>::type U;
  const U m0 = c0 ? ~U(0) : U(0), m1 = !c0 && c1 ? ~U(0) : U(0), m2 = ~(m0 | m1);
  return __builtin_bit_cast(T, (__builtin_bit_cast(U, a) & m0) | (__builtin_bit_cast(U, b) & m1) |
                                   (__builtin_bit_cast(U, c) & m2));
}

// The LDS copy of the sources seen through an LDS pointer: volatile reads
// through a generic pointer become flat loads, which count against the
// vector-memory counter, so each one waited for every field load in flight.
template <typename T>
__device__ __forceinline__ const volatile __attribute__((address_space(3))) Src<T> *
lds_src_ptr(const Src<T> *base, int t) {
  return (const volatile __attribute__((address_space(3))) Src<T> *)(base + t);
}

// Source t of the LDS copy, read field by field through volatile loads (so
// the compiler re-reads them where they are used instead of holding them).
template <typename T>
__device__ __forceinline__ Src<T> lds_src(const Src<T> *base, int t) {
  const volatile __attribute__((address_space(3))) Src<T> *p = lds_src_ptr(base, t);
  Src<T> s;
  s.comp = p->comp;
  s.mode = p->mode;
  s.taper = p->taper;
  s.i0 = p->i0;
  s.i1 = p->i1;
  s.j0 = p->j0;
  s.j1 = p->j1;
  s.k0 = p->k0;
  s.k1 = p->k1;
  s.ci0 = p->ci0;
  s.ci1 = p->ci1;
  s.cj0 = p->cj0;
  s.cj1 = p->cj1;
  s.ck0 = p->ck0;
  s.ck1 = p->ck1;
  s.amp = p->amp;
  s.re = p->re;
  return s;
}
// A buffer offset no slab reaches whose resource carries its size (psi
// slabs stay below 2 GB, checked on the host): stores there fall outside the
// resource's range and are dropped, loads return 0, neither touches memory.
constexpr unsigned kDropOffset = 0x80000000u;

// The engine's z-face record (ZRec in common.h) as a sweep argument; a
// null buf records nothing.  Blocks whose columns hold a face's k or k - 1
// (and whose rows and planes meet the record) run the recording variant:
// after its E update each owned cell there stores its new (Ex, Ey, Hx, Hy)
// with one 128-bit store.  Lanes and planes that record nothing get offsets
// past the record's resource, whose size is exact (below 1 GB): their
// stores are dropped, and no sum of the two drop offsets wraps back into
// range.
struct ZCap {
  float *buf;
  int kf0, kf1;  // the faces' k
  int i0, i1, j0, j1;
  unsigned bytes;
};
constexpr unsigned kCapDropLane = 0x80000000u;
constexpr unsigned kCapDropPlane = 0x40000000u;

__device__ __forceinline__ void bst_quad(Rsrc r, float a, float b, float c, float d, unsigned voff,
                                         unsigned soff) {
  typedef unsigned u4 __attribute__((ext_vector_type(4)));
  const u4 v = {__builtin_bit_cast(unsigned, a), __builtin_bit_cast(unsigned, b),
                __builtin_bit_cast(unsigned, c), __builtin_bit_cast(unsigned, d)};
  __builtin_amdgcn_raw_buffer_store_b128(v, r, voff, soff, 0);
}

// Tile shape: W columns of k (64 or 32) by 4 * RW * (64 / W) rows of j.
// A wavefront's slot holds 64 / W rows of W columns (lane = row * W + column
// within the slot), so narrow grids leave fewer lanes idle and the tile's
// recomputed halo shrinks as it gets squarer.
template <int RW, int W>
struct MarchShape {
  static_assert(W == 64 || W == 32, "march tiles are 64 or 32 columns wide");
  static constexpr int G = 64 / W;        // rows of one slot
  static constexpr int TJ = 4 * RW * G;   // rows of the tile
  static constexpr int C = W + 2;         // LDS columns: k0-1 ... k0+W
  static constexpr int R = TJ + 4;        // LDS rows: j0-1 ... j0+TJ, a trash and a spare row
};

// LDS of a block: new H and E of the planes, double-buffered by plane
// parity, on a grid of rows j0-1 ... j0+TJ (then a trash row and one spare
// row, so neighbour reads from the trash row stay inside the array) by
// columns k0-1 ... k0+W.
template <typename T, int RW, int W = 64>
struct MarchLds {
  static constexpr int R = MarchShape<RW, W>::R, C = MarchShape<RW, W>::C;
  T h[2][3][R][C];  // new H(i): [parity][Hx Hy Hz]
  T e[2][3][R][C];  // E(i) (old), the j+1 / k+1 neighbours of H
};

// Plane-i data of one slot, loaded one plane ahead in two parts: the H part
// at the top of plane i-1, the E part after its H half, so each is in
// flight for a full half-plane while the registers of the part it replaces
// are still being read.
template <typename T>
struct MarchH {
  T ex, ey, ez;  // E(i+1) at the cell (becomes E of the next plane)
  T hx, hy, hz;  // old H(i)
  T p[6];        // old psi of the H slots (PML tiles)
};

template <typename T>
struct MarchE {
  T cx, cy, cz;  // ce(i)
  // material (kCeIndex) or eps_r bits (kCeEps) of the cells at i+1, j+1, k+1
  unsigned mi1, mj1, mk1;
  T ax, ay, az;  // ca(i) (lossy grids)
  T p[6];        // old psi of the E slots (PML tiles)
};

// Fixed data of one slot.  The cell is clamped into the grid: values
// computed for cells outside it are never stored nor used.
struct MarchSlot {
  int j, k;
  unsigned o4;   // byte offset of (j, k) inside a plane
  unsigned qy4;  // byte offset inside a y-stripe plane (kDropOffset outside the stripe)
  unsigned qz4;  // byte offset inside the z-stripe rows of a plane (kDropOffset outside)
  unsigned o1, o1j, o1k;  // ce-source byte offsets of (j, k), (j+1, k), (j, k+1), clamped
  bool iny, inz;
  bool own;      // an owned cell inside the grid
  int lh, le;    // LDS element index (row * LDS columns + column) of its H and E
};

// Raw buffer loads and stores; NT marks them non-temporal (streamed once).
template <typename T, bool NT>
__device__ __forceinline__ T bld(Rsrc r, unsigned voff, unsigned soff) {
  constexpr int aux = NT ? 2 : 0;
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, aux));
  } else {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, aux));
  }
}

template <typename T, bool NT>
__device__ __forceinline__ void bst(Rsrc r, T v, unsigned voff, unsigned soff) {
  constexpr int aux = NT ? 2 : 0;
  if constexpr (sizeof(T) == 4) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, aux);
  } else {
    typedef unsigned u2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, voff, soff, aux);
  }
}

// Bytes of a six-volume field slab.
template <typename T>
__device__ __forceinline__ unsigned long long slab_bytes(const Geo &g) {
  return 6ull * (unsigned long long)g.vs * sizeof(T);
}

// psi addressing of one plane: whether it lies in an x-stripe, and the
// plane parts (in elements) of the x-, y- and z-stripe offsets.
struct PsiPlane {
  bool inx;
  unsigned xs, ys, zs;
  bool dead;  // a plane past the chunk: its loads are dropped
};

// A point probe of the batched sweep: one E component of one cell.  The
// block that owns the cell keeps the component's new value and, after its
// planes, adds value x phase to the probe's DFT (FP64), as the probe
// kernels of the split sweeps do with the stored value.
struct MarchProbe {
  int i, j, k, comp;
  int nf;             // frequencies
  const double *ph;   // nf x {re, im}: the phase factors of this step
  double *acc;        // nf x {re, im}
};

template <typename T, int RW, bool LOSSY, bool X, bool Y, bool Z, bool NT, int CEM = kCeVolumes,
          bool CAP = false, bool BIG = false, bool PRB = false, int W = 64>
struct March {
  static constexpr bool MAT = CEM != kCeVolumes;  // ce from a per-cell volume, not the ce volumes
  static constexpr unsigned CB = CEM == kCeEps ? 4u : 1u;  // bytes per cell of that volume
  static constexpr int S = RW + 1;  // RW owned rows + the halo slot
  static constexpr int TJ = MarchShape<RW, W>::TJ;
  static constexpr int G = MarchShape<RW, W>::G;
  static constexpr int LC = MarchShape<RW, W>::C;  // LDS row stride
  static constexpr unsigned E = sizeof(T);

  Geo g;
  T dtm_dl, dtm, inv_dl;
  Rsrc r0, r1, rc, rq0, rq1;  // slab 0, slab 1, coefficients, psi 0, psi 1
  const T *__restrict__ prof;
  const Src<T> *lsrc;  // sources of the step (LDS copy)
  int nsrc;
  int src_i0, src_i1;  // planes on which some source can reach this tile
  // indexed grids: one byte of material per cell, ce of an edge from the
  // table of its two cells' materials (LDS copy); eps_r grids: the eps_r
  // volume, ce = cscale / (eps_r averaged over the edge's two cells)
  Rsrc rmat;
  const float *mtab;
  int nmat;
  double cscale;    // dt / eps0 (kCeEps)
  unsigned mc[RW];  // material (or eps_r bits) of the owned cells on the current plane
  MarchSlot sl[S];
  Prof1<T> pyh[S], pye[S], pzh[S], pze[S];
  T hyc[RW], hzc[RW];  // new H(i-1) at the owned cells
  T oe[S][3];          // E(i) at every slot (copied from the loads of plane i-1)
  // the halo slot's per-lane offsets of its six field loads (E at i+1, old H
  // at i) and of its x-stripe psi loads: a lane loads only what its role
  // uses (see run), the others get kDropOffset and move no data
  unsigned hoff[6], hoffx;
  // z-face record (CAP): its resource and extents, and the per-lane byte
  // offset of an owned cell's quad in it
  Rsrc rcap;
  ZCap zc;
  unsigned zo[RW];
  // probe (PRB): the probe, whether owned slot q holds its cell, and the
  // component's new value there
  MarchProbe pr;
  bool phit[RW];
  T pv;
  unsigned padb;       // store offset of lanes and planes outside the grid
  int iend;            // first plane past this block's chunk
  bool sized;          // slab resources carry their sizes (slabs below 2 GB)

  // BIG: slabs of 4 GB and more, past 32-bit offsets.  Every access then
  // goes through a resource on its own volume (each below 2 GB), made from
  // the slab's base pointer where it is used.
  const T *s0p, *cp;
  T *s1p;

  __device__ __forceinline__ unsigned vsb() const { return (unsigned)g.vs * E; }
  __device__ __forceinline__ Rsrc vol_rsrc(const T *base, int c) const {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(base) + (size_t)c * g.vs, (short)0,
                                             (int)(g.vs * E), 0x00020000);
  }
  // volume c (Ex .. Hz) of the slab read, at byte offsets voff (per lane)
  // and pl (wave-uniform)
  __device__ __forceinline__ T ld0(int c, unsigned voff, unsigned pl) const {
    if constexpr (BIG) return bld<T, false>(vol_rsrc(s0p, c), voff, pl);
    else return bld<T, false>(r0, voff, (unsigned)c * vsb() + pl);
  }
  // coefficient volume c (ce_x .. ca_z), streamed
  __device__ __forceinline__ T ldc(int c, unsigned voff, unsigned pl) const {
    if constexpr (BIG) return bld<T, NT>(vol_rsrc(cp, c), voff, pl);
    else return bld<T, NT>(rc, voff, (unsigned)c * vsb() + pl);
  }
  // volume c of the slab written
  __device__ __forceinline__ void st1(int c, T v, unsigned voff) const {
    if constexpr (BIG) bst<T, NT>(vol_rsrc(s1p, c), v, voff, 0u);
    else bst<T, NT>(r1, v, voff, (unsigned)c * vsb());
  }
  __device__ __forceinline__ unsigned plane_b(int i) const {
    return (unsigned)min(i, g.nx - 1) * (unsigned)(g.ny * g.sk) * E;
  }
  // byte offset of psi slot s plus a part given in elements
  __device__ __forceinline__ unsigned psib(int s, unsigned elems) const {
    return (psi_slot(g, s) + elems) * E;
  }
  __device__ __forceinline__ PsiPlane psi_plane(int i) const {
    const int n = g.npml, ic = min(i, g.nx - 1);
    PsiPlane pp;
    pp.inx = n > 0 && (ic < n || ic >= g.nx - n);
    pp.xs = pp.inx ? (unsigned)stripe_local(ic, n, g.nx) * (unsigned)(g.ny * g.sk) : 0u;
    pp.ys = (unsigned)ic * 2u * (unsigned)n * (unsigned)g.sk;
    pp.zs = (unsigned)ic * (unsigned)g.ny * 2u * (unsigned)n;
    pp.dead = sized && i >= iend;
    return pp;
  }
  // plane part of a field load: past the resource's end for a plane after
  // the chunk (the load is dropped and returns 0), else the clamped plane
  __device__ __forceinline__ unsigned load_plane(int i, bool dead) const {
    return dead ? kDropOffset : plane_b(i);
  }

  // H part of plane i.  psi of the x and y stripes is loaded on every plane
  // of an X block, psi_y on every plane of a Y block, psi_z on every plane
  // of a Z block; lanes outside a
  // stripe load from past the end of the psi slab's resource, which returns
  // 0 without touching memory (h_math ignores the value).  Pointing them all
  // at one valid element instead would send every wave of the grid to the
  // same cache line.
  __device__ __forceinline__ void load_h(MarchH<T> &s, const MarchSlot &m, unsigned pl,
                                         unsigned pl1, const PsiPlane &pp, bool halo) const {
    const unsigned vs = vsb();
    s.ex = ld0(0, halo ? hoff[0] : m.o4, pl1);
    s.ey = ld0(1, halo ? hoff[1] : m.o4, pl1);
    s.ez = ld0(2, halo ? hoff[2] : m.o4, pl1);
    s.hx = ld0(3, halo ? hoff[3] : m.o4, pl);
    s.hy = ld0(4, halo ? hoff[4] : m.o4, pl);
    s.hz = ld0(5, halo ? hoff[5] : m.o4, pl);
    if constexpr (X) {
      const unsigned vx = pp.inx && !pp.dead ? (halo ? hoffx : m.o4) : kDropOffset;
      s.p[0] = bld<T, false>(rq0, vx, psib(0, pp.xs));
      s.p[1] = bld<T, false>(rq0, vx, psib(1, pp.xs));
    }
    if constexpr (Y) {
      const unsigned vy = pp.dead ? kDropOffset : m.qy4;
      s.p[2] = bld<T, false>(rq0, vy, psib(2, pp.ys));
      s.p[3] = bld<T, false>(rq0, vy, psib(3, pp.ys));
    }
    if constexpr (Z) {
      const unsigned vz = pp.dead ? kDropOffset : m.qz4;
      s.p[4] = bld<T, false>(rq0, vz, psib(4, pp.zs));
      s.p[5] = bld<T, false>(rq0, vz, psib(5, pp.zs));
    }
  }

  // plane part of a record store: dropped outside [i0, i1] and past the chunk
  __device__ __forceinline__ unsigned cap_plane(int i, bool live) const {
    return live && i >= zc.i0 && i <= zc.i1
               ? (unsigned)(i - zc.i0) * (unsigned)(zc.j1 - zc.j0 + 1) * 16u
               : kCapDropPlane;
  }

  __device__ __forceinline__ unsigned mat_plane(int i, bool dead) const {
    return dead ? kDropOffset : (unsigned)min(i, g.nx - 1) * (unsigned)(g.ny * g.sk) * CB;
  }
  // one value of the material index or of the eps_r volume
  __device__ __forceinline__ unsigned ld_mat(unsigned voff, unsigned soff) const {
    if constexpr (CEM == kCeEps) return __builtin_amdgcn_raw_buffer_load_b32(rmat, voff, soff, 0);
    else return __builtin_amdgcn_raw_buffer_load_b8(rmat, voff, soff, 0);
  }
  // ce of an edge between cells of eps_r a and b (bit patterns), as the
  // coefficient kernel computes it
  __device__ __forceinline__ float ce_eps(unsigned a, unsigned b) const {
    const double ea = (double)__builtin_bit_cast(float, a), eb = (double)__builtin_bit_cast(float, b);
    return (float)(cscale / (0.5 * (ea + eb)));
  }

  __device__ __forceinline__ void load_e(MarchE<T> &s, const MarchSlot &m, unsigned pl,
                                         unsigned pm, unsigned pm1, const PsiPlane &pp) const {
    const unsigned vs = vsb();
    if constexpr (MAT) {
      s.mi1 = ld_mat(m.o1, pm1);
      s.mj1 = ld_mat(m.o1j, pm);
      s.mk1 = ld_mat(m.o1k, pm);
    } else {
      s.cx = ldc(0, m.o4, pl);
      s.cy = ldc(1, m.o4, pl);
      s.cz = ldc(2, m.o4, pl);
    }
    if constexpr (LOSSY) {
      s.ax = ldc(3, m.o4, pl);
      s.ay = ldc(4, m.o4, pl);
      s.az = ldc(5, m.o4, pl);
    }
    if constexpr (X) {
      const unsigned vx = pp.inx && !pp.dead ? m.o4 : kDropOffset;
      s.p[0] = bld<T, false>(rq0, vx, psib(6, pp.xs));
      s.p[1] = bld<T, false>(rq0, vx, psib(7, pp.xs));
    }
    if constexpr (Y) {
      const unsigned vy = pp.dead ? kDropOffset : m.qy4;
      s.p[2] = bld<T, false>(rq0, vy, psib(8, pp.ys));
      s.p[3] = bld<T, false>(rq0, vy, psib(9, pp.ys));
    }
    if constexpr (Z) {
      const unsigned vz = pp.dead ? kDropOffset : m.qz4;
      s.p[4] = bld<T, false>(rq0, vz, psib(10, pp.zs));
      s.p[5] = bld<T, false>(rq0, vz, psib(11, pp.zs));
    }
  }

  __device__ __forceinline__ void load_h_all(MarchH<T> (&s)[S], int i) const {
    const PsiPlane pp = psi_plane(i);
    const unsigned pl = load_plane(i, pp.dead), pl1 = load_plane(i + 1, pp.dead);
#pragma unroll
    for (int q = 0; q < S; ++q) load_h(s[q], sl[q], pl, pl1, pp, q == RW);
  }

  __device__ __forceinline__ void load_e_all(MarchE<T> (&s)[RW], int i) const {
    const PsiPlane pp = psi_plane(i);
    const unsigned pl = load_plane(i, pp.dead);
    const unsigned pm = MAT ? mat_plane(i, pp.dead) : 0u, pm1 = MAT ? mat_plane(i + 1, pp.dead) : 0u;
#pragma unroll
    for (int q = 0; q < RW; ++q) load_e(s[q], sl[q], pl, pm, pm1, pp);
  }

  // psi store of an owned cell to stripe slot s at byte offset a (slot base
  // excluded, the wave-uniform part).  Lanes that must not store get an
  // offset past the end of the psi slab, whose resource carries its true
  // size, and the hardware drops their store; every store stays
  // unconditional, so the compiler's counted waits stay exact.
  __device__ __forceinline__ void psi_st(T v, int s, bool ok, unsigned a) const {
    bst<T, false>(rq1, v, ok ? a : kDropOffset, psi_slot(g, s) * E);
  }

  __device__ __forceinline__ StripeIdx stripes(const MarchSlot &m, bool inx) const {
    StripeIdx st;
    st.inx = X && inx;
    st.iny = m.iny;
    st.inz = m.inz;
    st.qx = st.qy = st.qz = 0;
    return st;
  }

  // Plane i: H on every slot, barrier, E on the owned slots.  hc / ec hold
  // the loads of plane i; hn / en, no longer read, receive those of plane
  // i+1.  E of the plane is carried in oe, so no value of a set stays live
  // while its next loads are in flight: the sets keep their registers from
  // one plane to the next and the loop needs no copies of in-flight values.
  __device__ __forceinline__ void step(MarchLds<T, RW, W> &L, int buf, int i, MarchH<T> (&hc)[S],
                                       MarchH<T> (&hn)[S], MarchE<T> (&ec)[RW],
                                       MarchE<T> (&en)[RW]) {
    const unsigned sxy = (unsigned)(g.ny * g.sk);
    const unsigned vs = vsb();
    const int ic = min(i, g.nx - 1);
    const bool live = i < iend;
    // keep these loads out of the previous plane's code, where the old
    // values of hn are still live
    __builtin_amdgcn_sched_barrier(0);
    load_h_all(hn, i + 1);
    __builtin_amdgcn_sched_barrier(0);

    // ── H(i) ──
    const PsiPlane pp = psi_plane(ic);
    const Prof1<T> pxh = prof1(g, prof, 1, 0, ic), pxe = prof1(g, prof, 0, 0, ic);
    const unsigned pl = (unsigned)ic * sxy * E;
    T h[S][3];
#pragma unroll
    for (int q = 0; q < S; ++q) {
      const MarchSlot &m = sl[q];
      const StripeIdx st = stripes(m, pp.inx);
      T p[6] = {0, 0, 0, 0, 0, 0};
      if constexpr (X) {
        p[0] = hc[q].p[0];
        p[1] = hc[q].p[1];
      }
      if constexpr (Y) {
        p[2] = hc[q].p[2];
        p[3] = hc[q].p[3];
      }
      if constexpr (Z) {
        p[4] = hc[q].p[4];
        p[5] = hc[q].p[5];
      }
      h[q][0] = hc[q].hx;
      h[q][1] = hc[q].hy;
      h[q][2] = hc[q].hz;
      // E(i) at j+1 and k+1 (written to LDS during the previous plane); past
      // the grid's last row / column they are clamped copies h_math ignores
      const T exj = (&L.e[buf][0][0][0])[m.le + LC];
      const T ezj = (&L.e[buf][2][0][0])[m.le + LC];
      const T exk = (&L.e[buf][0][0][0])[m.le + 1];
      const T eyk = (&L.e[buf][1][0][0])[m.le + 1];
      h_math(g, ic, m.j, m.k, st, pxh, pyh[q], pzh[q], oe[q][0], oe[q][1], oe[q][2], exj, ezj,
             exk, eyk, hc[q].ey, hc[q].ez, h[q][0], h[q][1], h[q][2], p, dtm_dl, dtm);
      (&L.h[buf][0][0][0])[m.lh] = h[q][0];
      (&L.h[buf][1][0][0])[m.lh] = h[q][1];
      (&L.h[buf][2][0][0])[m.lh] = h[q][2];
      if (q < RW) {
        const bool ok = m.own && live;
        const unsigned v = ok ? m.o4 + pl : padb;
        st1(3, h[q][0], v);
        st1(4, h[q][1], v);
        st1(5, h[q][2], v);
        if constexpr (X) {
          psi_st(p[0], 0, ok && pp.inx, m.o4 + pp.xs * E);
          psi_st(p[1], 1, ok && pp.inx, m.o4 + pp.xs * E);
        }
        if constexpr (Y) {
          psi_st(p[2], 2, ok && m.iny, m.qy4 + pp.ys * E);
          psi_st(p[3], 3, ok && m.iny, m.qy4 + pp.ys * E);
        }
        if constexpr (Z) {
          psi_st(p[4], 4, ok && m.inz, m.qz4 + pp.zs * E);
          psi_st(p[5], 5, ok && m.inz, m.qz4 + pp.zs * E);
        }
      }
    }
    // E(i+1) of every slot for the next plane's H neighbours
#pragma unroll
    for (int q = 0; q < S; ++q) {
      (&L.e[buf ^ 1][0][0][0])[sl[q].le] = hc[q].ex;
      (&L.e[buf ^ 1][1][0][0])[sl[q].le] = hc[q].ey;
      (&L.e[buf ^ 1][2][0][0])[sl[q].le] = hc[q].ez;
    }
    // the E part of plane i+1, in flight across the barrier and the E half
    __builtin_amdgcn_sched_barrier(0);
    load_e_all(en, i + 1);
    __builtin_amdgcn_sched_barrier(0);
    lds_barrier();

    // ── E(i) ──
#pragma unroll
    for (int q = 0; q < RW; ++q) {
      const MarchSlot &m = sl[q];
      const T hz_jm = (&L.h[buf][2][0][0])[m.lh - LC];
      const T hx_jm = (&L.h[buf][0][0][0])[m.lh - LC];
      const T hy_km = (&L.h[buf][1][0][0])[m.lh - 1], hx_km = (&L.h[buf][0][0][0])[m.lh - 1];
      const StripeIdx st = stripes(m, pp.inx);
      T p[6] = {0, 0, 0, 0, 0, 0};
      if constexpr (X) {
        p[0] = ec[q].p[0];
        p[1] = ec[q].p[1];
      }
      if constexpr (Y) {
        p[2] = ec[q].p[2];
        p[3] = ec[q].p[3];
      }
      if constexpr (Z) {
        p[4] = ec[q].p[4];
        p[5] = ec[q].p[5];
      }
      T ex = oe[q][0], ey = oe[q][1], ez = oe[q][2];
      T cx, cy, cz;
      if constexpr (CEM == kCeIndex) {
        const float *tb = mtab + mc[q] * (unsigned)nmat;
        cx = tb[ec[q].mi1];
        cy = tb[ec[q].mj1];
        cz = tb[ec[q].mk1];
        mc[q] = ec[q].mi1;  // the cell's material on the next plane
      } else if constexpr (CEM == kCeEps) {
        // on the grid's last plane, row or column the neighbour offsets are
        // clamped to the cell itself: 0.5 (e + e) = e, the coefficient
        // kernel's value there
        cx = ce_eps(mc[q], ec[q].mi1);
        cy = ce_eps(mc[q], ec[q].mj1);
        cz = ce_eps(mc[q], ec[q].mk1);
        mc[q] = ec[q].mi1;
      } else {
        cx = ec[q].cx;
        cy = ec[q].cy;
        cz = ec[q].cz;
      }
      // sources of this step on the cell, added before the update in list
      // order, as the reference adds them between its H and E kernels; the
      // plane and row tests are wave-uniform
      for (int t = 0; t < (ic >= src_i0 && ic < src_i1 ? nsrc : 0); ++t) {
        const auto *vp = lds_src_ptr(lsrc, t);
        if (ic < vp->ci0 || ic >= vp->ci1 || m.j < vp->cj0 || m.j >= vp->cj1) continue;
        const Src<T> sq = lds_src(lsrc, t);
        if (m.k >= sq.ck0 && m.k < sq.ck1) {
          const int c = sq.comp;
          const bool c0 = c == 0, c1 = c == 1;
          // ce of the component by masking bits: a select of the three values
          // is turned into an indexed load from a scratch copy of them
          const T ce = pick3(c0, c1, cx, cy, cz);
          const T v = src_term<T>(sq, ic, m.j, m.k, ce, sq.amp);
          ex = c0 ? ex + v : ex;
          ey = c1 ? ey + v : ey;
          ez = !c0 && !c1 ? ez + v : ez;
        }
      }
      const T ax = LOSSY ? ec[q].ax : T(1), ay = LOSSY ? ec[q].ay : T(1);
      const T az = LOSSY ? ec[q].az : T(1);
      e_math(g, ic, m.j, m.k, st, pxe, pye[q], pze[q], h[q][0], h[q][1], h[q][2], hx_jm, hz_jm,
             hx_km, hy_km, hyc[q], hzc[q], ex, ey, ez, ax, ay, az, cx, cy, cz, p, inv_dl);
      const bool ok = m.own && live;
      const unsigned v = ok ? m.o4 + pl : padb;
      st1(0, ex, v);
      st1(1, ey, v);
      st1(2, ez, v);
      if constexpr (X) {
        psi_st(p[0], 6, ok && pp.inx, m.o4 + pp.xs * E);
        psi_st(p[1], 7, ok && pp.inx, m.o4 + pp.xs * E);
      }
      if constexpr (Y) {
        psi_st(p[2], 8, ok && m.iny, m.qy4 + pp.ys * E);
        psi_st(p[3], 9, ok && m.iny, m.qy4 + pp.ys * E);
      }
      if constexpr (Z) {
        psi_st(p[4], 10, ok && m.inz, m.qz4 + pp.zs * E);
        psi_st(p[5], 11, ok && m.inz, m.qz4 + pp.zs * E);
      }
      if constexpr (CAP) bst_quad(rcap, ex, ey, h[q][0], h[q][1], zo[q], cap_plane(i, live));
      if constexpr (PRB) {
        if (ic == pr.i && live) pv = phit[q] ? (pr.comp == 0 ? ex : pr.comp == 1 ? ey : ez) : pv;
      }
      hyc[q] = h[q][1];
      hzc[q] = h[q][2];
    }
#pragma unroll
    for (int q = 0; q < S; ++q) {
      oe[q][0] = hc[q].ex;
      oe[q][1] = hc[q].ey;
      oe[q][2] = hc[q].ez;
    }
  }

  // Slot set-up, the carry H(i0-1) and the first loads, then the planes.
  __device__ __forceinline__ void run(MarchLds<T, RW, W> &L, int planes, int k0, int j0, int i0,
                                      int i1) {
    iend = i1;
    if constexpr (PRB) pv = T(0);
    sized = BIG ? (unsigned long long)g.vs * E < 0x80000000ull : slab_bytes<T>(g) < 0x80000000ull;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int n = g.npml;
    // halo slot lanes of wave w (QW = W/4 columns per wave, PE rows per
    // wave): 0 .. QW-1 H on row j0-1 (columns k0+QW w ...); 16 .. 16+RW G-1
    // H on column k0-1 (the wave's rows); 32 .. 32+QW-1 E only on row j0+TJ
    // (columns k0+QW w ...); 48 .. 48+PE-1 E only on column k0+W (rows
    // j0-1+PE w ...); lane 48+PE of wave 0 E only at (j0+TJ, k0-1); the rest
    // idle.  H of E-only lanes and both of idle lanes go to the trash row.
    constexpr int trash_row = TJ + 2;
    constexpr int QW = W / 4;
    constexpr int PE = (TJ + 2 + 3) / 4 > 3 ? (TJ + 2 + 3) / 4 : 3;
    const int tc = lane % LC;  // a trash-row column
    // fields a halo lane's role reads, bit c for Ex Ey Ez Hx Hy Hz: H on row
    // j0-1 is needed for Hx and Hz (E's j-1 neighbours), H on column k0-1
    // for Hx and Hy (E's k-1 neighbours), both from all of E; E-only lanes
    // need the components their H neighbours read across the tile edge
    unsigned need = 0;
#pragma unroll
    for (int q = 0; q <= RW; ++q) {
      int j, k, hr, hc, er, ec;
      bool own = false;
      if (q < RW) {
        const int lr = (RW * w + q) * G + lane / W, c = lane % W;  // tile row and column
        j = j0 + lr;
        k = k0 + c;
        hr = er = lr + 1;
        hc = ec = c + 1;
        own = j < g.ny && k < g.nz;
      } else if (lane < QW) {
        j = j0 - 1;
        k = k0 + QW * w + lane;
        hr = er = 0;
        hc = ec = QW * w + lane + 1;
        need = 0b101111;  // E, Hx, Hz
      } else if (lane >= 16 && lane < 16 + RW * G) {
        j = j0 + RW * G * w + lane - 16;
        k = k0 - 1;
        hr = er = RW * G * w + lane - 16 + 1;
        hc = ec = 0;
        need = 0b011111;  // E, Hx, Hy
      } else if (lane >= 32 && lane < 32 + QW) {
        j = j0 + TJ;
        k = k0 + QW * w + lane - 32;
        hr = trash_row;
        hc = tc;
        er = TJ + 1;
        ec = QW * w + lane - 32 + 1;
        need = 0b000101;  // Ex, Ez: j+1 neighbours of Hz, Hx on row j0+TJ-1
      } else if (lane >= 48 && lane < 48 + PE && PE * w + lane - 48 <= TJ) {
        j = j0 - 1 + PE * w + lane - 48;
        k = k0 + W;
        hr = trash_row;
        hc = tc;
        er = PE * w + lane - 48;
        ec = W + 1;
        need = 0b000011;  // Ex, Ey: k+1 neighbours of Hy, Hx on column k0+W-1
      } else if (lane == 48 + PE && w == 0) {
        j = j0 + TJ;
        k = k0 - 1;
        hr = trash_row;
        hc = tc;
        er = TJ + 1;
        ec = 0;
        need = 0b000100;  // Ez: j+1 neighbour of Hx at (j0+TJ-1, k0-1)
      } else {
        j = j0;
        k = k0;
        hr = er = trash_row;
        hc = ec = tc;
      }
      MarchSlot &m = sl[q];
      m.j = min(max(j, 0), g.ny - 1);
      m.k = min(max(k, 0), g.nz - 1);
      m.o4 = (unsigned)(m.j * g.sk + m.k) * E;
      m.o1 = (unsigned)(m.j * g.sk + m.k) * CB;
      m.o1j = (unsigned)(min(m.j + 1, g.ny - 1) * g.sk + m.k) * CB;
      m.o1k = (unsigned)(m.j * g.sk + min(m.k + 1, g.nz - 1)) * CB;
      // the halo slot's psi only for lanes that compute a needed H
      const bool hpsi = q < RW || (need & 0b011000) != 0;
      m.iny = Y && n > 0 && hpsi && (m.j < n || m.j >= g.ny - n);
      m.inz = Z && n > 0 && hpsi && (m.k < n || m.k >= g.nz - n);
      m.qy4 = m.iny ? (unsigned)(stripe_local(m.j, n, g.ny) * g.sk + m.k) * E : kDropOffset;
      m.qz4 = m.inz ? (unsigned)(m.j * 2 * n + stripe_local(m.k, n, g.nz)) * E : kDropOffset;
      m.lh = hr * LC + hc;
      m.le = er * LC + ec;
      m.own = own;
      if constexpr (CAP) {
        if (q < RW) {
          // record slot 2 f + 1 for the face's k, 2 f for its k - 1
          const int f = (k == zc.kf0 || k == zc.kf0 - 1) ? 0 : (k == zc.kf1 || k == zc.kf1 - 1) ? 1 : -1;
          const int kf = f == 0 ? zc.kf0 : zc.kf1;
          const bool rec = f >= 0 && own && j >= zc.j0 && j <= zc.j1;
          const unsigned P = (unsigned)(zc.i1 - zc.i0 + 1) * (unsigned)(zc.j1 - zc.j0 + 1);
          zo[q] = rec ? ((unsigned)(2 * f + (k == kf ? 1 : 0)) * P + (unsigned)(j - zc.j0)) * 16u
                      : kCapDropLane;
        }
      }
      if constexpr (PRB) {
        if (q < RW) phit[q] = own && j == pr.j && k == pr.k;
      }
      pyh[q] = prof1(g, prof, 1, 1, m.j);
      pye[q] = prof1(g, prof, 0, 1, m.j);
      pzh[q] = prof1(g, prof, 1, 2, m.k);
      pze[q] = prof1(g, prof, 0, 2, m.k);
    }
    // dropped by the hardware when the slab resource carries its size (slabs
    // below 2 GB), else this lane's slot in the volume padding
    if (sized) padb = kDropOffset;
    else padb = ((unsigned)((size_t)g.nx * g.ny * g.sk) + (unsigned)lane) * E;
    // the halo slot's field loads a lane's role does not read are dropped
    // (only with sized resources; a dropped load returns 0, and on a dead
    // plane, whose plane offset is already past the resource, the sum may
    // wrap to an address inside the slab: a load of valid memory whose
    // value is never used)
#pragma unroll
    for (int c = 0; c < 6; ++c) hoff[c] = !sized || ((need >> c) & 1) ? sl[RW].o4 : kDropOffset;
    hoffx = !sized || (need & 0b011000) ? sl[RW].o4 : kDropOffset;

    const unsigned vs = vsb();
    MarchH<T> A[S], B[S];
    MarchE<T> EA[RW], EB[RW];
    // E(i0) of every slot, then the loads of the first plane: they are in
    // flight while the carry below waits for its own loads
    {
      const unsigned p0 = plane_b(i0);
#pragma unroll
      for (int q = 0; q < S; ++q) {
        oe[q][0] = ld0(0, sl[q].o4, p0);
        oe[q][1] = ld0(1, sl[q].o4, p0);
        oe[q][2] = ld0(2, sl[q].o4, p0);
      }
      if constexpr (MAT) {
        const unsigned pm = mat_plane(i0, false);
#pragma unroll
        for (int q = 0; q < RW; ++q) mc[q] = ld_mat(sl[q].o1, pm);
      }
    }
    load_h_all(A, i0);
    load_e_all(EA, i0);
    // new H(i0-1) at the owned cells (unused when i0 == 0: e_math drops the
    // i-1 term on plane 0)
    {
      const int im = max(i0 - 1, 0);
      const unsigned pm = plane_b(im), nzb = (unsigned)g.sk * E;
      const PsiPlane pp = psi_plane(im);
      const Prof1<T> pxh = prof1(g, prof, 1, 0, im);
#pragma unroll
      for (int q = 0; q < RW; ++q) {
        const MarchSlot &m = sl[q];
        StripeIdx st = stripes(m, pp.inx);
        st.inx = pp.inx;  // the plane before the chunk may lie in an x-stripe in any variant
        T p[6] = {0, 0, 0, 0, 0, 0};
        if (st.inx) {
          p[0] = bld<T, false>(rq0, m.o4, psib(0, pp.xs));
          p[1] = bld<T, false>(rq0, m.o4, psib(1, pp.xs));
        }
        if (st.iny) {
          p[2] = bld<T, false>(rq0, m.qy4, psib(2, pp.ys));
          p[3] = bld<T, false>(rq0, m.qy4, psib(3, pp.ys));
        }
        if (st.inz) {
          p[4] = bld<T, false>(rq0, m.qz4, psib(4, pp.zs));
          p[5] = bld<T, false>(rq0, m.qz4, psib(5, pp.zs));
        }
        T hx = ld0(3, m.o4, pm), hy = ld0(4, m.o4, pm);
        T hz = ld0(5, m.o4, pm);
        h_math(g, im, m.j, m.k, st, pxh, pyh[q], pzh[q], ld0(0, m.o4, pm), ld0(1, m.o4, pm),
               ld0(2, m.o4, pm), ld0(0, m.o4, pm + nzb), ld0(2, m.o4, pm + nzb),
               ld0(0, m.o4 + E, pm), ld0(1, m.o4 + E, pm), oe[q][1],
               oe[q][2], hx, hy, hz, p, dtm_dl, dtm);
        hyc[q] = hy;
        hzc[q] = hz;
      }
    }
#pragma unroll
    for (int q = 0; q < S; ++q) {
      (&L.e[0][0][0][0])[sl[q].le] = oe[q][0];
      (&L.e[0][1][0][0])[sl[q].le] = oe[q][1];
      (&L.e[0][2][0][0])[sl[q].le] = oe[q][2];
    }
    lds_barrier();
    // the last plane's loads read the (clamped) plane after the chunk and go
    // unused; a peeled last pair makes the register allocator shuffle the
    // sets at the loop head, which waits for every load in flight
    int t = 0;  // planes >= 2 and even
#pragma unroll 1
    do {
      step(L, 0, i0 + t, A, B, EA, EB);
      step(L, 1, i0 + t + 1, B, A, EB, EA);
      t += 2;
    } while (t < planes);
    if constexpr (PRB) {
      bool hit = false;
#pragma unroll
      for (int q = 0; q < RW; ++q) hit = hit || phit[q];
      if (hit) {  // one lane of the grid: the probe's plane lies in [i0, i1)
        const double v = (double)pv;
        for (int f = 0; f < pr.nf; ++f) {
          pr.acc[2 * f] += v * pr.ph[2 * f];
          pr.acc[2 * f + 1] += v * pr.ph[2 * f + 1];
        }
      }
    }
  }
};

// Chunks of i planes: with PML, the x-stripes get chunks of their own
// ([0, n) and [nx - n, nx)) and the planes between them chunks of `planes`
// (the last one shorter), so only those chunks carry x-stripe psi.
struct MarchChunks {
  int split;   // 1: x-stripe chunks of their own
  int inner;   // chunks between the x-stripes (all chunks when split == 0)
  __device__ __forceinline__ void range(const Geo &g, int planes, int c, int &i0, int &i1) const {
    const int n = g.npml;
    if (!split) {
      i0 = c * planes;
      i1 = min(i0 + planes, g.nx);
    } else if (c == 0) {
      i0 = 0;
      i1 = n;
    } else if (c == inner + 1) {
      i0 = g.nx - n;
      i1 = g.nx;
    } else {
      i0 = n + (c - 1) * planes;
      i1 = min(i0 + planes, g.nx - n);
    }
  }
};

// One tile of the sweep: rows j0 ... j0+TJ-1, columns k0 ... k0+W-1,
// planes [i0, i1).  The block takes the variant with the psi it needs: X
// when its chunk meets an x-stripe, Y when its rows (with the halo row)
// meet a y-stripe, Z when its columns (with the halo column) meet a
// z-stripe.  (The plane before the chunk, recomputed once, reads its psi in
// any variant.)  Fewer psi operations per plane also keep a plane's loads
// and stores within the 63 a wavefront can count.  EXTRA names the one
// optional per-block feature a kernel carries: 1 the z-face record (taken
// by blocks that meet it), 2 the probe (taken by the block that owns its
// cell); `extra` says whether this block takes it.
template <typename T, int RW, bool LOSSY, bool NT, int CEM, bool BIG, int EXTRA, int W = 64>
__device__ __forceinline__ void march_tile(MarchLds<T, RW, W> &L, const Src<T> *lsrc, int nsrc,
                                           int si0, int si1, const float *ltab, const Geo &g,
                                           int k0, int j0, int i0, int i1, T dtm_dl, T dtm,
                                           T inv_dl, const T *__restrict__ s0,
                                           T *__restrict__ s1, const T *__restrict__ coef,
                                           const T *__restrict__ prof,
                                           const T *__restrict__ q0, T *__restrict__ q1,
                                           const void *__restrict__ mat, int nmat, double cscale,
                                           const ZCap &zc, bool extra, const MarchProbe &pr) {
  const int np = (i1 - i0 + 1) / 2 * 2;  // even; planes past i1 are masked
  const int n = g.npml;
  const bool x = n > 0 && (i0 < n || i1 > g.nx - n);
  constexpr int TJ = MarchShape<RW, W>::TJ;
  const bool y = n > 0 && (j0 - 1 < n || min(j0 + TJ - 1, g.ny - 1) >= g.ny - n);
  const bool z = n > 0 && (k0 - 1 < n || min(k0 + W - 1, g.nz - 1) >= g.nz - n);
  const unsigned long long sb = slab_bytes<T>(g);
  const int sbr = sb < 0x80000000ull ? (int)sb : -1;  // the coefficient slab has the same size
  const Rsrc r0 = __builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(s0), (short)0, sbr, 0x00020000);
  const Rsrc rc = __builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(coef), (short)0, sbr,
                                                    0x00020000);
  const Rsrc r1 = __builtin_amdgcn_make_buffer_rsrc(s1, (short)0, sbr, 0x00020000);
  const int psi_bytes = (int)((psi_slot(g, PSI_COUNT) + 64u) * (unsigned)sizeof(T));
  const Rsrc rq0 = __builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(q0), (short)0, psi_bytes,
                                                     0x00020000);
  const Rsrc rq1 = __builtin_amdgcn_make_buffer_rsrc(q1, (short)0, psi_bytes, 0x00020000);
  // the material index (one byte per cell) or the eps_r volume (four),
  // sized so loads of dead planes drop (the eps_r volume stays below 2 GB,
  // checked on the host)
  constexpr unsigned long long cb = CEM == kCeEps ? 4ull : 1ull;
  const Rsrc rmat = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<void *>(mat), (short)0,
      CEM != kCeVolumes ? (int)min(((unsigned long long)g.nx * g.ny * g.sk + 64ull) * cb, 0x7fffffffull)
                        : 0,
      0x00020000);
  // the z-face record (FP32)
  constexpr bool kCapable = EXTRA == 1 && sizeof(T) == 4 && !BIG;
  const Rsrc rcap = __builtin_amdgcn_make_buffer_rsrc(zc.buf, (short)0, (int)zc.bytes, 0x00020000);
#define HF_MARCH_RUN1(X, Y, Z, C, P)                                                           \
  {                                                                                           \
    March<T, RW, LOSSY, X, Y, Z, NT, CEM, C, BIG, P, W> m{g, dtm_dl, dtm, inv_dl, r0,    r1,   \
                                                       rc,  rq0,    rq1,  prof,   lsrc,   nsrc, \
                                                       si0,    si1,  rmat, ltab, nmat};         \
    m.cscale = cscale;                                                                        \
    m.s0p = s0;                                                                               \
    m.s1p = s1;                                                                               \
    m.cp = coef;                                                                              \
    m.rcap = rcap;                                                                            \
    m.zc = zc;                                                                                \
    m.pr = pr;                                                                                \
    m.run(L, np, k0, j0, i0, i1);                                                             \
  }
#define HF_MARCH_RUN(X, Y, Z)                                                                  \
  {                                                                                           \
    if constexpr (kCapable) {                                                                 \
      if (extra) HF_MARCH_RUN1(X, Y, Z, true, false) else HF_MARCH_RUN1(X, Y, Z, false, false) \
    } else if constexpr (EXTRA == 2) {                                                        \
      if (extra) HF_MARCH_RUN1(X, Y, Z, false, true) else HF_MARCH_RUN1(X, Y, Z, false, false) \
    } else {                                                                                  \
      HF_MARCH_RUN1(X, Y, Z, false, false)                                                    \
    }                                                                                         \
  }
  if (x) {
    if (y) {
      if (z) HF_MARCH_RUN(true, true, true) else HF_MARCH_RUN(true, true, false)
    } else {
      if (z) HF_MARCH_RUN(true, false, true) else HF_MARCH_RUN(true, false, false)
    }
  } else {
    if (y) {
      if (z) HF_MARCH_RUN(false, true, true) else HF_MARCH_RUN(false, true, false)
    } else {
      if (z) HF_MARCH_RUN(false, false, true) else HF_MARCH_RUN(false, false, false)
    }
  }
#undef HF_MARCH_RUN
#undef HF_MARCH_RUN1
}

// Tile B of a grid (k tiles fastest, then j tiles, then i chunks): its
// first column, first row and plane range.  XCD x runs the same j-tiles of
// every chunk unless the tile order rotates from chunk to chunk: then the
// CPML rows at both ends of j land on a different XCD in each chunk instead
// of loading the same two XCDs.
template <int RW, int W = 64>
__device__ __forceinline__ void march_tile_of(const Geo &g, int B, int planes, int tiles_k,
                                              int tiles_j, const MarchChunks &ch, int run,
                                              int &k0, int &j0, int &i0, int &i1) {
  const int bk = B % tiles_k, bc = B / (tiles_k * tiles_j);
  const int bj = ((B / tiles_k) % tiles_j + bc * max(1, run / tiles_k)) % tiles_j;
  k0 = bk * W;
  j0 = bj * MarchShape<RW, W>::TJ;
  ch.range(g, planes, bc, i0, i1);
}

// Planes on which one of the step's sources can reach a slot of the tile:
// the sources whose box meets rows j0-1 ... j0+TJ and columns k0-1 ...
// k0+W (every slot's clamped cell lies there); the per-cell tests of the
// sweep still decide, so a tile no source touches skips the source loop.
template <typename T, int RW, int W = 64>
__device__ __forceinline__ void march_src_window(const Src<T> &s, bool on, int k0, int j0,
                                                 int &si0, int &si1) {
  if (on && s.cj0 <= j0 + MarchShape<RW, W>::TJ && s.cj1 > j0 - 1 && s.ck0 <= k0 + W &&
      s.ck1 > k0 - 1) {
    si0 = min(si0, s.ci0);
    si1 = max(si1, s.ci1);
  }
}

// One block per (k tile, j tile, i chunk).
template <typename T, int RW, bool LOSSY, bool NT, int CEM, bool BIG = false, int W = 64>
__global__ __launch_bounds__(kMarchThreads, sizeof(T) == 4 && RW == 1 ? 4 : 2) void k_march(Geo g, int planes, T dtm_dl, T dtm,
                                                         T inv_dl, const T *__restrict__ s0,
                                                         T *__restrict__ s1,
                                                         const T *__restrict__ coef,
                                                         const T *__restrict__ prof,
                                                         const T *__restrict__ q0,
                                                         T *__restrict__ q1, int tiles_k,
                                                         int tiles_j, MarchChunks ch, int xcd,
                                                         int run, const MarchSources<T> src,
                                                         const void *__restrict__ mat,
                                                         const float *__restrict__ mtab,
                                                         int nmat, double cscale,
                                                         const ZCap zc, int nsweep,
                                                         const GatherTail gt) {
  if constexpr (sizeof(T) == 4 && !BIG) {
    // blocks past the sweep's: a held Huygens-box gather of the step before
    // (it reads s0 and the other record, which this sweep does not write)
    if ((int)blockIdx.x >= nsweep) {
      gather_block(gt, (int)blockIdx.x - nsweep, (int)threadIdx.x);
      return;
    }
  }
  __shared__ MarchLds<T, RW, W> L;
  __shared__ Src<T> lsrc[kMarchSources];
  __shared__ float ltab[CEM == kCeIndex ? kMatMax * kMatMax : 1];
  if (threadIdx.x == 0) {
#pragma unroll
    for (int t = 0; t < kMarchSources; ++t) lsrc[t] = src.s[t];
  }  // visible after the barrier that ends the prologue
  if constexpr (CEM == kCeIndex) {
    for (int t = threadIdx.x; t < nmat * nmat; t += kMarchThreads) ltab[t] = mtab[t];
  }
  const int B = xcd_runs(blockIdx.x, nsweep, xcd, run);
  int k0, j0, i0, i1;
  march_tile_of<RW, W>(g, B, planes, tiles_k, tiles_j, ch, run, k0, j0, i0, i1);
  int si0 = g.nx, si1 = 0;
#pragma unroll
  for (int t = 0; t < kMarchSources; ++t)
    march_src_window<T, RW, W>(src.s[t], t < src.n, k0, j0, si0, si1);
  // the z-face record (FP32): blocks whose columns, rows and planes meet it
  bool cap = false;
  if (sizeof(T) == 4 && !BIG && zc.buf) {
    auto col = [&](int c) { return c >= k0 && c < k0 + W; };
    cap = (col(zc.kf0) || col(zc.kf0 - 1) || col(zc.kf1) || col(zc.kf1 - 1)) &&
          j0 <= zc.j1 && j0 + MarchShape<RW, W>::TJ - 1 >= zc.j0 && i0 <= zc.i1 &&
          i1 - 1 >= zc.i0;
  }
  march_tile<T, RW, LOSSY, NT, CEM, BIG, 1, W>(L, lsrc, src.n, si0, si1, ltab, g, k0, j0, i0, i1,
                                            dtm_dl, dtm, inv_dl, s0, s1, coef, prof, q0, q1, mat,
                                            nmat, cscale, zc, cap, MarchProbe{});
}

}  // namespace hfdtd
