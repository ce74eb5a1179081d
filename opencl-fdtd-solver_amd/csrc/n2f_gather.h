// The Huygens-box gather's device code, shared by its own kernel
// (k_n2f_gather, n2f.hip) and by the march sweep, which runs a held gather
// of the step before in blocks after its own: face-sample decoding, the
// co-location of the tangential fields at the face centres
// (dft_farfield.cl:141-227), and one block of the gather into a K-step
// buffer row.
#pragma once

#include "common.h"

namespace hfdtd {

struct Box {
  int ix0, ix1, iy0, iy1, iz0, iz1;
  int off[6];
  int n_face;
};

// Packed sample -> face id and absolute (i, j, k) (dft_farfield.cl:141-164).
__device__ __forceinline__ int decode_face(const Box &b, int s, int &i, int &j, int &k) {
  const int nyf = b.iy1 - b.iy0 + 1, nzf = b.iz1 - b.iz0 + 1;
  int f, loc;
  if (s < b.off[2]) {
    f = s < b.off[1] ? 0 : 1;
    loc = s - b.off[f];
    i = f == 0 ? b.ix0 : b.ix1;
    j = b.iy0 + loc / nzf;
    k = b.iz0 + loc % nzf;
  } else if (s < b.off[4]) {
    f = s < b.off[3] ? 2 : 3;
    loc = s - b.off[f];
    i = b.ix0 + loc / nzf;
    j = f == 2 ? b.iy0 : b.iy1;
    k = b.iz0 + loc % nzf;
  } else {
    f = s < b.off[5] ? 4 : 5;
    loc = s - b.off[f];
    i = b.ix0 + loc / nyf;
    j = b.iy0 + loc % nyf;
    k = f == 4 ? b.iz0 : b.iz1;
  }
  return f;
}

struct Dims3 {
  int nx, ny, nz;
  int sk;  // elements between j rows of the field volumes (nz, or a row-padded run's)
};

template <typename T>
__device__ __forceinline__ T avg2(T a, T b, bool ok) {
  return ok ? T(0.5) * (a + b) : a;
}

template <typename T>
__device__ __forceinline__ T avg4(T a, T b, T c, T d) {
  return T(0.25) * (a + b + c + d);
}

// Tangential E/H of one face sample co-located at the face centre
// (dft_farfield.cl:171-227): E as 2-point, H as 4-point averages with
// clamped neighbours.  c[] receives the component ids (0..5 = Ex..Hz).
template <typename T>
__device__ __forceinline__ void colocate_impl(const Dims3 &g, int f, int i, int j, int k,
                                              const T *Ex, const T *Ey, const T *Ez,
                                              const T *Hx, const T *Hy, const T *Hz, T v[4],
                                              int c[4]) {
  auto at = [&](const T *F, int a, int b, int cc) {
    return F[((size_t)a * g.ny + b) * g.sk + cc];
  };
  if (f <= 1) {
    const int k1 = k + 1 < g.nz ? k + 1 : k;
    const int j1 = j + 1 < g.ny ? j + 1 : j;
    const int il = i > 0 ? i - 1 : i;
    v[0] = avg2(at(Ey, i, j, k), at(Ey, i, j, k1), k1 != k);
    v[1] = avg2(at(Ez, i, j, k), at(Ez, i, j1, k), j1 != j);
    v[2] = avg4(at(Hy, i, j, k), at(Hy, il, j, k), at(Hy, i, j1, k), at(Hy, il, j1, k));
    v[3] = avg4(at(Hz, i, j, k), at(Hz, il, j, k), at(Hz, i, j, k1), at(Hz, il, j, k1));
    c[0] = 1; c[1] = 2; c[2] = 4; c[3] = 5;
  } else if (f <= 3) {
    const int i1 = i + 1 < g.nx ? i + 1 : i;
    const int k1 = k + 1 < g.nz ? k + 1 : k;
    const int jl = j > 0 ? j - 1 : j;
    v[0] = avg2(at(Ex, i, j, k), at(Ex, i, j, k1), k1 != k);
    v[1] = avg2(at(Ez, i, j, k), at(Ez, i1, j, k), i1 != i);
    v[2] = avg4(at(Hx, i, j, k), at(Hx, i, jl, k), at(Hx, i1, j, k), at(Hx, i1, jl, k));
    v[3] = avg4(at(Hz, i, j, k), at(Hz, i, jl, k), at(Hz, i, j, k1), at(Hz, i, jl, k1));
    c[0] = 0; c[1] = 2; c[2] = 3; c[3] = 5;
  } else {
    const int i1 = i + 1 < g.nx ? i + 1 : i;
    const int j1 = j + 1 < g.ny ? j + 1 : j;
    const int kl = k > 0 ? k - 1 : k;
    v[0] = avg2(at(Ex, i, j, k), at(Ex, i, j1, k), j1 != j);
    v[1] = avg2(at(Ey, i, j, k), at(Ey, i1, j, k), i1 != i);
    v[2] = avg4(at(Hx, i, j, k), at(Hx, i, j, kl), at(Hx, i, j1, k), at(Hx, i, j1, kl));
    v[3] = avg4(at(Hy, i, j, k), at(Hy, i, j, kl), at(Hy, i1, j, k), at(Hy, i1, j, kl));
    c[0] = 0; c[1] = 1; c[2] = 3; c[3] = 4;
  }
}

struct PhaseRowEH {
  float e[32], h[32];  // cos, sin pairs (times dt); zero beyond 2 nfreq
};

// The engine's z-face record as the gather reads it (buf == nullptr: read
// the z faces from the fields).
struct ZRead {
  const float4 *buf;
  int i0, j0, nj;
  long long quads;  // quads per (face, column) array
};

// Co-located z-face sample from the record: the same values and the same
// operations as colocate_impl's z branch (the faces' k are at least 1, so
// its k - 1 is the record's first column).
__device__ __forceinline__ void colocate_zrec(const Dims3 &g, const ZRead &z, int f, int i, int j,
                                              float v[4]) {
  const int i1 = i + 1 < g.nx ? i + 1 : i;
  const int j1 = j + 1 < g.ny ? j + 1 : j;
  auto at = [&](int col, int a, int bj) {  // (Ex, Ey, Hx, Hy) at column k - 1 + col
    return z.buf[(size_t)(2 * (f - 4) + col) * z.quads + (size_t)(a - z.i0) * z.nj + (bj - z.j0)];
  };
  const float4 c0 = at(1, i, j), cj = at(1, i, j1), ci = at(1, i1, j);
  const float4 l0 = at(0, i, j), lj = at(0, i, j1), li = at(0, i1, j);
  v[0] = avg2(c0.x, cj.x, j1 != j);
  v[1] = avg2(c0.y, ci.y, i1 != i);
  v[2] = avg4(c0.z, l0.z, cj.z, lj.z);
  v[3] = avg4(c0.w, l0.w, ci.w, li.w);
}

// One step's gather into a K-step buffer row: its launch arguments, also
// held by the engine until the next march sweep runs it (Engine::pgather).
struct GatherTail {
  bool held = false;  // a gather waits for the next launch
  Dims3 g;
  Box b;
  const float *f[6];  // Ex .. Hz as the step left them
  float *row;         // row kfill of the buffer
  long long pe;       // start of the H rows
  PhaseRowEH ph;
  float *we, *wh;     // row kfill of the phase rows
  ZRead zr;
  int blocks() const { return (int)ceil_div((size_t)b.n_face, 256); }
};

// Block blk of a gather (256 threads): the co-located samples s = 256 blk +
// thread into the row, and block 0 copies the phase row.
__device__ __forceinline__ void gather_block(const GatherTail &t, int blk, int tid) {
  if (blk == 0 && tid < 32) {
    t.we[tid] = t.ph.e[tid];
    t.wh[tid] = t.ph.h[tid];
  }
  const int s = blk * 256 + tid;
  if (s >= t.b.n_face) return;
  int i, j, k;
  const int f = decode_face(t.b, s, i, j, k);
  float v[4];
  int c[4];
  if (t.zr.buf && f >= 4)
    colocate_zrec(t.g, t.zr, f, i, j, v);
  else
    colocate_impl<float>(t.g, f, i, j, k, t.f[0], t.f[1], t.f[2], t.f[3], t.f[4], t.f[5], v, c);
  t.row[2 * s] = v[0];
  t.row[2 * s + 1] = v[1];
  t.row[t.pe + 2 * s] = v[2];
  t.row[t.pe + 2 * s + 1] = v[3];
}

}  // namespace hfdtd
