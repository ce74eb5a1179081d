// Yee H/E updates with face-stripe CPML, source injection, and the fused
// H+sources+E sweep.
//
// Per-cell arithmetic follows the reference kernels term for term
// (kernels/yee_update.cl:11-238, kernels/sources.cl:7-49): interior cells use
// the folded dtm_dl / inv_dl form, PML-shell cells the kappa-scaled form with
// psi recursions on the face stripes.  Both the split sweeps and the fused
// sweep call the same h_cell / e_cell device functions, so they produce the
// same values.
//
// Layout: C-order (Nx, Ny, Nz), idx = (i*Ny + j)*Nz + k, k contiguous.  One
// wavefront (64 lanes) spans 64 consecutive k, so every field load and store
// is a 256-byte coalesced segment (FP32).

#include "common.h"

namespace hfdtd {

struct Dims {
  int nx, ny, nz, npml;
};

template <typename T>
struct Prof {  // one stagger: b, c, 1/(kappa dl) per axis
  const T *b[3];
  const T *c[3];
  const T *ik[3];
};

template <typename T>
struct Psi6 {  // H: Hy_x Hz_x Hx_y Hz_y Hx_z Hy_z | E: Ey_x Ez_x Ex_y Ez_y Ex_z Ey_z
  T *p[6];
};

// local stripe offsets (yee_update.cl:152, 160, 168)
__device__ __forceinline__ int stripe_local(int i, int n, int N) { return i < n ? i : n + i - (N - n); }

__device__ __forceinline__ bool is_interior(const Dims &d, int i, int j, int k) {
  const int n = d.npml;
  return n > 0 && i >= n && i < d.nx - n && j >= n && j < d.ny - n && k >= n && k < d.nz - n;
}

// H at (i,j,k) from E.  e_*: E at the cell; ex_j1/ez_j1 at j+1; ex_k1/ey_k1 at
// k+1; ey_i1/ez_i1 at i+1.  Neighbour values past the last index are ignored
// (forward difference 0, yee_update.cl:136-141).  psi is read from pin and,
// when WRITE, stored to pout (pin == pout for the in-place split sweep).
template <typename T, bool WRITE>
__device__ __forceinline__ void h_cell(const Dims &d, int i, int j, int k, T ex, T ey, T ez,
                                       T ex_j1, T ez_j1, T ex_k1, T ey_k1, T ey_i1, T ez_i1,
                                       T &hx, T &hy, T &hz, T dtm_dl, T dtm, const Prof<T> &P,
                                       const Psi6<T> &pin, const Psi6<T> &pout) {
  if (is_interior(d, i, j, k)) {  // yee_update.cl:11-40
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
  // PML shell or npml == 0 (yee_update.cl:42-118)
  const int n = d.npml;
  const T dEz_dy = j < d.ny - 1 ? ez_j1 - ez : T(0);
  const T dEy_dz = k < d.nz - 1 ? ey_k1 - ey : T(0);
  const T dEx_dz = k < d.nz - 1 ? ex_k1 - ex : T(0);
  const T dEz_dx = i < d.nx - 1 ? ez_i1 - ez : T(0);
  const T dEy_dx = i < d.nx - 1 ? ey_i1 - ey : T(0);
  const T dEx_dy = j < d.ny - 1 ? ex_j1 - ex : T(0);
  T pHx_y = 0, pHx_z = 0, pHy_x = 0, pHy_z = 0, pHz_x = 0, pHz_y = 0;
  if (i < n || i >= d.nx - n) {
    const int il = stripe_local(i, n, d.nx);
    const size_t q = ((size_t)il * d.ny + j) * d.nz + k;
    const T b = P.b[0][i], c = P.c[0][i];
    pHy_x = b * pin.p[0][q] + c * dEz_dx;
    pHz_x = b * pin.p[1][q] + c * dEy_dx;
    if (WRITE) {
      pout.p[0][q] = pHy_x;
      pout.p[1][q] = pHz_x;
    }
  }
  if (j < n || j >= d.ny - n) {
    const int jl = stripe_local(j, n, d.ny);
    const size_t q = ((size_t)i * 2 * n + jl) * d.nz + k;
    const T b = P.b[1][j], c = P.c[1][j];
    pHx_y = b * pin.p[2][q] + c * dEz_dy;
    pHz_y = b * pin.p[3][q] + c * dEx_dy;
    if (WRITE) {
      pout.p[2][q] = pHx_y;
      pout.p[3][q] = pHz_y;
    }
  }
  if (k < n || k >= d.nz - n) {
    const int kl = stripe_local(k, n, d.nz);
    const size_t q = ((size_t)i * d.ny + j) * (2 * n) + kl;
    const T b = P.b[2][k], c = P.c[2][k];
    pHx_z = b * pin.p[4][q] + c * dEy_dz;
    pHy_z = b * pin.p[5][q] + c * dEx_dz;
    if (WRITE) {
      pout.p[4][q] = pHx_z;
      pout.p[5][q] = pHy_z;
    }
  }
  const T ikx = P.ik[0][i], iky = P.ik[1][j], ikz = P.ik[2][k];
  hx -= dtm * (dEz_dy * iky + pHx_y - dEy_dz * ikz - pHx_z);
  hy -= dtm * (dEx_dz * ikz + pHy_z - dEz_dx * ikx - pHy_x);
  hz -= dtm * (dEy_dx * ikx + pHz_x - dEx_dy * iky - pHz_y);
}

// E at (i,j,k) from H (already advanced).  h*: H at the cell; hx_jm/hz_jm at
// j-1; hx_km/hy_km at k-1; hy_im/hz_im at i-1 (backward difference 0 on
// index 0, yee_update.cl:256-261).
template <typename T>
__device__ __forceinline__ void e_cell(const Dims &d, int i, int j, int k, T hx, T hy, T hz,
                                       T hx_jm, T hz_jm, T hx_km, T hy_km, T hy_im, T hz_im,
                                       T &ex, T &ey, T &ez, T cax, T cay, T caz, T cex, T cey,
                                       T cez, T inv_dl, const Prof<T> &P, const Psi6<T> &pin,
                                       const Psi6<T> &pout) {
  if (is_interior(d, i, j, k)) {  // yee_update.cl:120-155
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
  const int n = d.npml;  // yee_update.cl:157-238
  const T dHz_dy = j > 0 ? hz - hz_jm : T(0);
  const T dHy_dz = k > 0 ? hy - hy_km : T(0);
  const T dHx_dz = k > 0 ? hx - hx_km : T(0);
  const T dHz_dx = i > 0 ? hz - hz_im : T(0);
  const T dHy_dx = i > 0 ? hy - hy_im : T(0);
  const T dHx_dy = j > 0 ? hx - hx_jm : T(0);
  T pEx_y = 0, pEx_z = 0, pEy_x = 0, pEy_z = 0, pEz_x = 0, pEz_y = 0;
  if (i < n || i >= d.nx - n) {
    const int il = stripe_local(i, n, d.nx);
    const size_t q = ((size_t)il * d.ny + j) * d.nz + k;
    const T b = P.b[0][i], c = P.c[0][i];
    pEy_x = b * pin.p[0][q] + c * dHz_dx;
    pEz_x = b * pin.p[1][q] + c * dHy_dx;
    pout.p[0][q] = pEy_x;
    pout.p[1][q] = pEz_x;
  }
  if (j < n || j >= d.ny - n) {
    const int jl = stripe_local(j, n, d.ny);
    const size_t q = ((size_t)i * 2 * n + jl) * d.nz + k;
    const T b = P.b[1][j], c = P.c[1][j];
    pEx_y = b * pin.p[2][q] + c * dHz_dy;
    pEz_y = b * pin.p[3][q] + c * dHx_dy;
    pout.p[2][q] = pEx_y;
    pout.p[3][q] = pEz_y;
  }
  if (k < n || k >= d.nz - n) {
    const int kl = stripe_local(k, n, d.nz);
    const size_t q = ((size_t)i * d.ny + j) * (2 * n) + kl;
    const T b = P.b[2][k], c = P.c[2][k];
    pEx_z = b * pin.p[4][q] + c * dHy_dz;
    pEy_z = b * pin.p[5][q] + c * dHx_dz;
    pout.p[4][q] = pEx_z;
    pout.p[5][q] = pEy_z;
  }
  const T ikx = P.ik[0][i], iky = P.ik[1][j], ikz = P.ik[2][k];
  ex = cax * ex + cex * (dHz_dy * iky + pEx_y - dHy_dz * ikz - pEx_z);
  ey = cay * ey + cey * (dHx_dz * ikz + pEy_z - dHz_dx * ikx - pEy_x);
  ez = caz * ez + cez * (dHy_dx * ikx + pEz_x - dHx_dy * iky - pEz_y);
}

// ── split sweeps (reference order: all of H, then sources, then all of E) ──

constexpr int kSplitRows = 4;

template <typename T>
__global__ __launch_bounds__(64 * kSplitRows) void k_update_H(
    Dims d, T dtm_dl, T dtm, const T *__restrict__ Ex, const T *__restrict__ Ey,
    const T *__restrict__ Ez, T *__restrict__ Hx, T *__restrict__ Hy, T *__restrict__ Hz,
    Prof<T> P, Psi6<T> psi) {
  const int k = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * kSplitRows + threadIdx.y;
  const int i = blockIdx.z;
  if (k >= d.nz || j >= d.ny) return;
  const size_t sxy = (size_t)d.ny * d.nz;
  const size_t idx = (size_t)i * sxy + (size_t)j * d.nz + k;
  const T ex = Ex[idx], ey = Ey[idx], ez = Ez[idx];
  const bool jn = j + 1 < d.ny, kn = k + 1 < d.nz, in = i + 1 < d.nx;
  const T ex_j1 = jn ? Ex[idx + d.nz] : T(0);
  const T ez_j1 = jn ? Ez[idx + d.nz] : T(0);
  const T ex_k1 = kn ? Ex[idx + 1] : T(0);
  const T ey_k1 = kn ? Ey[idx + 1] : T(0);
  const T ey_i1 = in ? Ey[idx + sxy] : T(0);
  const T ez_i1 = in ? Ez[idx + sxy] : T(0);
  T hx = Hx[idx], hy = Hy[idx], hz = Hz[idx];
  h_cell<T, true>(d, i, j, k, ex, ey, ez, ex_j1, ez_j1, ex_k1, ey_k1, ey_i1, ez_i1, hx, hy, hz,
                  dtm_dl, dtm, P, psi, psi);
  Hx[idx] = hx;
  Hy[idx] = hy;
  Hz[idx] = hz;
}

template <typename T, bool LOSSY>
__global__ __launch_bounds__(64 * kSplitRows) void k_update_E(
    Dims d, T inv_dl, const T *__restrict__ Hx, const T *__restrict__ Hy,
    const T *__restrict__ Hz, T *__restrict__ Ex, T *__restrict__ Ey, T *__restrict__ Ez,
    const T *__restrict__ cex, const T *__restrict__ cey, const T *__restrict__ cez,
    const T *__restrict__ cax, const T *__restrict__ cay, const T *__restrict__ caz, Prof<T> P,
    Psi6<T> psi) {
  const int k = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * kSplitRows + threadIdx.y;
  const int i = blockIdx.z;
  if (k >= d.nz || j >= d.ny) return;
  const size_t sxy = (size_t)d.ny * d.nz;
  const size_t idx = (size_t)i * sxy + (size_t)j * d.nz + k;
  const T hx = Hx[idx], hy = Hy[idx], hz = Hz[idx];
  const T hx_jm = j > 0 ? Hx[idx - d.nz] : T(0);
  const T hz_jm = j > 0 ? Hz[idx - d.nz] : T(0);
  const T hx_km = k > 0 ? Hx[idx - 1] : T(0);
  const T hy_km = k > 0 ? Hy[idx - 1] : T(0);
  const T hy_im = i > 0 ? Hy[idx - sxy] : T(0);
  const T hz_im = i > 0 ? Hz[idx - sxy] : T(0);
  T ex = Ex[idx], ey = Ey[idx], ez = Ez[idx];
  const T ax = LOSSY ? cax[idx] : T(1), ay = LOSSY ? cay[idx] : T(1), az = LOSSY ? caz[idx] : T(1);
  e_cell<T>(d, i, j, k, hx, hy, hz, hx_jm, hz_jm, hx_km, hy_km, hy_im, hz_im, ex, ey, ez, ax, ay,
            az, cex[idx], cey[idx], cez[idx], inv_dl, P, psi, psi);
  Ex[idx] = ex;
  Ey[idx] = ey;
  Ez[idx] = ez;
}

// ── sources ───────────────────────────────────────────────────────────────

constexpr int kMaxSources = 16;

template <typename T>
struct Src {
  int comp, mode, taper;
  int i0, i1, j0, j1, k0, k1;     // unclipped (edge tests)
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
__device__ __forceinline__ T src_term(const Src<T> &s, int i, int j, int k, T ce) {
  if (s.mode == HFDTD_SRC_SOFT) return s.amp;
  T w = T(1);
  if ((s.taper & 1) && (i == s.i0 || i == s.i1 - 1)) w *= s.re;
  if ((s.taper & 2) && (j == s.j0 || j == s.j1 - 1)) w *= s.re;
  if ((s.taper & 4) && (k == s.k0 || k == s.k1 - 1)) w *= s.re;
  return -ce * s.amp * w;
}

template <typename T>
__global__ void k_add_source(Dims d, Src<T> s, T *__restrict__ F, const T *__restrict__ ce) {
  const int ni = s.ci1 - s.ci0, nj = s.cj1 - s.cj0, nk = s.ck1 - s.ck0;
  const size_t total = (size_t)ni * nj * nk;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int k = s.ck0 + (int)(t % nk);
  const int j = s.cj0 + (int)((t / nk) % nj);
  const int i = s.ci0 + (int)(t / ((size_t)nk * nj));
  const size_t idx = ((size_t)i * d.ny + j) * d.nz + k;
  F[idx] += src_term<T>(s, i, j, k, ce[idx]);
}

template <typename T>
static bool make_src(const Engine *e, const hfdtd_source &in, Src<T> &o) {
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

// ── fused sweep ─────────────────────────────────────────────────────────────
//
// One block owns a 64(k) x TJ(j) column tile and marches along i over
// `planes` planes.  State is ping-ponged (reads *0, writes *1), so values a
// neighbouring block needs are never overwritten while it runs.  Per plane:
//   1. E(i+1) of the own cell is prefetched into registers (it is E(i) of the
//      next plane, so every E value is read from HBM once);
//   2. H(i) of the tile is computed from E(i) (LDS tile + global for the
//      +1 row/column outside the tile) and written out;
//   3. H(i) on the j0-1 row and k0-1 column is recomputed (not written) so the
//      E update needs no other block's output;
//   4. sources are added to E(i) and E(i) is advanced from H(i) (LDS) and
//      H(i-1) (registers).
// HBM traffic per cell is the compulsory one: read E, H, ce (ca when lossy),
// psi; write E, H, psi.

template <typename T, int TJ, bool LOSSY>
__global__ __launch_bounds__(64 * TJ) void k_fused(
    Dims d, int planes, T dtm_dl, T dtm, T inv_dl, const T *__restrict__ Ex0,
    const T *__restrict__ Ey0, const T *__restrict__ Ez0, const T *__restrict__ Hx0,
    const T *__restrict__ Hy0, const T *__restrict__ Hz0, T *__restrict__ Ex1,
    T *__restrict__ Ey1, T *__restrict__ Ez1, T *__restrict__ Hx1, T *__restrict__ Hy1,
    T *__restrict__ Hz1, const T *__restrict__ cex, const T *__restrict__ cey,
    const T *__restrict__ cez, const T *__restrict__ cax, const T *__restrict__ cay,
    const T *__restrict__ caz, Prof<T> PH, Prof<T> PE, Psi6<T> pH0, Psi6<T> pH1, Psi6<T> pE0,
    Psi6<T> pE1, SrcList<T> src) {
  __shared__ T sE[3][TJ][64];
  __shared__ T sH[3][TJ + 1][65];

  const int tk = threadIdx.x, tj = threadIdx.y;
  const int k0 = blockIdx.x * 64, j0 = blockIdx.y * TJ;
  const int k = k0 + tk, j = j0 + tj;
  const int ib = blockIdx.z * planes;
  const int ie = min(ib + planes, d.nx);
  const bool act = k < d.nz && j < d.ny;
  const size_t sxy = (size_t)d.ny * d.nz;
  const size_t nz = (size_t)d.nz;
  const bool jn = j + 1 < d.ny, kn = k + 1 < d.nz;

  // sources whose (j, k) footprint contains this column
  unsigned smask = 0;
  if (act)
    for (int s = 0; s < src.n; ++s)
      if (j >= src.s[s].cj0 && j < src.s[s].cj1 && k >= src.s[s].ck0 && k < src.s[s].ck1)
        smask |= 1u << s;

  size_t idx = (size_t)ib * sxy + (size_t)j * nz + k;
  T e_x = 0, e_y = 0, e_z = 0;  // E(i) of the own cell (pre-source)
  T hyp = 0, hzp = 0;           // H(i-1) of the own cell, advanced
  if (act) {
    e_x = Ex0[idx];
    e_y = Ey0[idx];
    e_z = Ez0[idx];
    if (ib > 0) {  // recompute H(ib-1) of the own cell (not written)
      const size_t q = idx - sxy;
      const T px = Ex0[q], py = Ey0[q], pz = Ez0[q];
      T hx = Hx0[q], hy = Hy0[q], hz = Hz0[q];
      h_cell<T, false>(d, ib - 1, j, k, px, py, pz, jn ? Ex0[q + nz] : T(0),
                       jn ? Ez0[q + nz] : T(0), kn ? Ex0[q + 1] : T(0), kn ? Ey0[q + 1] : T(0),
                       e_y, e_z, hx, hy, hz, dtm_dl, dtm, PH, pH0, pH0);
      hyp = hy;
      hzp = hz;
    }
  }

  for (int i = ib; i < ie; ++i, idx += sxy) {
    sE[0][tj][tk] = e_x;
    sE[1][tj][tk] = e_y;
    sE[2][tj][tk] = e_z;
    const bool in = i + 1 < d.nx;
    T n_x = 0, n_y = 0, n_z = 0, hx = 0, hy = 0, hz = 0;
    if (act) {
      if (in) {
        n_x = Ex0[idx + sxy];
        n_y = Ey0[idx + sxy];
        n_z = Ez0[idx + sxy];
      }
      hx = Hx0[idx];
      hy = Hy0[idx];
      hz = Hz0[idx];
    }
    __syncthreads();

    // 2. own H(i)
    if (act) {
      const T ex_j1 = tj + 1 < TJ ? sE[0][tj + 1][tk] : (jn ? Ex0[idx + nz] : T(0));
      const T ez_j1 = tj + 1 < TJ ? sE[2][tj + 1][tk] : (jn ? Ez0[idx + nz] : T(0));
      const T ex_k1 = tk + 1 < 64 ? sE[0][tj][tk + 1] : (kn ? Ex0[idx + 1] : T(0));
      const T ey_k1 = tk + 1 < 64 ? sE[1][tj][tk + 1] : (kn ? Ey0[idx + 1] : T(0));
      h_cell<T, true>(d, i, j, k, e_x, e_y, e_z, ex_j1, ez_j1, ex_k1, ey_k1, n_y, n_z, hx, hy,
                      hz, dtm_dl, dtm, PH, pH0, pH1);
      Hx1[idx] = hx;
      Hy1[idx] = hy;
      Hz1[idx] = hz;
    }
    sH[0][tj + 1][tk + 1] = hx;
    sH[1][tj + 1][tk + 1] = hy;
    sH[2][tj + 1][tk + 1] = hz;

    // 3a. halo row j0-1 (wave tj == 0)
    if (tj == 0 && j0 > 0 && k < d.nz) {
      const size_t q = idx - nz;  // (i, j0-1, k)
      const T ax = Ex0[q], ay = Ey0[q], az = Ez0[q];
      T gx = Hx0[q], gy = Hy0[q], gz = Hz0[q];
      const T a_kx = kn ? Ex0[q + 1] : T(0), a_ky = kn ? Ey0[q + 1] : T(0);
      const T a_iy = in ? Ey0[q + sxy] : T(0), a_iz = in ? Ez0[q + sxy] : T(0);
      // E at (i, j0, k) is the own E(i) of this lane (row j0 exists: j0 < ny)
      h_cell<T, false>(d, i, j0 - 1, k, ax, ay, az, e_x, e_z, a_kx, a_ky, a_iy, a_iz, gx, gy, gz,
                       dtm_dl, dtm, PH, pH0, pH0);
      sH[0][0][tk + 1] = gx;
      sH[1][0][tk + 1] = gy;
      sH[2][0][tk + 1] = gz;
    }
    // 3b. halo column k0-1 (lane tk == 0 of each row)
    if (tk == 0 && k0 > 0 && j < d.ny) {
      const size_t q = idx - 1;  // (i, j, k0-1)
      const T ax = Ex0[q], ay = Ey0[q], az = Ez0[q];
      T gx = Hx0[q], gy = Hy0[q], gz = Hz0[q];
      const T a_jx = jn ? Ex0[q + nz] : T(0), a_jz = jn ? Ez0[q + nz] : T(0);
      const T a_iy = in ? Ey0[q + sxy] : T(0), a_iz = in ? Ez0[q + sxy] : T(0);
      // E at (i, j, k0) is the own E(i) of this lane
      h_cell<T, false>(d, i, j, k0 - 1, ax, ay, az, a_jx, a_jz, e_x, e_y, a_iy, a_iz, gx, gy, gz,
                       dtm_dl, dtm, PH, pH0, pH0);
      sH[0][tj + 1][0] = gx;
      sH[1][tj + 1][0] = gy;
      sH[2][tj + 1][0] = gz;
    }
    __syncthreads();

    // 4. sources then E(i)
    if (act) {
      const T cx = cex[idx], cy = cey[idx], cz = cez[idx];
      if (smask) {
        for (int s = 0; s < src.n; ++s) {
          if (!((smask >> s) & 1u) || i < src.s[s].ci0 || i >= src.s[s].ci1) continue;
          const int c = src.s[s].comp;
          const T ce = c == 0 ? cx : (c == 1 ? cy : cz);
          const T v = src_term<T>(src.s[s], i, j, k, ce);
          if (c == 0) e_x += v;
          else if (c == 1) e_y += v;
          else e_z += v;
        }
      }
      const T ax = LOSSY ? cax[idx] : T(1), ay = LOSSY ? cay[idx] : T(1),
              az = LOSSY ? caz[idx] : T(1);
      e_cell<T>(d, i, j, k, hx, hy, hz, sH[0][tj][tk + 1], sH[2][tj][tk + 1], sH[0][tj + 1][tk],
                sH[1][tj + 1][tk], hyp, hzp, e_x, e_y, e_z, ax, ay, az, cx, cy, cz, inv_dl, PE,
                pE0, pE1);
      Ex1[idx] = e_x;
      Ey1[idx] = e_y;
      Ez1[idx] = e_z;
    }
    hyp = hy;
    hzp = hz;
    e_x = n_x;
    e_y = n_y;
    e_z = n_z;
  }
}

// ── launchers ─────────────────────────────────────────────────────────────

template <typename T>
static Prof<T> prof_of(const Engine *e, int stag) {
  Prof<T> p;
  for (int a = 0; a < 3; ++a) {
    p.b[a] = (const T *)e->prof[stag][a][0];
    p.c[a] = (const T *)e->prof[stag][a][1];
    p.ik[a] = (const T *)e->prof[stag][a][2];
  }
  return p;
}

template <typename T>
static Psi6<T> psi_of(void *const *slots, int base) {
  Psi6<T> p;
  for (int s = 0; s < 6; ++s) p.p[s] = (T *)slots[base + s];
  return p;
}

static Dims dims_of(const Engine *e) { return Dims{e->nx, e->ny, e->nz, e->npml}; }

template <typename T>
static int update_H_t(Engine *e) {
  dim3 grid(ceil_div(e->nz, 64), ceil_div(e->ny, kSplitRows), e->nx);
  int rc = timing_mark(e);
  if (rc) return rc;
  hipLaunchKernelGGL(k_update_H<T>, grid, dim3(64, kSplitRows), 0, e->stream, dims_of(e),
                     T(e->dtm_dl), T(e->dtm), (const T *)e->field[0], (const T *)e->field[1],
                     (const T *)e->field[2], (T *)e->field[3], (T *)e->field[4],
                     (T *)e->field[5], prof_of<T>(e, 1), psi_of<T>(e->psi, 0));
  HF_HIP(hipGetLastError());
  return timing_mark(e);
}

template <typename T>
static int update_E_t(Engine *e) {
  dim3 grid(ceil_div(e->nz, 64), ceil_div(e->ny, kSplitRows), e->nx);
  auto kern = e->lossless ? k_update_E<T, false> : k_update_E<T, true>;
  int rc = timing_mark(e);
  if (rc) return rc;
  hipLaunchKernelGGL(kern, grid, dim3(64, kSplitRows), 0, e->stream, dims_of(e), T(e->inv_dl),
                     (const T *)e->field[3], (const T *)e->field[4], (const T *)e->field[5],
                     (T *)e->field[0], (T *)e->field[1], (T *)e->field[2], (const T *)e->ce[0],
                     (const T *)e->ce[1], (const T *)e->ce[2], (const T *)e->ca[0],
                     (const T *)e->ca[1], (const T *)e->ca[2], prof_of<T>(e, 0),
                     psi_of<T>(e->psi, 6));
  HF_HIP(hipGetLastError());
  return timing_mark(e);
}

template <typename T>
static int sources_t(Engine *e, const hfdtd_source *src, int n) {
  for (int q = 0; q < n; ++q) {
    Src<T> s;
    if (!make_src<T>(e, src[q], s)) return fail(HFDTD_EINVAL, "invalid source descriptor");
    const size_t total = (size_t)(s.ci1 - s.ci0) * (s.cj1 - s.cj0) * (s.ck1 - s.ck0);
    if (total == 0) continue;
    hipLaunchKernelGGL(k_add_source<T>, dim3(ceil_div(total, 256)), dim3(256), 0, e->stream,
                       dims_of(e), s, (T *)e->field[s.comp], (const T *)e->ce[s.comp]);
    HF_HIP(hipGetLastError());
  }
  return HFDTD_OK;
}

template <typename T, int TJ>
static int fused_tj(Engine *e, const SrcList<T> &sl) {
  const int planes = e->tile_planes < 1 ? 1 : e->tile_planes;
  dim3 grid(ceil_div(e->nz, 64), ceil_div(e->ny, TJ), ceil_div(e->nx, planes));
  auto kern = e->lossless ? k_fused<T, TJ, false> : k_fused<T, TJ, true>;
  int rc = timing_mark(e);
  if (rc) return rc;
  hipLaunchKernelGGL(kern, grid, dim3(64, TJ), 0, e->stream, dims_of(e), planes, T(e->dtm_dl),
                     T(e->dtm), T(e->inv_dl), (const T *)e->field[0], (const T *)e->field[1],
                     (const T *)e->field[2], (const T *)e->field[3], (const T *)e->field[4],
                     (const T *)e->field[5], (T *)e->field_alt[0], (T *)e->field_alt[1],
                     (T *)e->field_alt[2], (T *)e->field_alt[3], (T *)e->field_alt[4],
                     (T *)e->field_alt[5], (const T *)e->ce[0], (const T *)e->ce[1],
                     (const T *)e->ce[2], (const T *)e->ca[0], (const T *)e->ca[1],
                     (const T *)e->ca[2], prof_of<T>(e, 1), prof_of<T>(e, 0),
                     psi_of<T>(e->psi, 0), psi_of<T>(e->psi_alt, 0), psi_of<T>(e->psi, 6),
                     psi_of<T>(e->psi_alt, 6), sl);
  HF_HIP(hipGetLastError());
  return timing_mark(e);
}

template <typename T>
static int fused_t(Engine *e, const hfdtd_source *src, int n) {
  if (n > kMaxSources) return fail(HFDTD_EINVAL, "at most 16 sources per fused step");
  SrcList<T> sl;
  sl.n = 0;
  for (int q = 0; q < n; ++q) {
    Src<T> s;
    if (!make_src<T>(e, src[q], s)) return fail(HFDTD_EINVAL, "invalid source descriptor");
    if (s.ci1 > s.ci0 && s.cj1 > s.cj0 && s.ck1 > s.ck0) sl.s[sl.n++] = s;
  }
  int rc = ensure_alt(e);
  if (rc) return rc;
  switch (e->tile_rows) {
    case 1: rc = fused_tj<T, 1>(e, sl); break;
    case 2: rc = fused_tj<T, 2>(e, sl); break;
    case 8: rc = fused_tj<T, 8>(e, sl); break;
    case 16: rc = fused_tj<T, 16>(e, sl); break;
    default: rc = fused_tj<T, 4>(e, sl); break;
  }
  if (rc) return rc;
  for (int f = 0; f < 6; ++f) std::swap(e->field[f], e->field_alt[f]);
  for (int s = 0; s < PSI_COUNT; ++s) std::swap(e->psi[s], e->psi_alt[s]);
  return HFDTD_OK;
}

int launch_update_H(Engine *e) {
  HF_HIP(hipSetDevice(e->device));
  return e->fp64 ? update_H_t<double>(e) : update_H_t<float>(e);
}

int launch_update_E(Engine *e) {
  HF_HIP(hipSetDevice(e->device));
  return e->fp64 ? update_E_t<double>(e) : update_E_t<float>(e);
}

int launch_sources(Engine *e, const hfdtd_source *src, int n) {
  HF_HIP(hipSetDevice(e->device));
  return e->fp64 ? sources_t<double>(e, src, n) : sources_t<float>(e, src, n);
}

int launch_fused_step(Engine *e, const hfdtd_source *src, int n) {
  HF_HIP(hipSetDevice(e->device));
  return e->fp64 ? fused_t<double>(e, src, n) : fused_t<float>(e, src, n);
}

}  // namespace hfdtd
