// Discrete adjoint of the Yee + CPML step, probe-DFT objectives and the
// checkpointed reverse sweep that gives dL/d(eps) for inverse design.
//
// The reference has no adjoint.  With the lossless forward step
// (kernels/yee_update.cl:11-238, step order engine.py:724-737)
//
//   H half:  psi_H' = b psi_H + c D E,   H' = H - dtm (ik D E + S psi_H')
//   E half:  psi_E' = b psi_E + c D'H,   E' = E + ce (ik D'H + S psi_E')
//
// (D forward difference zero on the last index, D' backward difference zero
// on index 0), the transposed step, taken in reverse order, is
//
//   E half:  w = s ce E~,   g = psi~_E + w,  H~ += D'^T (ik w + c g),  psi~_E = b g
//   H half:  w = -s dtm H~, g = psi~_H + w,  E~ += D^T  (ik w + c g),  psi~_H = b g
//
// with s = +-1 the sign of the term in the curl, (D^T u)_m = [m>=1] u_{m-1} -
// [m<=N-2] u_m and (D'^T u)_m = [m>=1] u_m - [m<=N-2] u_{m+1}.  Each half is
// one gather sweep shaped like the forward split sweeps:
//
//   pass 1 (one thread per H cell): H~ += D'^T(...) from E~ and psi~_E;
//   pass 2 (one thread per E cell): E~ += D^T(...)  from H~ and psi~_H,
//
// with threads over the flattened (j, k) plane and XCD-contiguous plane
// slabs as in the split sweeps.
//
// The stored E adjoint is scaled, ce E~ (with E~ the adjoint above), which
// drops ce from pass 1; in the interior the two passes are then the forward
// H and E stencils with -inv_dl for dtm_dl and -ce dtm_dl for ce inv_dl.
// Every cell uses the kappa-scaled form (ik = 1/dl and c = 0 off the PML),
// so a wave runs one code path whether or not it touches a face stripe.
//
// A thread may update psi~ only where no other thread of the same pass reads
// it, so each psi~ update is deferred to the pass that does not read it:
// pass 2 settles psi~_E of its own cell (from the E~ it is about to
// overwrite) and pass 1 of the next reverse step settles psi~_H of its own
// cell (from the H~ it is about to add to).  Between reverse steps E~, H~
// and psi~_E are final and psi~_H holds the value one step behind, which
// pass 1 completes as  psi~_H = b (psi~_H - s dtm H~).
//
// Gradient: current sources add -ce J, so for a lossless step
// dE^n/dce = (E^n - E^{n-1})/ce and dL/d(eps_edge) = -G/eps_edge with
// G = sum_n E~^n (E^n - E^{n-1}); G is summed in FP64 over a box of edges
// (the design box grown one cell toward lower indices, so every edge that
// touches a design cell is in it) from E snapshots recorded per step.

#include <algorithm>
#include <cmath>
#include <vector>

#include "fdtd_core.h"

namespace hfdtd {

namespace {

struct Box {
  int i0, j0, k0, ni, nj, nk;
  long long n;
};

__device__ __forceinline__ int coord_of(int a, int i, int j, int k) {
  return a == 0 ? i : (a == 1 ? j : k);
}

__device__ __forceinline__ bool in_stripe(const Geo &g, int a, int x) {
  const int N = a == 0 ? g.nx : (a == 1 ? g.ny : g.nz);
  return x < g.npml || x >= N - g.npml;
}

// element offset of cell (i,j,k) inside the face stripe normal to axis a
__device__ __forceinline__ int stripe_off(const Geo &g, int a, int i, int j, int k) {
  const int n = g.npml;
  if (a == 0) return (stripe_local(i, n, g.nx) * g.ny + j) * g.nz + k;
  if (a == 1) return (i * 2 * n + stripe_local(j, n, g.ny)) * g.nz + k;
  return (i * g.ny + j) * (2 * n) + stripe_local(k, n, g.nz);
}

__device__ __forceinline__ size_t cell(const Geo &g, int i, int j, int k) {
  return ((size_t)i * g.ny + j) * g.nz + k;
}

// All loads of the two passes are unconditional: a neighbour outside the
// grid reads the cell itself (and is masked out of the sum), a psi~ entry
// off its stripe reads a zero element of the slab's padding (its c is 0
// anyway), a profile index outside the axis reads the axis' edge.  The
// compiler can then issue every load of a cell at once instead of waiting at
// the join of each divergent branch.  psi~ stores stay predicated on the
// stripe.

// first element of the padding after the psi~ slab that nothing writes
__device__ __forceinline__ int zero_slot(const Geo &g) { return g.psi_off[PSI_COUNT] + 32; }

__device__ __forceinline__ int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

// profile entry q (0 = b, 1 = c, 2 = ik) of stagger st along axis a at index
// x clamped to the axis
template <typename T>
__device__ __forceinline__ T prof_c(const Geo &g, const T *__restrict__ prof, int st, int q, int a,
                                    int x) {
  const int N = a == 0 ? g.nx : (a == 1 ? g.ny : g.nz);
  return prof_at(g, prof, st, q, a, clampi(x, 0, N - 1));
}

// Per-axis data of one cell index: s = ik + c, c, b (stagger st) and the
// psi~ element offsets of two slots of this axis' stripe (zero_slot off it).
template <typename T>
struct AxisT {
  T s, c, b;
  bool in;
  int o0, o1;
};

template <typename T>
__device__ __forceinline__ AxisT<T> axis_t(const Geo &g, const T *__restrict__ prof, int st, int a,
                                           int slot0, int slot1, int i, int j, int k) {
  const int x = coord_of(a, i, j, k);
  AxisT<T> p;
  p.c = prof_c(g, prof, st, 1, a, x);
  p.s = prof_c(g, prof, st, 2, a, x) + p.c;
  p.b = prof_c(g, prof, st, 0, a, x);
  const int N = a == 0 ? g.nx : (a == 1 ? g.ny : g.nz);
  p.in = x >= 0 && x < N && in_stripe(g, a, x);
  const int so = p.in ? stripe_off(g, a, i, j, k) : 0;
  p.o0 = p.in ? g.psi_off[slot0] + so : zero_slot(g);
  p.o1 = p.in ? g.psi_off[slot1] + so : zero_slot(g);
  return p;
}

// psi~ <- b (psi~ + w) where the cell lies on the stripe
template <typename T>
__device__ __forceinline__ void settle_at(T *__restrict__ apsi, bool in, int o, T b, T old, T w) {
  if (in) apsi[o] = b * (old + w);
}

// Pass 1: settle psi~_H of the previous reverse step, then the transposed E
// half into H~.  adj = [ce E~x, ce E~y, ce E~z, H~x, H~y, H~z], psi~ slab in
// PsiSlot order.  For E term (component e, axis a, sign s) with w = s ce E~
// at cell c, u(c) = w (ik + c) + c psi~ and H~ += [m_a >= 1] u(m) -
// [m_a <= N-2] u(m + e_a).  In the interior (c = 0, ik = 1/dl) this is the
// forward H stencil with -inv_dl in place of dtm_dl.
template <typename T>
__global__ __launch_bounds__(kFlatBlock) void k_adj_pass1(Geo g, T dtm, T *__restrict__ adj,
                                                          const T *__restrict__ prof,
                                                          T *__restrict__ apsi, int bpp,
                                                          int xcd, int run) {
  const int L = xcd_runs(blockIdx.x, gridDim.x, xcd, run);
  const int i = L / bpp;
  const int q = (L - i * bpp) * kFlatBlock + threadIdx.x;
  if (q >= g.ny * g.nz) return;
  const int j = q / g.nz, k = q - j * g.nz;
  const size_t idx = cell(g, i, j, k);
  const size_t sxy = (size_t)g.ny * g.nz;
  const T *Ex = adj, *Ey = adj + g.vs, *Ez = adj + 2 * g.vs;
  const bool i1 = i + 1 < g.nx, j1 = j + 1 < g.ny, k1 = k + 1 < g.nz;
  const bool i0 = i >= 1, j0 = j >= 1, k0 = k >= 1;
  const T Z = T(0);
  // fields: own cell, +x/+y/+z neighbours (the cell itself past the edge)
  T hx = adj[3 * g.vs + idx], hy = adj[4 * g.vs + idx], hz = adj[5 * g.vs + idx];
  const T ex = Ex[idx], ey = Ey[idx], ez = Ez[idx];
  const T ex_j1 = Ex[j1 ? idx + g.nz : idx], ex_k1 = Ex[k1 ? idx + 1 : idx];
  const T ey_i1 = Ey[i1 ? idx + sxy : idx], ey_k1 = Ey[k1 ? idx + 1 : idx];
  const T ez_i1 = Ez[i1 ? idx + sxy : idx], ez_j1 = Ez[j1 ? idx + g.nz : idx];
  // H-side psi~ of this cell (settled below) and E-side psi~ at the cell and
  // its +1 neighbours
  const AxisT<T> hxa = axis_t(g, prof, 1, 0, PSI_HY_X, PSI_HZ_X, i, j, k);
  const AxisT<T> hya = axis_t(g, prof, 1, 1, PSI_HX_Y, PSI_HZ_Y, i, j, k);
  const AxisT<T> hza = axis_t(g, prof, 1, 2, PSI_HX_Z, PSI_HY_Z, i, j, k);
  const AxisT<T> x0 = axis_t(g, prof, 0, 0, PSI_EY_X, PSI_EZ_X, i, j, k);
  const AxisT<T> x1 = axis_t(g, prof, 0, 0, PSI_EY_X, PSI_EZ_X, i + 1, j, k);
  const AxisT<T> y0 = axis_t(g, prof, 0, 1, PSI_EX_Y, PSI_EZ_Y, i, j, k);
  const AxisT<T> y1 = axis_t(g, prof, 0, 1, PSI_EX_Y, PSI_EZ_Y, i, j + 1, k);
  const AxisT<T> z0 = axis_t(g, prof, 0, 2, PSI_EX_Z, PSI_EY_Z, i, j, k);
  const AxisT<T> z1 = axis_t(g, prof, 0, 2, PSI_EX_Z, PSI_EY_Z, i, j, k + 1);
  const T q_hyx = apsi[hxa.o0], q_hzx = apsi[hxa.o1];
  const T q_hxy = apsi[hya.o0], q_hzy = apsi[hya.o1];
  const T q_hxz = apsi[hza.o0], q_hyz = apsi[hza.o1];
  const T q_eyx0 = apsi[x0.o0], q_ezx0 = apsi[x0.o1], q_eyx1 = apsi[x1.o0], q_ezx1 = apsi[x1.o1];
  const T q_exy0 = apsi[y0.o0], q_ezy0 = apsi[y0.o1], q_exy1 = apsi[y1.o0], q_ezy1 = apsi[y1.o1];
  const T q_exz0 = apsi[z0.o0], q_eyz0 = apsi[z0.o1], q_exz1 = apsi[z1.o0], q_eyz1 = apsi[z1.o1];
  // settle: H terms Hx (+ y, - z), Hy (+ z, - x), Hz (+ x, - y); w = -s dtm H~
  settle_at(apsi, hxa.in, hxa.o0, hxa.b, q_hyx, dtm * hy);
  settle_at(apsi, hxa.in, hxa.o1, hxa.b, q_hzx, -dtm * hz);
  settle_at(apsi, hya.in, hya.o0, hya.b, q_hxy, -dtm * hx);
  settle_at(apsi, hya.in, hya.o1, hya.b, q_hzy, dtm * hz);
  settle_at(apsi, hza.in, hza.o0, hza.b, q_hxz, dtm * hx);
  settle_at(apsi, hza.in, hza.o1, hza.b, q_hyz, -dtm * hy);
  // u = w (ik + c) + c psi~ for each E term at the cell (0) and neighbour (1)
  // Hx: Ey (+ z), Ez (- y)
  hx += (k0 ? ey * z0.s + z0.c * q_eyz0 : Z) - (k1 ? ey_k1 * z1.s + z1.c * q_eyz1 : Z) +
        (j0 ? -ez * y0.s + y0.c * q_ezy0 : Z) - (j1 ? -ez_j1 * y1.s + y1.c * q_ezy1 : Z);
  // Hy: Ex (- z), Ez (+ x)
  hy += (k0 ? -ex * z0.s + z0.c * q_exz0 : Z) - (k1 ? -ex_k1 * z1.s + z1.c * q_exz1 : Z) +
        (i0 ? ez * x0.s + x0.c * q_ezx0 : Z) - (i1 ? ez_i1 * x1.s + x1.c * q_ezx1 : Z);
  // Hz: Ex (+ y), Ey (- x)
  hz += (j0 ? ex * y0.s + y0.c * q_exy0 : Z) - (j1 ? ex_j1 * y1.s + y1.c * q_exy1 : Z) +
        (i0 ? -ey * x0.s + x0.c * q_eyx0 : Z) - (i1 ? -ey_i1 * x1.s + x1.c * q_eyx1 : Z);
  adj[3 * g.vs + idx] = hx;
  adj[4 * g.vs + idx] = hy;
  adj[5 * g.vs + idx] = hz;
}

// Pass 2: settle psi~_E of this cell from the E~ about to be replaced, then
// the transposed H half into E~ (lossless: ca = 1).  For H term (component
// h, axis a, sign s) with w = -s dtm H~ at cell c, u(c) = w (ik + c) + c psi~
// and E~ += ce ([m_a >= 1] u(m - e_a) - [m_a <= N-2] u(m)).  In the interior
// this is the forward E stencil with -ce dtm_dl in place of ce inv_dl.
template <typename T>
__global__ __launch_bounds__(kFlatBlock) void k_adj_pass2(Geo g, T dtm, T *__restrict__ adj,
                                                          const T *__restrict__ coef,
                                                          const T *__restrict__ prof,
                                                          T *__restrict__ apsi, int bpp,
                                                          int xcd, int run) {
  const int L = xcd_runs(blockIdx.x, gridDim.x, xcd, run);
  const int i = L / bpp;
  const int q = (L - i * bpp) * kFlatBlock + threadIdx.x;
  if (q >= g.ny * g.nz) return;
  const int j = q / g.nz, k = q - j * g.nz;
  const size_t idx = cell(g, i, j, k);
  const size_t sxy = (size_t)g.ny * g.nz;
  const T *Hx = adj + 3 * g.vs, *Hy = adj + 4 * g.vs, *Hz = adj + 5 * g.vs;
  const bool i0 = i >= 1, j0 = j >= 1, k0 = k >= 1;
  const bool i1 = i + 1 < g.nx, j1 = j + 1 < g.ny, k1 = k + 1 < g.nz;
  const T Z = T(0);
  const T ex = adj[idx], ey = adj[g.vs + idx], ez = adj[2 * g.vs + idx];
  const T cx = coef[idx], cy = coef[g.vs + idx], cz = coef[2 * g.vs + idx];
  // w = -dtm H~ at the cell and its -x/-y/-z neighbours (the cell itself at
  // the low edge); the sign s of each term is applied below
  const T hx = -dtm * Hx[idx], hy = -dtm * Hy[idx], hz = -dtm * Hz[idx];
  const T hx_jm = -dtm * Hx[j0 ? idx - g.nz : idx], hx_km = -dtm * Hx[k0 ? idx - 1 : idx];
  const T hy_im = -dtm * Hy[i0 ? idx - sxy : idx], hy_km = -dtm * Hy[k0 ? idx - 1 : idx];
  const T hz_im = -dtm * Hz[i0 ? idx - sxy : idx], hz_jm = -dtm * Hz[j0 ? idx - g.nz : idx];
  const AxisT<T> exa = axis_t(g, prof, 0, 0, PSI_EY_X, PSI_EZ_X, i, j, k);
  const AxisT<T> eya = axis_t(g, prof, 0, 1, PSI_EX_Y, PSI_EZ_Y, i, j, k);
  const AxisT<T> eza = axis_t(g, prof, 0, 2, PSI_EX_Z, PSI_EY_Z, i, j, k);
  const AxisT<T> x0 = axis_t(g, prof, 1, 0, PSI_HY_X, PSI_HZ_X, i, j, k);
  const AxisT<T> xm = axis_t(g, prof, 1, 0, PSI_HY_X, PSI_HZ_X, i - 1, j, k);
  const AxisT<T> y0 = axis_t(g, prof, 1, 1, PSI_HX_Y, PSI_HZ_Y, i, j, k);
  const AxisT<T> ym = axis_t(g, prof, 1, 1, PSI_HX_Y, PSI_HZ_Y, i, j - 1, k);
  const AxisT<T> z0 = axis_t(g, prof, 1, 2, PSI_HX_Z, PSI_HY_Z, i, j, k);
  const AxisT<T> zm = axis_t(g, prof, 1, 2, PSI_HX_Z, PSI_HY_Z, i, j, k - 1);
  const T q_eyx = apsi[exa.o0], q_ezx = apsi[exa.o1];
  const T q_exy = apsi[eya.o0], q_ezy = apsi[eya.o1];
  const T q_exz = apsi[eza.o0], q_eyz = apsi[eza.o1];
  const T q_hyx0 = apsi[x0.o0], q_hzx0 = apsi[x0.o1], q_hyxm = apsi[xm.o0], q_hzxm = apsi[xm.o1];
  const T q_hxy0 = apsi[y0.o0], q_hzy0 = apsi[y0.o1], q_hxym = apsi[ym.o0], q_hzym = apsi[ym.o1];
  const T q_hxz0 = apsi[z0.o0], q_hyz0 = apsi[z0.o1], q_hxzm = apsi[zm.o0], q_hyzm = apsi[zm.o1];
  // settle: E terms Ex (+ y, - z), Ey (+ z, - x), Ez (+ x, - y); w = s ce E~
  settle_at(apsi, exa.in, exa.o0, exa.b, q_eyx, -ey);
  settle_at(apsi, exa.in, exa.o1, exa.b, q_ezx, ez);
  settle_at(apsi, eya.in, eya.o0, eya.b, q_exy, ex);
  settle_at(apsi, eya.in, eya.o1, eya.b, q_ezy, -ez);
  settle_at(apsi, eza.in, eza.o0, eza.b, q_exz, -ex);
  settle_at(apsi, eza.in, eza.o1, eza.b, q_eyz, ey);
  // Ex: Hy (+ z), Hz (- y)
  const T ax = (k0 ? hy_km * zm.s + zm.c * q_hyzm : Z) - (k1 ? hy * z0.s + z0.c * q_hyz0 : Z) +
               (j0 ? -hz_jm * ym.s + ym.c * q_hzym : Z) - (j1 ? -hz * y0.s + y0.c * q_hzy0 : Z);
  // Ey: Hx (- z), Hz (+ x)
  const T ay = (k0 ? -hx_km * zm.s + zm.c * q_hxzm : Z) - (k1 ? -hx * z0.s + z0.c * q_hxz0 : Z) +
               (i0 ? hz_im * xm.s + xm.c * q_hzxm : Z) - (i1 ? hz * x0.s + x0.c * q_hzx0 : Z);
  // Ez: Hx (+ y), Hy (- x)
  const T az = (j0 ? hx_jm * ym.s + ym.c * q_hxym : Z) - (j1 ? hx * y0.s + y0.c * q_hxy0 : Z) +
               (i0 ? -hy_im * xm.s + xm.c * q_hyxm : Z) - (i1 ? -hy * x0.s + x0.c * q_hyx0 : Z);
  adj[idx] = ex + cx * ax;
  adj[g.vs + idx] = ey + cy * ay;
  adj[2 * g.vs + idx] = ez + cz * az;
}

__device__ __forceinline__ size_t box_cell(const Geo &g, const Box &b, long long q) {
  const int bk = (int)(q % b.nk);
  const int bj = (int)((q / b.nk) % b.nj);
  const int bi = (int)(q / ((long long)b.nk * b.nj));
  return cell(g, b.i0 + bi, b.j0 + bj, b.k0 + bk);
}

// hist[c][q] = E_c at box cell q
template <typename T>
__global__ void k_adj_record(Geo g, Box b, const T *__restrict__ st, T *__restrict__ hist) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= b.n) return;
  const size_t idx = box_cell(g, b, q);
  hist[q] = st[idx];
  hist[b.n + q] = st[g.vs + idx];
  hist[2 * b.n + q] = st[2 * g.vs + idx];
}

// G_c += E~_c (E^n_c - E^{n-1}_c) over the box, in FP64 (adj holds ce E~)
template <typename T>
__global__ void k_adj_grad(Geo g, Box b, const T *__restrict__ adj, const T *__restrict__ coef,
                           const T *__restrict__ h1, const T *__restrict__ h0,
                           double *__restrict__ G) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= b.n) return;
  const size_t idx = box_cell(g, b, q);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const long long o = c * b.n + q;
    const double eb = (double)adj[c * g.vs + idx] / (double)coef[c * g.vs + idx];
    G[o] += eb * ((double)h1[o] - (double)h0[o]);
  }
}

// F[p][f] += E_p e^{j w_f t} dt (FP64), one thread per (probe, frequency)
template <typename T>
__global__ void k_adj_probe_dft(Geo g, int np, int nf, const int *__restrict__ comp,
                                const long long *__restrict__ where, const double *__restrict__ om,
                                double t, double dt, const T *__restrict__ st,
                                double *__restrict__ F) {
  const int th = blockIdx.x * blockDim.x + threadIdx.x;
  if (th >= np * nf) return;
  const int p = th / nf, f = th % nf;
  const double v = (double)st[comp[p] * g.vs + where[p]] * dt;
  double s, c;
  sincos(om[f] * t, &s, &c);
  F[2 * th] += v * c;
  F[2 * th + 1] += v * s;
}

// E~_p += sum_f Re(coef_pf e^{j w_f t}) dt, one thread per probe (probes are
// distinct cells, checked on the host); adj holds ce E~
template <typename T>
__global__ void k_adj_inject(Geo g, int np, int nf, const int *__restrict__ comp,
                             const long long *__restrict__ where, const double *__restrict__ om,
                             const double *__restrict__ coef, double t, double dt,
                             const T *__restrict__ ce, T *__restrict__ adj) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  double acc = 0.0;
  for (int f = 0; f < nf; ++f) {
    double s, c;
    sincos(om[f] * t, &s, &c);
    acc += coef[2 * (p * nf + f)] * c - coef[2 * (p * nf + f) + 1] * s;
  }
  const size_t at = comp[p] * g.vs + where[p];
  adj[at] += T((double)ce[at] * (acc * dt));
}

template <typename T>
__global__ void k_adj_fill(T *p, size_t n, T v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

}  // namespace

}  // namespace hfdtd

using namespace hfdtd;

struct hfdtd_adjoint {
  Engine *e = nullptr;
  Box box{};
  int design[6] = {};
  int hist_slots = 0;
  int n_ckpt = 0;
  void *adj = nullptr;   // E~ H~ slab (engine layout)
  void *apsi = nullptr;  // psi~ slab (engine layout)
  void *hist = nullptr;  // hist_slots x 3 x box.n
  double *G = nullptr;   // 3 x box.n
  void *ckpt = nullptr;  // n_ckpt x (state slab + psi slab)
  size_t ckpt_state = 0, ckpt_psi = 0;
  int np = 0, nf = 0;
  int *pcomp = nullptr;
  long long *pidx = nullptr;
  double *om = nullptr;
  double *F = nullptr;     // np x nf x {re, im}
  double *coef = nullptr;  // np x nf x {re, im}
  bool coef_set = false;
  uint64_t bytes = 0;
  uint64_t probe_bytes = 0;
};

namespace {

int amalloc(hfdtd_adjoint *a, void **p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) return HFDTD_OK;
  hipError_t err = hipMalloc(p, bytes);
  if (err != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    char buf[160];
    std::snprintf(buf, sizeof buf, "adjoint: device allocation of %.3f GB failed (%s)",
                  bytes / 1073741824.0, hipGetErrorString(err));
    return fail(HFDTD_ENOMEM, buf);
  }
  a->bytes += bytes;
  return HFDTD_OK;
}

void release(hfdtd_adjoint *a) {
  if (a->e) {
    (void)hipSetDevice(a->e->device);
    (void)hipStreamSynchronize(a->e->stream);
  }
  void **ptrs[] = {&a->adj, &a->apsi, &a->hist, (void **)&a->G, &a->ckpt, (void **)&a->pcomp,
                   (void **)&a->pidx, (void **)&a->om, (void **)&a->F, (void **)&a->coef};
  for (void **p : ptrs) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
}

template <typename T>
int fill_t(Engine *e, void *p, size_t n, T v) {
  if (n == 0) return HFDTD_OK;
  unsigned blocks = ceil_div(n, 256);
  if (blocks > 65536u * 4u) blocks = 65536u * 4u;
  hipLaunchKernelGGL(k_adj_fill<T>, dim3(blocks), dim3(256), 0, e->stream, (T *)p, n, v);
  HF_HIP(hipGetLastError());
  return HFDTD_OK;
}

size_t psi_elems(const Engine *e) { return e->psi_off[PSI_COUNT] + 64; }

template <typename T>
int record_t(hfdtd_adjoint *a, int slot) {
  Engine *e = a->e;
  HF_SETTLE(e);
  T *dst = (T *)a->hist + (size_t)slot * 3 * a->box.n;
  hipLaunchKernelGGL(k_adj_record<T>, dim3(ceil_div(a->box.n, 256)), dim3(256), 0, e->stream,
                     geo_of(e), a->box, (const T *)e->state[e->cur], dst);
  HF_HIP(hipGetLastError());
  return HFDTD_OK;
}

int record(hfdtd_adjoint *a, int slot) {
  return a->e->fp64 ? record_t<double>(a, slot) : record_t<float>(a, slot);
}

template <typename T>
int probe_t(hfdtd_adjoint *a, long long n) {
  Engine *e = a->e;
  HF_SETTLE(e);
  const int th = a->np * a->nf;
  hipLaunchKernelGGL(k_adj_probe_dft<T>, dim3(ceil_div(th, 64)), dim3(64), 0, e->stream,
                     geo_of(e), a->np, a->nf, a->pcomp, a->pidx, a->om, (double)n * e->dt, e->dt,
                     (const T *)e->state[e->cur], a->F);
  HF_HIP(hipGetLastError());
  return HFDTD_OK;
}

template <typename T>
int inject_t(hfdtd_adjoint *a, long long n) {
  Engine *e = a->e;
  HF_SETTLE(e);
  hipLaunchKernelGGL(k_adj_inject<T>, dim3(ceil_div(a->np, 64)), dim3(64), 0, e->stream,
                     geo_of(e), a->np, a->nf, a->pcomp, a->pidx, a->om, a->coef,
                     (double)n * e->dt, e->dt, (const T *)e->coef, (T *)a->adj);
  HF_HIP(hipGetLastError());
  return HFDTD_OK;
}

template <typename T>
int reverse_step_t(hfdtd_adjoint *a, long long n, int slot) {
  Engine *e = a->e;
  HF_SETTLE(e);
  const Geo g = geo_of(e);
  const size_t hs = (size_t)3 * a->box.n;
  const T *h1 = (const T *)a->hist + (size_t)slot * hs;
  const T *h0 = (const T *)a->hist + (size_t)(slot - 1) * hs;
  hipLaunchKernelGGL(k_adj_grad<T>, dim3(ceil_div(a->box.n, 256)), dim3(256), 0, e->stream, g,
                     a->box, (const T *)a->adj, (const T *)e->coef, h1, h0, a->G);
  HF_HIP(hipGetLastError());
  const unsigned bpp = ceil_div((size_t)e->ny * e->nz, kFlatBlock);
  const dim3 grid(bpp * (unsigned)e->nx);
  hipLaunchKernelGGL(k_adj_pass1<T>, grid, dim3(kFlatBlock), 0, e->stream, g, T(e->dtm),
                     (T *)a->adj, (const T *)e->prof, (T *)a->apsi, (int)bpp, e->xcd, e->split_run);
  HF_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_adj_pass2<T>, grid, dim3(kFlatBlock), 0, e->stream, g, T(e->dtm),
                     (T *)a->adj, (const T *)e->coef, (const T *)e->prof, (T *)a->apsi,
                     (int)bpp, e->xcd, e->split_run);
  HF_HIP(hipGetLastError());
  if (n - 1 >= 1) return inject_t<T>(a, n - 1);
  return HFDTD_OK;
}

}  // namespace

extern "C" {

int hfdtd_adjoint_create(hfdtd_engine *h, const int *design_box, int hist_slots, int n_ckpt,
                         hfdtd_adjoint **out) {
  if (!h || !design_box || !out) return fail(HFDTD_EINVAL, "NULL argument");
  *out = nullptr;
  Engine *e = &h->s;
  HF_SETTLE(e);
  const int ext[3] = {e->nx, e->ny, e->nz};
  for (int a = 0; a < 3; ++a) {
    const int lo = design_box[2 * a], hi = design_box[2 * a + 1];
    if (lo < 0 || hi > ext[a] || hi <= lo)
      return fail(HFDTD_EINVAL, "design box must be a non-empty box inside the grid");
  }
  if (hist_slots < 2) return fail(HFDTD_EINVAL, "hist_slots must be >= 2");
  if (n_ckpt < 0) return fail(HFDTD_EINVAL, "n_ckpt must be >= 0");
  if (!e->lossless)
    return fail(HFDTD_EUNSUPPORTED, "the adjoint gradient needs a lossless grid (no sigma)");
  HF_HIP(hipSetDevice(e->device));
  hfdtd_adjoint *a = new hfdtd_adjoint();
  a->e = e;
  for (int q = 0; q < 6; ++q) a->design[q] = design_box[q];
  a->box.i0 = std::max(design_box[0] - 1, 0);
  a->box.j0 = std::max(design_box[2] - 1, 0);
  a->box.k0 = std::max(design_box[4] - 1, 0);
  a->box.ni = design_box[1] - a->box.i0;
  a->box.nj = design_box[3] - a->box.j0;
  a->box.nk = design_box[5] - a->box.k0;
  a->box.n = (long long)a->box.ni * a->box.nj * a->box.nk;
  a->hist_slots = hist_slots;
  a->n_ckpt = n_ckpt;
  a->ckpt_state = 6 * e->vstride * e->elem;
  a->ckpt_psi = psi_elems(e) * e->elem;
  int rc;
  auto bail = [&](int code) {
    release(a);
    delete a;
    return code;
  };
  if ((rc = amalloc(a, &a->adj, 6 * e->vstride * e->elem))) return bail(rc);
  if ((rc = amalloc(a, &a->apsi, psi_elems(e) * e->elem))) return bail(rc);
  if ((rc = amalloc(a, &a->hist, (size_t)hist_slots * 3 * a->box.n * e->elem))) return bail(rc);
  if ((rc = amalloc(a, (void **)&a->G, (size_t)3 * a->box.n * sizeof(double)))) return bail(rc);
  if ((rc = amalloc(a, &a->ckpt, (size_t)n_ckpt * (a->ckpt_state + a->ckpt_psi)))) return bail(rc);
  *out = a;
  return HFDTD_OK;
}

int hfdtd_adjoint_destroy(hfdtd_adjoint *a) {
  if (!a) return HFDTD_OK;
  release(a);
  delete a;
  return HFDTD_OK;
}

int hfdtd_adjoint_memory_bytes(const hfdtd_adjoint *a, uint64_t *bytes) {
  if (!a || !bytes) return fail(HFDTD_EINVAL, "NULL argument");
  *bytes = a->bytes;
  return HFDTD_OK;
}

int hfdtd_adjoint_grad_box(const hfdtd_adjoint *a, int *box) {
  if (!a || !box) return fail(HFDTD_EINVAL, "NULL argument");
  box[0] = a->box.i0;
  box[1] = a->box.i0 + a->box.ni;
  box[2] = a->box.j0;
  box[3] = a->box.j0 + a->box.nj;
  box[4] = a->box.k0;
  box[5] = a->box.k0 + a->box.nk;
  return HFDTD_OK;
}

int hfdtd_adjoint_set_probes(hfdtd_adjoint *a, int n, const int *ijkc, int nfreq,
                             const double *omega) {
  if (!a || n < 1 || !ijkc || nfreq < 1 || !omega) return fail(HFDTD_EINVAL, "bad probe list");
  Engine *e = a->e;
  HF_SETTLE(e);
  std::vector<int> comp(n);
  std::vector<long long> where(n);
  for (int p = 0; p < n; ++p) {
    const int i = ijkc[4 * p], j = ijkc[4 * p + 1], k = ijkc[4 * p + 2], c = ijkc[4 * p + 3];
    if (i < 0 || i >= e->nx || j < 0 || j >= e->ny || k < 0 || k >= e->nz || c < 0 || c > 2)
      return fail(HFDTD_EINVAL, "probe cell outside the grid or component not Ex/Ey/Ez");
    comp[p] = c;
    where[p] = ((long long)i * e->ny + j) * e->nz + k;
    for (int q = 0; q < p; ++q)
      if (comp[q] == c && where[q] == where[p]) return fail(HFDTD_EINVAL, "duplicate probe");
  }
  HF_HIP(hipSetDevice(e->device));
  HF_HIP(hipStreamSynchronize(e->stream));
  void **ptrs[] = {(void **)&a->pcomp, (void **)&a->pidx, (void **)&a->om, (void **)&a->F,
                   (void **)&a->coef};
  for (void **p : ptrs) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  a->bytes -= a->probe_bytes;
  a->np = a->nf = 0;
  const uint64_t before = a->bytes;
  int rc;
  const size_t nfc = (size_t)n * nfreq * 2 * sizeof(double);
  if ((rc = amalloc(a, (void **)&a->pcomp, n * sizeof(int)))) return rc;
  if ((rc = amalloc(a, (void **)&a->pidx, n * sizeof(long long)))) return rc;
  if ((rc = amalloc(a, (void **)&a->om, nfreq * sizeof(double)))) return rc;
  if ((rc = amalloc(a, (void **)&a->F, nfc))) return rc;
  if ((rc = amalloc(a, (void **)&a->coef, nfc))) return rc;
  HF_HIP(hipMemcpy(a->pcomp, comp.data(), n * sizeof(int), hipMemcpyHostToDevice));
  HF_HIP(hipMemcpy(a->pidx, where.data(), n * sizeof(long long), hipMemcpyHostToDevice));
  HF_HIP(hipMemcpy(a->om, omega, nfreq * sizeof(double), hipMemcpyHostToDevice));
  HF_HIP(hipMemset(a->F, 0, nfc));
  HF_HIP(hipMemset(a->coef, 0, nfc));
  a->probe_bytes = a->bytes - before;
  a->np = n;
  a->nf = nfreq;
  a->coef_set = false;
  return HFDTD_OK;
}

int hfdtd_adjoint_probe_reset(hfdtd_adjoint *a) {
  if (!a || !a->np) return fail(HFDTD_EINVAL, "no probes set");
  HF_HIP(hipSetDevice(a->e->device));
  return fill_t<double>(a->e, a->F, (size_t)a->np * a->nf * 2, 0.0);
}

int hfdtd_adjoint_probe_read(hfdtd_adjoint *a, double *out) {
  if (!a || !out || !a->np) return fail(HFDTD_EINVAL, "no probes set");
  Engine *e = a->e;
  HF_SETTLE(e);
  HF_HIP(hipSetDevice(e->device));
  HF_HIP(hipMemcpyAsync(out, a->F, (size_t)a->np * a->nf * 2 * sizeof(double),
                        hipMemcpyDeviceToHost, e->stream));
  HF_HIP(hipStreamSynchronize(e->stream));
  return HFDTD_OK;
}

int hfdtd_adjoint_set_objective_grad(hfdtd_adjoint *a, const double *coef) {
  if (!a || !coef || !a->np) return fail(HFDTD_EINVAL, "no probes set");
  Engine *e = a->e;
  HF_SETTLE(e);
  HF_HIP(hipSetDevice(e->device));
  HF_HIP(hipMemcpyAsync(a->coef, coef, (size_t)a->np * a->nf * 2 * sizeof(double),
                        hipMemcpyHostToDevice, e->stream));
  HF_HIP(hipStreamSynchronize(e->stream));
  a->coef_set = true;
  return HFDTD_OK;
}

int hfdtd_adjoint_save(hfdtd_adjoint *a, int c) {
  if (!a || c < 0 || c >= a->n_ckpt) return fail(HFDTD_EINVAL, "checkpoint index out of range");
  Engine *e = a->e;
  HF_SETTLE(e);
  HF_HIP(hipSetDevice(e->device));
  char *dst = (char *)a->ckpt + (size_t)c * (a->ckpt_state + a->ckpt_psi);
  HF_HIP(hipMemcpyAsync(dst, e->state[e->cur], a->ckpt_state, hipMemcpyDeviceToDevice,
                        e->stream));
  HF_HIP(hipMemcpyAsync(dst + a->ckpt_state, e->psi[e->cur], a->ckpt_psi,
                        hipMemcpyDeviceToDevice, e->stream));
  return HFDTD_OK;
}

int hfdtd_adjoint_restore(hfdtd_adjoint *a, int c) {
  if (!a || c < 0 || c >= a->n_ckpt) return fail(HFDTD_EINVAL, "checkpoint index out of range");
  Engine *e = a->e;
  HF_SETTLE(e);
  e->touch();
  HF_HIP(hipSetDevice(e->device));
  const char *src = (const char *)a->ckpt + (size_t)c * (a->ckpt_state + a->ckpt_psi);
  HF_HIP(hipMemcpyAsync(e->state[e->cur], src, a->ckpt_state, hipMemcpyDeviceToDevice,
                        e->stream));
  HF_HIP(hipMemcpyAsync(e->psi[e->cur], src + a->ckpt_state, a->ckpt_psi,
                        hipMemcpyDeviceToDevice, e->stream));
  return HFDTD_OK;
}

int hfdtd_adjoint_record(hfdtd_adjoint *a, int slot) {
  if (!a || slot < 0 || slot >= a->hist_slots) return fail(HFDTD_EINVAL, "slot out of range");
  HF_HIP(hipSetDevice(a->e->device));
  return record(a, slot);
}

int hfdtd_adjoint_forward(hfdtd_adjoint *a, const hfdtd_source *src, int nsrc,
                          const double *amp, int n_steps, int64_t step0, int slot0, int probe,
                          int path) {
  if (!a || n_steps < 0 || nsrc < 0 || nsrc > kMaxSources || (nsrc > 0 && (!src || !amp)))
    return fail(HFDTD_EINVAL, "bad forward arguments (at most 16 sources)");
  if (path != HFDTD_PATH_FUSED && path != HFDTD_PATH_SPLIT && path != HFDTD_PATH_MARCH)
    return fail(HFDTD_EINVAL, "unknown step path");
  if (slot0 >= 0 && slot0 + n_steps >= a->hist_slots)
    return fail(HFDTD_EINVAL, "recorded steps exceed the history slots");
  if (probe && !a->np) return fail(HFDTD_EINVAL, "no probes set");
  Engine *e = a->e;
  HF_SETTLE(e);
  HF_HIP(hipSetDevice(e->device));
  std::vector<hfdtd_source> s(src, src + nsrc);
  const bool one = path != HFDTD_PATH_SPLIT && e->fused_fits();
  for (int q = 0; q < n_steps; ++q) {
    for (int r = 0; r < nsrc; ++r) s[r].amp = amp[(size_t)q * nsrc + r];
    int rc;
    if (one) {
      rc = path == HFDTD_PATH_MARCH ? launch_march_step(e, s.data(), nsrc)
                                    : launch_fused_step(e, s.data(), nsrc);
    } else {
      rc = launch_update_H(e);
      if (!rc && nsrc) rc = launch_sources(e, s.data(), nsrc);
      if (!rc) rc = launch_update_E(e);
    }
    if (rc) return rc;
    const long long n = step0 + q + 1;
    if (probe && (rc = e->fp64 ? probe_t<double>(a, n) : probe_t<float>(a, n))) return rc;
    if (slot0 >= 0 && (rc = record(a, slot0 + q + 1))) return rc;
  }
  return HFDTD_OK;
}

int hfdtd_adjoint_begin(hfdtd_adjoint *a, int64_t n_last) {
  if (!a || !a->np || !a->coef_set)
    return fail(HFDTD_EINVAL, "set the probes and the objective gradient first");
  Engine *e = a->e;
  HF_SETTLE(e);
  e->touch();
  if (!e->lossless)
    return fail(HFDTD_EUNSUPPORTED, "the adjoint gradient needs a lossless grid (no sigma)");
  HF_HIP(hipSetDevice(e->device));
  int rc;
  if (e->fp64) {
    if ((rc = fill_t<double>(e, a->adj, 6 * e->vstride, 0.0))) return rc;
    if ((rc = fill_t<double>(e, a->apsi, psi_elems(e), 0.0))) return rc;
  } else {
    if ((rc = fill_t<float>(e, a->adj, 6 * e->vstride, 0.f))) return rc;
    if ((rc = fill_t<float>(e, a->apsi, psi_elems(e), 0.f))) return rc;
  }
  if ((rc = fill_t<double>(e, a->G, (size_t)3 * a->box.n, 0.0))) return rc;
  if (n_last >= 1) return e->fp64 ? inject_t<double>(a, n_last) : inject_t<float>(a, n_last);
  return HFDTD_OK;
}

int hfdtd_adjoint_reverse(hfdtd_adjoint *a, int64_t n_hi, int64_t n_lo, int slot_hi) {
  if (!a || n_lo < 0 || n_hi < n_lo) return fail(HFDTD_EINVAL, "need 0 <= n_lo <= n_hi");
  if (slot_hi >= a->hist_slots || slot_hi - (n_hi - n_lo) < 0)
    return fail(HFDTD_EINVAL, "history slots do not cover the reverse range");
  Engine *e = a->e;
  HF_SETTLE(e);
  e->touch();
  HF_HIP(hipSetDevice(e->device));
  for (long long n = n_hi; n > n_lo; --n) {
    const int slot = (int)(slot_hi - (n_hi - n));
    const int rc = e->fp64 ? reverse_step_t<double>(a, n, slot) : reverse_step_t<float>(a, n, slot);
    if (rc) return rc;
  }
  return HFDTD_OK;
}

int hfdtd_adjoint_read_grad(hfdtd_adjoint *a, double *out) {
  if (!a || !out) return fail(HFDTD_EINVAL, "NULL argument");
  Engine *e = a->e;
  HF_SETTLE(e);
  HF_HIP(hipSetDevice(e->device));
  HF_HIP(hipMemcpyAsync(out, a->G, (size_t)3 * a->box.n * sizeof(double), hipMemcpyDeviceToHost,
                        e->stream));
  HF_HIP(hipStreamSynchronize(e->stream));
  return HFDTD_OK;
}

int hfdtd_adjoint_read_volume(hfdtd_adjoint *a, int which, void *host) {
  if (!a || !host || which < 0 || which > 5) return fail(HFDTD_EINVAL, "bad volume id");
  Engine *e = a->e;
  HF_SETTLE(e);
  HF_HIP(hipSetDevice(e->device));
  HF_HIP(hipMemcpyAsync(host, (const char *)a->adj + (size_t)which * e->vstride * e->elem,
                        e->ncell * e->elem, hipMemcpyDeviceToHost, e->stream));
  HF_HIP(hipStreamSynchronize(e->stream));
  return HFDTD_OK;
}

}  // extern "C"
