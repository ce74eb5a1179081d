// Near-to-far (Huygens box) monitor: per-step multi-frequency DFT of the
// co-located tangential fields on the six faces, snapshot / relative change,
// and the surface-integral far-field transform.
//
// Replaces OpenCLNear2FarMonitor (monitors.py:383-804) and the kernels
// accumulate_dft_faces_fused, dft_rel_change_partial, farfield_accumulate_nl
// and farfield_nl_to_eh (kernels/dft_farfield.cl:98-635).  Differences by
// design: one launch updates every frequency of the monitor (fields are read
// once per step, not once per frequency), and the N/L surface sum is a
// deterministic two-stage reduction in FP64 instead of float CAS atomics.

#include <cmath>
#include <mutex>
#include <vector>

#include "common.h"
#include "n2f_gather.h"

struct hfdtd_n2f {
  hfdtd::Engine *e = nullptr;
  int box[6] = {};
  int nfreq = 0;
  int acc_mode = 0;
  int n_face = 0;
  int off[6] = {};
  size_t acc_elem = 4;           // bytes per real component of an accumulator
  void *acc = nullptr;           // [6 comps][nfreq][n_face] x {re, im}
  void *cmp = nullptr;           // Kahan compensation, same layout (float)
  void *snap = nullptr;          // snapshot of acc
  double *scratch = nullptr;     // reduction partials
  size_t scratch_len = 0;
  // K-step buffering (0 = per-step accumulation): co-located tangential
  // samples of K steps, rows [E: 2 per face sample | H: 2 per face sample]
  int K = 0, kfill = 0;
  long long pe = 0, ppad = 0;    // start of the H rows, padded row count
  float *buf = nullptr;          // [K][ppad]
  float *we = nullptr, *wh = nullptr;  // [K][32] phase rows (E, H)
};

namespace hfdtd {

constexpr int kMaxFreqPerLaunch = 16;


template <typename A>
struct PhaseArg {
  A v[kMaxFreqPerLaunch][4];  // e.re e.im h.re h.im
};





// acc += val * phase, in A, optionally Kahan-compensated (cmp holds the
// running compensation, sum_true = acc - cmp).
template <typename A, bool KAHAN>
__device__ __forceinline__ void dft_add(A *acc, A *cmp, size_t o, A val, A pr, A pi) {
  if (!KAHAN) {
    acc[o] += val * pr;
    acc[o + 1] += val * pi;
    return;
  }
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const A term = val * (q == 0 ? pr : pi);
    const A s = acc[o + q];
    const A y = term - cmp[o + q];
    const A t = s + y;
    cmp[o + q] = (t - s) - y;
    acc[o + q] = t;
  }
}

template <typename T, typename A, bool KAHAN>
__global__ __launch_bounds__(256) void k_n2f_accumulate(Dims3 g, Box b, int nfreq, int f0,
                                                        int nf, const T *__restrict__ Ex,
                                                        const T *__restrict__ Ey,
                                                        const T *__restrict__ Ez,
                                                        const T *__restrict__ Hx,
                                                        const T *__restrict__ Hy,
                                                        const T *__restrict__ Hz, A *acc,
                                                        A *cmp, PhaseArg<A> ph) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= b.n_face) return;
  int i, j, k;
  const int f = decode_face(b, s, i, j, k);
  T v[4];
  int c[4];
  colocate_impl<T>(g, f, i, j, k, Ex, Ey, Ez, Hx, Hy, Hz, v, c);
  const size_t nfc = (size_t)b.n_face;
  for (int q = 0; q < nf; ++q) {
    const int fr = f0 + q;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const bool is_e = c[m] < 3;
      const size_t o = (((size_t)c[m] * nfreq + fr) * nfc + s) * 2;
      dft_add<A, KAHAN>(acc, cmp, o, A(v[m]), ph.v[q][is_e ? 0 : 2], ph.v[q][is_e ? 1 : 3]);
    }
  }
}

// ── K-step buffering ───────────────────────────────────────────────────────
// Over K steps the face sums are the contraction
//   acc[c][f][s] += sum_k v_k(c, s) * phase_k(f)   (E phases or H phases),
// done once per K steps on the matrix cores (as in dft.hip): k_n2f_gather
// writes the co-located tangential samples of one step into row k of the
// buffer, k_n2f_contract folds the block into the sums.

__global__ __launch_bounds__(256) void k_n2f_gather(const GatherTail t) {
  gather_block(t, blockIdx.x, threadIdx.x);
}

typedef float n2f_f32x16 __attribute__((ext_vector_type(16)));

// One wavefront per 32 buffer rows (all E rows or all H rows: pe is a
// multiple of 32); v_mfma_f32_32x32x2f32 as in k_dft_contract, output
// D[i][j] in register r of lane l with i = (r/4)*8 + (l/32)*4 + r%4,
// j = l%32 = 2 f + {re, im}.  W rows at or beyond kvalid count as zero.
// The block goes through LDS to fold D into the sums: row i of a wavefront's
// tile is sample s0 + i/2, tangential component i%2, so for one component
// and frequency its 16 samples' (re, im) pairs are 32 consecutive sums, and
// each lane then reads and writes whole runs of them.
template <typename A, bool KAHAN>
__global__ __launch_bounds__(256) void k_n2f_contract(Box b, const float *__restrict__ buf,
                                                      long long ppad, long long pe,
                                                      const float *__restrict__ we,
                                                      const float *__restrict__ wh, int kvalid,
                                                      int nfreq, A *__restrict__ acc,
                                                      A *__restrict__ cmp) {
  __shared__ float tile[4][32][33];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long long p0 = ((long long)blockIdx.x * 4 + wave) * 32;
  const bool live = p0 < ppad;  // wave-uniform
  const bool hrow = p0 >= pe;
  const int half = lane >> 5, col = lane & 31;
  if (live) {
    const float *W = hrow ? wh : we;
    n2f_f32x16 d = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int pairs = (kvalid + 1) / 2;
#pragma unroll 4
    for (int t = 0; t < pairs; ++t) {
      const int k = 2 * t + half;
      const float a = buf[(long long)k * ppad + p0 + col];
      const float w = k < kvalid ? W[k * 32 + col] : 0.0f;
      d = __builtin_amdgcn_mfma_f32_32x32x2f32(a, w, d, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) tile[wave][(r / 4) * 8 + half * 4 + (r % 4)][col] = d[r];
  }
  __syncthreads();
  if (!live) return;
  const long long row0 = p0 - (hrow ? pe : 0);  // even: the tile's first row
  const int s0 = (int)(row0 >> 1);
  const size_t nfc = (size_t)b.n_face;
  for (int v = 0; v < nfreq; ++v) {
    // lanes 0-31: (component 0, frequency v); lanes 32-63: (component 1, v)
    const int tg = half, q = lane & 1, sl = (lane >> 1) & 15;
    const int s = s0 + sl;
    if (s >= b.n_face) continue;
    const int f = s < b.off[2] ? (s < b.off[1] ? 0 : 1)
                               : s < b.off[4] ? (s < b.off[3] ? 2 : 3) : (s < b.off[5] ? 4 : 5);
    // tangential components of the face: E (t=0,1) then H (t=0,1)
    const int ce = f <= 1 ? 1 + tg : (f <= 3 ? 2 * tg : tg);
    const int comp = hrow ? ce + 3 : ce;
    const size_t o = (((size_t)comp * nfreq + v) * nfc + s) * 2 + q;
    const A term = (A)tile[wave][2 * sl + tg][2 * v + q];
    if (!KAHAN) {
      acc[o] += term;
    } else {
      const A sv = acc[o];
      const A y = term - cmp[o];
      const A tt = sv + y;
      cmp[o] = (tt - sv) - y;
      acc[o] = tt;
    }
  }
}

// Per-block partial sums of |cur - prev|^2 and |cur|^2 over the 6 buffers of
// one frequency (dft_farfield.cl:237-290), in FP64.
template <typename A>
__global__ __launch_bounds__(256) void k_n2f_relchange(int n_face, int nfreq, int fr,
                                                      const A *__restrict__ cur,
                                                      const A *__restrict__ prev,
                                                      double *__restrict__ part) {
  __shared__ double sn[256], sd[256];
  const int s = blockIdx.x * 256 + threadIdx.x;
  double num = 0.0, den = 0.0;
  if (s < n_face) {
    for (int c = 0; c < 6; ++c) {
      const size_t o = (((size_t)c * nfreq + fr) * n_face + s) * 2;
      const double cr = cur[o], ci = cur[o + 1];
      const double dr = cr - (double)prev[o], di = ci - (double)prev[o + 1];
      num += dr * dr + di * di;
      den += cr * cr + ci * ci;
    }
  }
  sn[threadIdx.x] = num;
  sd[threadIdx.x] = den;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      sn[threadIdx.x] += sn[threadIdx.x + st];
      sd[threadIdx.x] += sd[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sn[0];
    part[2 * blockIdx.x + 1] = sd[0];
  }
}

// Stage 1 of the far-field: block (bx, obs) sums the N, L contributions of
// samples [bx*chunk, (bx+1)*chunk) for observation obs into 12 doubles.
// Per-sample terms follow face_sample_NL (dft_farfield.cl:372-474): phase
// k r.r' in G, e^{j phase} and products in A.
template <typename G, typename A>
__global__ __launch_bounds__(256) void k_n2f_nl(Box b, int nfreq, int fr, int chunk,
                                               const A *__restrict__ acc,
                                               const double *__restrict__ obs, G k_wave, G dl,
                                               double *__restrict__ part) {
  const int ob = blockIdx.y;
  const int nblk = gridDim.x;
  const G ox = G(obs[3 * ob]), oy = G(obs[3 * ob + 1]), oz = G(obs[3 * ob + 2]);
  G r = sqrt(ox * ox + oy * oy + oz * oz);
  if (r < G(1.0e-30)) r = G(1.0e-30);
  const G rx = ox / r, ry = oy / r, rz = oz / r;
  const G dA = dl * dl;
  double s6[12];
#pragma unroll
  for (int q = 0; q < 12; ++q) s6[q] = 0.0;
  const size_t nfc = (size_t)b.n_face;
  auto ld = [&](int comp, int s, A &re, A &im) {
    const size_t o = (((size_t)comp * nfreq + fr) * nfc + s) * 2;
    re = acc[o];
    im = acc[o + 1];
  };
  const int s_end = min(b.n_face, (int)((blockIdx.x + 1) * (size_t)chunk));
  for (int s = blockIdx.x * chunk + threadIdx.x; s < s_end; s += blockDim.x) {
    int i, j, k;
    const int f = decode_face(b, s, i, j, k);
    G xp, yp, zp;
    if (f <= 1) {
      xp = G(i) * dl;
      yp = (G(j) + G(0.5)) * dl;
      zp = (G(k) + G(0.5)) * dl;
    } else if (f <= 3) {
      xp = (G(i) + G(0.5)) * dl;
      yp = G(j) * dl;
      zp = (G(k) + G(0.5)) * dl;
    } else {
      xp = (G(i) + G(0.5)) * dl;
      yp = (G(j) + G(0.5)) * dl;
      zp = G(k) * dl;
    }
    const G phase = k_wave * (rx * xp + ry * yp + rz * zp);
    const A pc = cos(A(phase)), ps = sin(A(phase));
    G wj, wk;  // trapezoid weights (dft_farfield.cl:431-442)
    if (f <= 1) {
      wj = (j == b.iy0 || j == b.iy1) ? G(0.5) : G(1);
      wk = (k == b.iz0 || k == b.iz1) ? G(0.5) : G(1);
    } else if (f <= 3) {
      wj = (i == b.ix0 || i == b.ix1) ? G(0.5) : G(1);
      wk = (k == b.iz0 || k == b.iz1) ? G(0.5) : G(1);
    } else {
      wj = (i == b.ix0 || i == b.ix1) ? G(0.5) : G(1);
      wk = (j == b.iy0 || j == b.iy1) ? G(0.5) : G(1);
    }
    const A w = A(dA * wj * wk);
    const A nf = (f & 1) ? A(1) : A(-1);
    // J = n x H, M = -n x E (dft_farfield.cl:446-473); slots N 0..2, L 3..5
    int jc[2], js[2], mc[2], ms[2];  // component slots and source buffers
    A jsg[2], msg[2];
    if (f <= 1) {
      jc[0] = 1; js[0] = 5; jsg[0] = nf;   // Jy =  n Hz
      jc[1] = 2; js[1] = 4; jsg[1] = -nf;  // Jz = -n Hy
      mc[0] = 1; ms[0] = 2; msg[0] = -nf;  // My = -n Ez
      mc[1] = 2; ms[1] = 1; msg[1] = nf;   // Mz =  n Ey
    } else if (f <= 3) {
      jc[0] = 0; js[0] = 5; jsg[0] = -nf;  // Jx = -n Hz
      jc[1] = 2; js[1] = 3; jsg[1] = nf;   // Jz =  n Hx
      mc[0] = 0; ms[0] = 2; msg[0] = nf;   // Mx =  n Ez
      mc[1] = 2; ms[1] = 0; msg[1] = -nf;  // Mz = -n Ex
    } else {
      jc[0] = 0; js[0] = 4; jsg[0] = nf;   // Jx =  n Hy
      jc[1] = 1; js[1] = 3; jsg[1] = -nf;  // Jy = -n Hx
      mc[0] = 0; ms[0] = 1; msg[0] = -nf;  // Mx = -n Ey
      mc[1] = 1; ms[1] = 0; msg[1] = nf;   // My =  n Ex
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      A re, im;
      ld(js[q], s, re, im);
      re *= jsg[q];
      im *= jsg[q];
      const A tr = re * pc - im * ps, ti = re * ps + im * pc;
      s6[2 * jc[q]] += double(tr * w);
      s6[2 * jc[q] + 1] += double(ti * w);
      ld(ms[q], s, re, im);
      re *= msg[q];
      im *= msg[q];
      const A ur = re * pc - im * ps, ui = re * ps + im * pc;
      s6[6 + 2 * mc[q]] += double(ur * w);
      s6[6 + 2 * mc[q] + 1] += double(ui * w);
    }
  }
  // block reduction: wave shuffles, then the 4 wave sums in LDS
  __shared__ double red[4][12];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 12; ++q) {
    double v = s6[q];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) red[wave][q] = v;
  }
  __syncthreads();
  if (threadIdx.x < 12) {
    const int nw = blockDim.x >> 6;
    double v = 0.0;
    for (int w2 = 0; w2 < nw; ++w2) v += red[w2][threadIdx.x];
    part[((size_t)ob * nblk + blockIdx.x) * 12 + threadIdx.x] = v;
  }
}

// Stage 2: sum the partials of each observation in block order, then
// N, L -> far E, H (dft_farfield.cl:559-635) in A:
//   E = -j k e^{-jkr}/(4 pi r) (eta0 N_t + l_sign r x L),  H = -(r x E)/eta0.
template <typename A>
__global__ void k_n2f_eh(int n_obs, int nblk, const double *__restrict__ part,
                         const double *__restrict__ obs, A k_wave, A eta0, A l_sign,
                         double *__restrict__ eh) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_obs) return;
  double nl[12];
  for (int q = 0; q < 12; ++q) nl[q] = 0.0;
  for (int bq = 0; bq < nblk; ++bq)
    for (int q = 0; q < 12; ++q) nl[q] += part[((size_t)p * nblk + bq) * 12 + q];
  const A ox = A(obs[3 * p]), oy = A(obs[3 * p + 1]), oz = A(obs[3 * p + 2]);
  A r = sqrt(ox * ox + oy * oy + oz * oz);
  if (r < A(1.0e-30)) r = A(1.0e-30);
  const A rh[3] = {ox / r, oy / r, oz / r};
  A Nr[3], Ni[3], Lr[3], Li[3];
  for (int c = 0; c < 3; ++c) {
    Nr[c] = A(nl[2 * c]);
    Ni[c] = A(nl[2 * c + 1]);
    Lr[c] = A(nl[6 + 2 * c]);
    Li[c] = A(nl[6 + 2 * c + 1]);
  }
  const A ang = k_wave * r;
  const A er = cos(ang), ei = -sin(ang);
  const A scale = k_wave / (A(4) * A(3.14159265358979323846) * r);
  // pref = (0, -scale) * (er, ei)
  const A pr = scale * ei, pi = -scale * er;
  const A xLr[3] = {rh[1] * Lr[2] - rh[2] * Lr[1], rh[2] * Lr[0] - rh[0] * Lr[2],
                    rh[0] * Lr[1] - rh[1] * Lr[0]};
  const A xLi[3] = {rh[1] * Li[2] - rh[2] * Li[1], rh[2] * Li[0] - rh[0] * Li[2],
                    rh[0] * Li[1] - rh[1] * Li[0]};
  const A dr = rh[0] * Nr[0] + rh[1] * Nr[1] + rh[2] * Nr[2];
  const A di = rh[0] * Ni[0] + rh[1] * Ni[1] + rh[2] * Ni[2];
  A Er[3], Ei[3];
  for (int c = 0; c < 3; ++c) {
    const A tr = eta0 * (Nr[c] - dr * rh[c]) + l_sign * xLr[c];
    const A ti = eta0 * (Ni[c] - di * rh[c]) + l_sign * xLi[c];
    Er[c] = pr * tr - pi * ti;
    Ei[c] = pr * ti + pi * tr;
  }
  const A Hr[3] = {-(rh[1] * Er[2] - rh[2] * Er[1]) / eta0,
                   -(rh[2] * Er[0] - rh[0] * Er[2]) / eta0,
                   -(rh[0] * Er[1] - rh[1] * Er[0]) / eta0};
  const A Hi[3] = {-(rh[1] * Ei[2] - rh[2] * Ei[1]) / eta0,
                   -(rh[2] * Ei[0] - rh[0] * Ei[2]) / eta0,
                   -(rh[0] * Ei[1] - rh[1] * Ei[0]) / eta0};
  double *o = eh + (size_t)p * 12;
  for (int c = 0; c < 3; ++c) {
    o[2 * c] = Er[c];
    o[2 * c + 1] = Ei[c];
    o[6 + 2 * c] = Hr[c];
    o[6 + 2 * c + 1] = Hi[c];
  }
}

static Box box_of(const hfdtd_n2f *m) {
  Box b;
  b.ix0 = m->box[0]; b.ix1 = m->box[1];
  b.iy0 = m->box[2]; b.iy1 = m->box[3];
  b.iz0 = m->box[4]; b.iz1 = m->box[5];
  for (int q = 0; q < 6; ++q) b.off[q] = m->off[q];
  b.n_face = m->n_face;
  return b;
}

static int ensure_scratch(hfdtd_n2f *m, size_t n) {
  if (m->scratch_len >= n) return HFDTD_OK;
  if (m->scratch) (void)hipFree(m->scratch);
  m->scratch = nullptr;
  m->scratch_len = 0;
  HF_HIP(hipMalloc((void **)&m->scratch, n * sizeof(double)));
  m->scratch_len = n;
  return HFDTD_OK;
}

template <typename T, typename A, bool KAHAN>
static int accumulate_t(hfdtd_n2f *m, const double *phase) {
  Engine *e = m->e;  // the field volumes as the engine holds them (a row-padded run's too)
  const Dims3 g{e->nx, e->ny, e->nz, e->sk};
  const Box b = box_of(m);
  for (int f0 = 0; f0 < m->nfreq; f0 += kMaxFreqPerLaunch) {
    const int nf = std::min(kMaxFreqPerLaunch, m->nfreq - f0);
    PhaseArg<A> ph;
    for (int q = 0; q < nf; ++q)
      for (int c = 0; c < 4; ++c) ph.v[q][c] = A(phase[(f0 + q) * 4 + c]);
    hipLaunchKernelGGL((k_n2f_accumulate<T, A, KAHAN>), dim3(ceil_div(m->n_face, 256)),
                       dim3(256), 0, e->stream, g, b, m->nfreq, f0, nf,
                       (const T *)e->field(0), (const T *)e->field(1), (const T *)e->field(2),
                       (const T *)e->field(3), (const T *)e->field(4), (const T *)e->field(5),
                       (A *)m->acc, (A *)m->cmp, ph);
    HF_HIP(hipGetLastError());
  }
  return HFDTD_OK;
}

template <typename A, bool KAHAN>
static int contract_t(hfdtd_n2f *m, hipStream_t st) {
  const long long waves = m->ppad / 32;
  hipLaunchKernelGGL((k_n2f_contract<A, KAHAN>), dim3((unsigned)((waves + 3) / 4)), dim3(256), 0,
                     st, box_of(m), m->buf, m->ppad, m->pe, m->we, m->wh, m->kfill, m->nfreq,
                     (A *)m->acc, (A *)m->cmp);
  HF_HIP(hipGetLastError());
  m->kfill = 0;
  return HFDTD_OK;
}

// Fold the buffered rows into the sums, on the stream that wrote the rows.
static int contract(hfdtd_n2f *m, hipStream_t st) {
  if (m->acc_mode == HFDTD_ACC_FP64) return contract_t<double, false>(m, st);
  if (m->acc_mode == HFDTD_ACC_KAHAN) return contract_t<float, true>(m, st);
  return contract_t<float, false>(m, st);
}

static int launch_gather(Engine *e, const GatherTail &t) {
  hipLaunchKernelGGL(k_n2f_gather, dim3(t.blocks()), dim3(256), 0, e->stream, t);
  HF_HIP(hipGetLastError());
  return HFDTD_OK;
}

// Fold a partially filled block into the sums (before anything reads them);
// a held gather first writes its row.
static int n2f_flush(hfdtd_n2f *m) {
  HF_HIP(hipSetDevice(m->e->device));
  int rc = flush_gather(m->e);
  if (rc) return rc;
  if (m->K == 0 || m->kfill == 0) return HFDTD_OK;
  return contract(m, m->e->stream);
}

// Drop a monitor's buffers: nothing may still run on them.
static int quiesce(Engine *e) {
  HF_HIP(hipSetDevice(e->device));
  int rc = flush_gather(e);
  if (rc) return rc;
  HF_HIP(hipStreamSynchronize(e->stream));
  return HFDTD_OK;
}

// One step's co-located samples into row kfill, and the contraction when
// the block is full.  With the engine's fuse_gather on, a gather that does
// not fill the block is held: the next march sweep runs it in blocks after
// its own (it reads state[cur] and record cur, which that sweep only reads),
// and any other call runs it first on its own.
static int gather_step(hfdtd_n2f *m, const double *phase) {
  Engine *e = m->e;  // the field volumes as the engine holds them (a row-padded run's too)
  int rc = flush_gather(e);  // one held at a time (another monitor's, say)
  if (rc) return rc;
  GatherTail t;
  for (int q = 0; q < 32; ++q) {
    const int fr = q >> 1, c = q & 1;
    t.ph.e[q] = fr < m->nfreq ? (float)phase[fr * 4 + c] : 0.0f;
    t.ph.h[q] = fr < m->nfreq ? (float)phase[fr * 4 + 2 + c] : 0.0f;
  }
  t.g = Dims3{e->nx, e->ny, e->nz, e->sk};
  t.b = box_of(m);
  for (int c = 0; c < 6; ++c) t.f[c] = (const float *)e->field(c);
  t.row = m->buf + (size_t)m->kfill * m->ppad;
  t.pe = m->pe;
  t.we = m->we + (size_t)m->kfill * 32;
  t.wh = m->wh + (size_t)m->kfill * 32;
  t.zr = ZRead{};
  if (e->zrec_valid(m)) {  // the last step recorded this box's z faces
    t.zr.buf = (const float4 *)e->zrec.half(e->cur);
    t.zr.i0 = e->zrec.i0;
    t.zr.j0 = e->zrec.j0;
    t.zr.nj = e->zrec.j1 - e->zrec.j0 + 1;
    t.zr.quads = (long long)e->zrec.quads();
  }
  if (++m->kfill < m->K && e->fuse_gather && !e->view_shadow) {
    if (!e->pgather) e->pgather = new GatherTail();
    *e->pgather = t;
    e->pgather->held = true;
    return HFDTD_OK;
  }
  if ((rc = launch_gather(e, t))) return rc;
  return m->kfill == m->K ? contract(m, e->stream) : HFDTD_OK;
}

int flush_gather(Engine *e) {
  if (!e->pgather || !e->pgather->held) return HFDTD_OK;
  e->pgather->held = false;
  return launch_gather(e, *e->pgather);
}

void free_gather(Engine *e) {
  delete e->pgather;
  e->pgather = nullptr;
}

// The engine's z-face record for this monitor's box: asked for with K-step
// buffering on a float32 engine when no other monitor holds it, the z faces
// lie at k >= 1 and are at least two cells apart (so the four recorded
// columns are distinct); otherwise the gather reads the fields.
static void zrec_acquire(hfdtd_n2f *m) {
  Engine *e = m->e;
  if (e->zrec.owner || e->fp64) return;
  const int *b = m->box;
  if (b[4] < 1 || b[5] < b[4] + 2) return;
  ZRec z;
  z.kf[0] = b[4];
  z.kf[1] = b[5];
  z.i0 = b[0];
  z.i1 = std::min(b[1] + 1, e->nx - 1);
  z.j0 = b[2];
  z.j1 = std::min(b[3] + 1, e->ny - 1);
  z.bytes = 2 * 2 * z.quads() * 4 * sizeof(float);
  if (z.bytes >= 0x40000000ull) return;  // the sweep's drop offsets need a record below 1 GB
  if (hipMalloc((void **)&z.buf, 2 * z.bytes) != hipSuccess) {  // one record per slab parity
    (void)hipGetLastError();
    return;
  }
  z.owner = m;
  z.epoch = ~0ull;  // nothing recorded yet
  e->zrec = z;
}

static void zrec_release(hfdtd_n2f *m) {
  ZRec &z = m->e->zrec;
  if (z.owner != m) return;
  (void)quiesce(m->e);
  if (z.buf) (void)hipFree(z.buf);
  z = ZRec{};
}

static void free_buffering(hfdtd_n2f *m) {
  (void)quiesce(m->e);
  zrec_release(m);
  for (float *p : {m->buf, m->we, m->wh})
    if (p) (void)hipFree(p);
  m->buf = m->we = m->wh = nullptr;
  m->K = m->kfill = 0;
}

constexpr int kFarfieldChunk = 2048;  // face samples per stage-1 block

// Doubles of scratch one far-field call needs: the stage-1 partials
// (12 per observation and chunk of face samples), the observation points
// and the E/H results.
static size_t farfield_scratch_doubles(const hfdtd_n2f *m, int n_obs) {
  const size_t nblk = std::max<size_t>(1, ceil_div(m->n_face, kFarfieldChunk));
  return (size_t)n_obs * nblk * 12 + (size_t)n_obs * 3 + (size_t)n_obs * 12;
}

// The far field into ``eh`` with ``scratch`` (at least
// farfield_scratch_doubles doubles of device memory); every word of it the
// kernels read is written first in the same call.
template <typename G, typename A>
static int farfield_t(hfdtd_n2f *m, int fr, const double *obs, int n_obs, double k_wave,
                      double eta0, int l_sign, double *eh, double *scratch) {
  Engine *e = m->e;
  HF_SETTLE(e);
  const int chunk = kFarfieldChunk;
  const int nblk = std::max(1, (int)ceil_div(m->n_face, chunk));
  double *part = scratch;
  double *d_obs = part + (size_t)n_obs * nblk * 12;
  double *d_eh = d_obs + (size_t)n_obs * 3;
  HF_HIP(hipMemcpyAsync(d_obs, obs, (size_t)n_obs * 3 * sizeof(double), hipMemcpyHostToDevice,
                        e->stream));
  hipLaunchKernelGGL((k_n2f_nl<G, A>), dim3(nblk, n_obs), dim3(256), 0, e->stream, box_of(m),
                     m->nfreq, fr, chunk, (const A *)m->acc, d_obs, G(k_wave), G(e->dl), part);
  HF_HIP(hipGetLastError());
  hipLaunchKernelGGL((k_n2f_eh<A>), dim3(ceil_div(n_obs, 64)), dim3(64), 0, e->stream, n_obs,
                     nblk, part, d_obs, A(k_wave), A(eta0), A(l_sign), d_eh);
  HF_HIP(hipGetLastError());
  HF_HIP(hipMemcpyAsync(eh, d_eh, (size_t)n_obs * 12 * sizeof(double), hipMemcpyDeviceToHost,
                        e->stream));
  HF_HIP(hipStreamSynchronize(e->stream));
  return HFDTD_OK;
}

// Scratch of hfdtd_n2f_farfield_noalloc: one copy per device, held for the
// life of the process.  Its name stands for the device array only in device
// code; host code uses the address hipGetSymbolAddress returns for the
// current device.
constexpr size_t kNoallocScratchDoubles = size_t(1) << 20;  // 8 MiB
static __device__ __align__(256) double scratch[kNoallocScratchDoubles];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kMaxDevices] = {};

// Device address of ``scratch`` on ``device`` (looked up once per device,
// with that device current).
static int noalloc_scratch(int device, double **out) {
  if (device < 0 || device >= kMaxDevices) return fail(HFDTD_EUNSUPPORTED, "device ordinal too large");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    HF_HIP(hipSetDevice(device));
    void *ptr = nullptr;
    HF_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = (double *)ptr;
  }
  *out = g_scratch_ptr[device];
  return HFDTD_OK;
}

template <typename Run>
static int farfield_dispatch(hfdtd_n2f *m, int freq, int geom_fp64, Run run) {
  const bool g64 = geom_fp64 || m->e->fp64;
  if (m->acc_elem == 8) return g64 ? run(double{}, double{}) : run(float{}, double{});
  return g64 ? run(double{}, float{}) : run(float{}, float{});
}

}  // namespace hfdtd

using namespace hfdtd;

extern "C" {

int hfdtd_n2f_create(hfdtd_engine *h, const int box[6], int nfreq, int acc_mode,
                     hfdtd_n2f **out) {
  if (!h || !box || !out) return fail(HFDTD_EINVAL, "NULL argument");
  *out = nullptr;
  Engine *e = &h->s;
  HF_SETTLE(e);
  if (nfreq < 1) return fail(HFDTD_EINVAL, "nfreq must be >= 1");
  if (acc_mode < HFDTD_ACC_REAL || acc_mode > HFDTD_ACC_KAHAN)
    return fail(HFDTD_EINVAL, "unknown accumulator mode");
  if (acc_mode == HFDTD_ACC_KAHAN && e->fp64)
    return fail(HFDTD_EUNSUPPORTED, "Kahan accumulation is for float32 engines");
  const int lim[3] = {e->nx, e->ny, e->nz};
  for (int a = 0; a < 3; ++a)
    if (box[2 * a] < 0 || box[2 * a + 1] >= lim[a] || box[2 * a] > box[2 * a + 1])
      return fail(HFDTD_EINVAL, "monitor box outside the grid or empty");
  hfdtd_n2f *m = new hfdtd_n2f();
  m->e = e;
  for (int q = 0; q < 6; ++q) m->box[q] = box[q];
  m->nfreq = nfreq;
  m->acc_mode = acc_mode;
  const long nxf = box[1] - box[0] + 1, nyf = box[3] - box[2] + 1, nzf = box[5] - box[4] + 1;
  const long counts[6] = {nyf * nzf, nyf * nzf, nxf * nzf, nxf * nzf, nxf * nyf, nxf * nyf};
  long tot = 0;
  for (int q = 0; q < 6; ++q) {
    m->off[q] = (int)tot;
    tot += counts[q];
  }
  m->n_face = (int)tot;
  const bool acc64 = acc_mode == HFDTD_ACC_FP64 || (acc_mode == HFDTD_ACC_REAL && e->fp64);
  m->acc_elem = acc64 ? 8 : 4;
  const size_t bytes = (size_t)6 * nfreq * tot * 2 * m->acc_elem;
  hipError_t err = hipSetDevice(e->device);
  if (err == hipSuccess) err = hipMalloc(&m->acc, bytes);
  if (err == hipSuccess && acc_mode == HFDTD_ACC_KAHAN) err = hipMalloc(&m->cmp, bytes);
  if (err == hipSuccess) err = hipMemsetAsync(m->acc, 0, bytes, e->stream);
  if (err == hipSuccess && m->cmp) err = hipMemsetAsync(m->cmp, 0, bytes, e->stream);
  if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
  if (err != hipSuccess) {
    int rc = hip_fail(err, "hfdtd_n2f_create");
    if (m->acc) (void)hipFree(m->acc);
    if (m->cmp) (void)hipFree(m->cmp);
    delete m;
    return rc;
  }
  *out = m;
  return HFDTD_OK;
}

int hfdtd_n2f_destroy(hfdtd_n2f *m) {
  if (!m) return HFDTD_OK;
  (void)quiesce(m->e);
  if (m->acc) (void)hipFree(m->acc);
  if (m->cmp) (void)hipFree(m->cmp);
  if (m->snap) (void)hipFree(m->snap);
  if (m->scratch) (void)hipFree(m->scratch);
  free_buffering(m);
  delete m;
  return HFDTD_OK;
}

int hfdtd_n2f_layout(const hfdtd_n2f *m, int *n_face, int offsets[6]) {
  if (!m) return fail(HFDTD_EINVAL, "NULL argument");
  if (n_face) *n_face = m->n_face;
  if (offsets)
    for (int q = 0; q < 6; ++q) offsets[q] = m->off[q];
  return HFDTD_OK;
}

int hfdtd_n2f_accumulate(hfdtd_n2f *m, const double *phase) {
  if (!m || !phase) return fail(HFDTD_EINVAL, "NULL argument");
  Engine *e = m->e;
  HF_HIP(hipSetDevice(e->device));
  if (m->K > 0) return gather_step(m, phase);
  if (e->fp64) return accumulate_t<double, double, false>(m, phase);
  if (m->acc_mode == HFDTD_ACC_FP64) return accumulate_t<float, double, false>(m, phase);
  if (m->acc_mode == HFDTD_ACC_KAHAN) return accumulate_t<float, float, true>(m, phase);
  return accumulate_t<float, float, false>(m, phase);
}

int hfdtd_n2f_zrecord(const hfdtd_n2f *m, int *on) {
  if (!m || !on) return fail(HFDTD_EINVAL, "NULL argument");
  *on = m->e->zrec.owner == m ? 1 : 0;
  return HFDTD_OK;
}

int hfdtd_n2f_flush(hfdtd_n2f *m) {
  if (!m) return fail(HFDTD_EINVAL, "NULL argument");
  return n2f_flush(m);
}

int hfdtd_n2f_snapshot(hfdtd_n2f *m) {
  if (!m) return fail(HFDTD_EINVAL, "NULL argument");
  int rc = n2f_flush(m);
  if (rc) return rc;
  Engine *e = m->e;
  HF_SETTLE(e);
  HF_HIP(hipSetDevice(e->device));
  const size_t bytes = (size_t)6 * m->nfreq * m->n_face * 2 * m->acc_elem;
  if (!m->snap) HF_HIP(hipMalloc(&m->snap, bytes));
  HF_HIP(hipMemcpyAsync(m->snap, m->acc, bytes, hipMemcpyDeviceToDevice, e->stream));
  return HFDTD_OK;
}

int hfdtd_n2f_rel_change(hfdtd_n2f *m, int freq, double *num, double *den) {
  if (!m || !num || !den) return fail(HFDTD_EINVAL, "NULL argument");
  if (freq < 0 || freq >= m->nfreq) return fail(HFDTD_EINVAL, "frequency index out of range");
  if (!m->snap) return fail(HFDTD_EINVAL, "no snapshot");
  int rc = n2f_flush(m);
  if (rc) return rc;
  Engine *e = m->e;
  HF_SETTLE(e);
  HF_HIP(hipSetDevice(e->device));
  const int nblk = (int)ceil_div(m->n_face, 256);
  rc = ensure_scratch(m, (size_t)2 * nblk);
  if (rc) return rc;
  if (m->acc_elem == 8)
    hipLaunchKernelGGL((k_n2f_relchange<double>), dim3(nblk), dim3(256), 0, e->stream,
                       m->n_face, m->nfreq, freq, (const double *)m->acc,
                       (const double *)m->snap, m->scratch);
  else
    hipLaunchKernelGGL((k_n2f_relchange<float>), dim3(nblk), dim3(256), 0, e->stream,
                       m->n_face, m->nfreq, freq, (const float *)m->acc,
                       (const float *)m->snap, m->scratch);
  HF_HIP(hipGetLastError());
  std::vector<double> host((size_t)2 * nblk);
  HF_HIP(hipMemcpyAsync(host.data(), m->scratch, host.size() * sizeof(double),
                        hipMemcpyDeviceToHost, e->stream));
  HF_HIP(hipStreamSynchronize(e->stream));
  double sn = 0.0, sd = 0.0;
  for (int q = 0; q < nblk; ++q) {
    sn += host[2 * q];
    sd += host[2 * q + 1];
  }
  *num = sn;
  *den = sd;
  return HFDTD_OK;
}

int hfdtd_n2f_read(hfdtd_n2f *m, int comp, int freq, double *host) {
  if (!m || !host) return fail(HFDTD_EINVAL, "NULL argument");
  if (comp < 0 || comp > 5 || freq < 0 || freq >= m->nfreq)
    return fail(HFDTD_EINVAL, "component or frequency out of range");
  int rc = n2f_flush(m);
  if (rc) return rc;
  Engine *e = m->e;
  HF_SETTLE(e);
  HF_HIP(hipSetDevice(e->device));
  const size_t n2 = (size_t)m->n_face * 2;
  const size_t off = ((size_t)comp * m->nfreq + freq) * n2;
  if (m->acc_elem == 8) {
    HF_HIP(hipMemcpyAsync(host, (const double *)m->acc + off, n2 * 8, hipMemcpyDeviceToHost,
                          e->stream));
    HF_HIP(hipStreamSynchronize(e->stream));
  } else {
    std::vector<float> tmp(n2);
    HF_HIP(hipMemcpyAsync(tmp.data(), (const float *)m->acc + off, n2 * 4,
                          hipMemcpyDeviceToHost, e->stream));
    HF_HIP(hipStreamSynchronize(e->stream));
    for (size_t q = 0; q < n2; ++q) host[q] = tmp[q];
  }
  return HFDTD_OK;
}

int hfdtd_n2f_farfield(hfdtd_n2f *m, int freq, const double *obs, int n_obs, double k_wave,
                       double eta0, int l_sign, int geom_fp64, double *eh) {
  if (!m || !obs || !eh) return fail(HFDTD_EINVAL, "NULL argument");
  if (n_obs < 1) return fail(HFDTD_EINVAL, "n_obs must be >= 1");
  if (freq < 0 || freq >= m->nfreq) return fail(HFDTD_EINVAL, "frequency index out of range");
  if (l_sign != 1 && l_sign != -1) return fail(HFDTD_EINVAL, "l_sign must be +1 or -1");
  int rc = n2f_flush(m);
  if (rc) return rc;
  Engine *e = m->e;
  HF_SETTLE(e);
  HF_HIP(hipSetDevice(e->device));
  rc = ensure_scratch(m, farfield_scratch_doubles(m, n_obs));
  if (rc) return rc;
  return farfield_dispatch(m, freq, geom_fp64, [&](auto g, auto a) {
    return farfield_t<decltype(g), decltype(a)>(m, freq, obs, n_obs, k_wave, eta0, l_sign, eh,
                                                m->scratch);
  });
}

int hfdtd_n2f_farfield_noalloc(hfdtd_n2f *m, int freq, const double *obs, int n_obs,
                               double k_wave, double eta0, int l_sign, int geom_fp64,
                               double *eh) {
  if (!m || !obs || !eh) return fail(HFDTD_EINVAL, "NULL argument");
  if (n_obs < 1) return fail(HFDTD_EINVAL, "n_obs must be >= 1");
  if (freq < 0 || freq >= m->nfreq) return fail(HFDTD_EINVAL, "frequency index out of range");
  if (l_sign != 1 && l_sign != -1) return fail(HFDTD_EINVAL, "l_sign must be +1 or -1");
  if (farfield_scratch_doubles(m, n_obs) > kNoallocScratchDoubles)
    return fail(HFDTD_EINVAL,
                "n_obs * (12 * ceil(n_face / 2048) + 15) exceeds the 1048576 doubles of the "
                "allocation-free far field's scratch");
  double *buf = nullptr;
  int rc = noalloc_scratch(m->e->device, &buf);
  if (rc) return rc;
  rc = n2f_flush(m);
  if (rc) return rc;
  HF_HIP(hipSetDevice(m->e->device));
  return farfield_dispatch(m, freq, geom_fp64, [&](auto g, auto a) {
    return farfield_t<decltype(g), decltype(a)>(m, freq, obs, n_obs, k_wave, eta0, l_sign, eh,
                                                buf);
  });
}

int hfdtd_n2f_set_buffering(hfdtd_n2f *m, int K) {
  if (!m) return fail(HFDTD_EINVAL, "NULL argument");
  if (K != 0 && (K < 2 || K > 256 || (K & 1))) return fail(HFDTD_EINVAL, "K must be 0 or even, 2..256");
  if (K != 0 && m->e->fp64)
    return fail(HFDTD_EUNSUPPORTED, "buffered face DFT needs a float32 engine");
  if (K != 0 && m->nfreq > 16) return fail(HFDTD_EUNSUPPORTED, "buffered face DFT: at most 16 frequencies");
  int rc = n2f_flush(m);
  if (rc) return rc;
  HF_HIP(hipSetDevice(m->e->device));
  HF_HIP(hipStreamSynchronize(m->e->stream));
  free_buffering(m);
  if (K == 0) return HFDTD_OK;
  const long long rows = 2LL * m->n_face;
  m->pe = (rows + 31) / 32 * 32;
  m->ppad = 2 * m->pe;
  hipError_t err = hipMalloc((void **)&m->buf, (size_t)K * m->ppad * sizeof(float));
  if (err == hipSuccess) err = hipMalloc((void **)&m->we, (size_t)K * 32 * sizeof(float));
  if (err == hipSuccess) err = hipMalloc((void **)&m->wh, (size_t)K * 32 * sizeof(float));
  if (err == hipSuccess) err = hipMemsetAsync(m->buf, 0, (size_t)K * m->ppad * sizeof(float), m->e->stream);
  if (err == hipSuccess) err = hipMemsetAsync(m->we, 0, (size_t)K * 32 * sizeof(float), m->e->stream);
  if (err == hipSuccess) err = hipMemsetAsync(m->wh, 0, (size_t)K * 32 * sizeof(float), m->e->stream);
  if (err == hipSuccess) err = hipStreamSynchronize(m->e->stream);
  if (err != hipSuccess) {
    free_buffering(m);
    return hip_fail(err, "hfdtd_n2f_set_buffering");
  }
  m->K = K;
  zrec_acquire(m);
  return HFDTD_OK;
}

}  // extern "C"
