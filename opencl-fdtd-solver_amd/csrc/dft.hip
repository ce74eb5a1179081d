// Multi-frequency DFT of field planes, buffered over K steps.
//
// The running sums of the reference's DFT monitors (acc += v(t) e^{jwt} dt,
// e.g. the plane convention of tests/harness cases.py:179-197) are, over a
// block of K steps, a dense contraction:
//     acc[p][2f + {re, im}] += sum_k buf[k][p] * W[k][2f + {re, im}],
//     W[k][2f] = cos(w_f t_k) dt,  W[k][2f + 1] = sin(w_f t_k) dt.
// Each step, k_dft_gather copies the plane samples into row k of buf (and
// the step's phases into row k of W); every K steps k_dft_contract performs
// the contraction on the matrix cores (v_mfma_f32_32x32x2f32: 32 samples x
// 32 phase columns per wavefront, two steps per instruction) and adds the
// block sums into the accumulators.  Accumulators are read and written once
// per block instead of once per step.  The direct mode (k_dft_direct) is
// the per-step form, with FP32, Kahan-compensated FP32 or FP64 sums, kept
// for comparison and for FP64 engines.

#include <algorithm>
#include <vector>

#include "fdtd_core.h"

namespace hfdtd {

constexpr int kDftMaxFreq = 16;  // 2F <= 32 phase columns (one MFMA tile)

// Sample p = c * (n0 * n1) + a * n1 + b is component vol[c] at element
// base + a * sa + b * sb.
struct Plane {
  int ncomp;
  int vol[6];
  long long base, sa, sb;
  int n0, n1;
};

struct PhaseRow {
  float v[2 * kDftMaxFreq];  // cos, sin pairs (times dt), zero beyond 2F
};

template <typename A>
struct PhaseDirect {
  A re[kDftMaxFreq], im[kDftMaxFreq];
};

__device__ __forceinline__ long long plane_addr(const Plane &pl, long long vs, long long p) {
  const long long per = (long long)pl.n0 * pl.n1;
  const int c = (int)(p / per);
  const long long r = p - c * per;
  const long long a = r / pl.n1, b = r - a * pl.n1;
  return pl.vol[c] * vs + pl.base + a * pl.sa + b * pl.sb;
}

// Row k of the sample buffer (and of the phase table, by block 0).
template <typename T>
__global__ __launch_bounds__(256) void k_dft_gather(const T *__restrict__ st, long long vs,
                                                    Plane pl, long long P, float *__restrict__ row,
                                                    PhaseRow ph, float *__restrict__ wrow) {
  const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
  if (blockIdx.x == 0 && threadIdx.x < 2 * kDftMaxFreq) wrow[threadIdx.x] = ph.v[threadIdx.x];
  if (p >= P) return;
  row[p] = (float)st[plane_addr(pl, vs, p)];
}

// acc[p][2f..] += v(p) * phase, per step.
template <typename T, typename A, bool KAHAN>
__global__ __launch_bounds__(256) void k_dft_direct(const T *__restrict__ st, long long vs,
                                                    Plane pl, long long P, int F, PhaseDirect<A> ph,
                                                    A *__restrict__ acc, A *__restrict__ cmp) {
  const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
  if (p >= P) return;
  const A v = (A)st[plane_addr(pl, vs, p)];
  for (int f = 0; f < F; ++f) {
    const size_t o = (size_t)p * (2 * F) + 2 * f;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const A term = v * (q == 0 ? ph.re[f] : ph.im[f]);
      if (!KAHAN) {
        acc[o + q] += term;
      } else {
        const A s = acc[o + q];
        const A y = term - cmp[o + q];
        const A t = s + y;
        cmp[o + q] = (t - s) - y;
        acc[o + q] = t;
      }
    }
  }
}

typedef float f32x16 __attribute__((ext_vector_type(16)));

// Block contraction.  One wavefront per 32 samples; lanes 0-31 feed step
// 2t, lanes 32-63 step 2t+1 of each MFMA (A: buf[step][p0 + lane % 32],
// B: W[step][lane % 32]).  Rows of W at or beyond `kvalid` count as zero.
// Output D[i][j] sits in register r of lane l with i = (r/4)*8 + (l/32)*4 +
// r%4, j = l%32 (the gfx950 layout of v_mfma_f32_32x32x2f32).
template <typename A, bool KAHAN>
__global__ __launch_bounds__(256) void k_dft_contract(const float *__restrict__ buf,
                                                      long long Ppad, const float *__restrict__ W,
                                                      int kvalid, long long P, int F2,
                                                      A *__restrict__ acc, A *__restrict__ cmp) {
  const int lane = threadIdx.x & 63;
  const long long p0 = ((long long)blockIdx.x * 4 + (threadIdx.x >> 6)) * 32;
  if (p0 >= Ppad) return;  // wave-uniform
  const int half = lane >> 5, col = lane & 31;
  f32x16 d = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  const int pairs = (kvalid + 1) / 2;
  for (int t = 0; t < pairs; ++t) {
    const int k = 2 * t + half;
    const float a = buf[(long long)k * Ppad + p0 + col];
    const float b = k < kvalid ? W[k * 32 + col] : 0.0f;
    d = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, d, 0, 0, 0);
  }
  if (col >= F2) return;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const long long p = p0 + (r / 4) * 8 + half * 4 + (r % 4);
    if (p >= P) continue;
    const size_t o = (size_t)p * F2 + col;
    const A term = (A)d[r];
    if (!KAHAN) {
      acc[o] += term;
    } else {
      const A s = acc[o];
      const A y = term - cmp[o];
      const A tt = s + y;
      cmp[o] = (tt - s) - y;
      acc[o] = tt;
    }
  }
}

}  // namespace hfdtd

struct hfdtd_dft {
  hfdtd::Engine *e = nullptr;
  hfdtd::Plane pl{};
  long long P = 0, Ppad = 0;
  int F = 0, K = 0, mode = 0, accum = 0;
  int k = 0;                // steps in the current block
  float *buf = nullptr;     // [K][Ppad]
  float *W = nullptr;       // [K][32]
  void *acc = nullptr;      // [P][2F], float or double
  float *cmp = nullptr;     // Kahan compensation, [P][2F]
  size_t acc_elem = 4;
};

namespace hfdtd {

template <typename A, bool KAHAN>
static int contract_t(hfdtd_dft *d) {
  const long long waves = d->Ppad / 32;
  hipLaunchKernelGGL((k_dft_contract<A, KAHAN>), dim3((unsigned)((waves + 3) / 4)), dim3(256), 0,
                     d->e->stream, d->buf, d->Ppad, d->W, d->k, d->P, 2 * d->F, (A *)d->acc,
                     (A *)d->cmp);
  HF_HIP(hipGetLastError());
  d->k = 0;
  return HFDTD_OK;
}

static int dft_flush(hfdtd_dft *d) {
  if (d->mode != HFDTD_DFT_BUFFERED || d->k == 0) return HFDTD_OK;
  HF_HIP(hipSetDevice(d->e->device));
  if (d->accum == HFDTD_ACC_FP64) return contract_t<double, false>(d);
  if (d->accum == HFDTD_ACC_KAHAN) return contract_t<float, true>(d);
  return contract_t<float, false>(d);
}

template <typename T, typename A, bool KAHAN>
static int direct_t(hfdtd_dft *d, const double *phase) {
  PhaseDirect<A> ph;
  for (int f = 0; f < kDftMaxFreq; ++f) {
    ph.re[f] = f < d->F ? A(phase[2 * f]) : A(0);
    ph.im[f] = f < d->F ? A(phase[2 * f + 1]) : A(0);
  }
  Engine *e = d->e;
  HF_SETTLE(e);
  hipLaunchKernelGGL((k_dft_direct<T, A, KAHAN>), dim3((unsigned)ceil_div(d->P, 256)), dim3(256),
                     0, e->stream, (const T *)e->field(0), (long long)e->vstride, d->pl,
                     d->P, d->F, ph, (A *)d->acc, (A *)d->cmp);
  HF_HIP(hipGetLastError());
  return HFDTD_OK;
}

}  // namespace hfdtd

using namespace hfdtd;

extern "C" {

int hfdtd_dft_plane_create(hfdtd_engine *h, int axis, int index, int comp_mask, int nfreq, int K,
                           int mode, int accum, hfdtd_dft **out) {
  if (!h || !out) return fail(HFDTD_EINVAL, "NULL argument");
  *out = nullptr;
  Engine *e = &h->s;
  HF_SETTLE(e);
  if (axis < 0 || axis > 2) return fail(HFDTD_EINVAL, "axis must be 0 (x), 1 (y) or 2 (z)");
  const int nax = axis == 0 ? e->nx : (axis == 1 ? e->ny : e->nz);
  if (index < 0 || index >= nax) return fail(HFDTD_EINVAL, "plane index outside the grid");
  if (comp_mask <= 0 || comp_mask >= 64) return fail(HFDTD_EINVAL, "comp_mask selects none of Ex..Hz");
  if (nfreq < 1 || nfreq > kDftMaxFreq) return fail(HFDTD_EINVAL, "1 to 16 frequencies per DFT");
  if (mode != HFDTD_DFT_DIRECT && mode != HFDTD_DFT_BUFFERED) return fail(HFDTD_EINVAL, "bad mode");
  if (accum != HFDTD_ACC_REAL && accum != HFDTD_ACC_FP64 && accum != HFDTD_ACC_KAHAN)
    return fail(HFDTD_EINVAL, "bad accumulator mode");
  if (mode == HFDTD_DFT_BUFFERED && (K < 2 || K > 256 || (K & 1)))
    return fail(HFDTD_EINVAL, "K must be even, 2..256");
  if (mode == HFDTD_DFT_BUFFERED && e->fp64)
    return fail(HFDTD_EUNSUPPORTED, "the buffered (FP32 matrix-core) DFT needs an FP32 engine");
  hfdtd_dft *d = new hfdtd_dft();
  d->e = e;
  Plane &pl = d->pl;
  pl.ncomp = 0;
  for (int c = 0; c < 6; ++c)
    if (comp_mask & (1 << c)) pl.vol[pl.ncomp++] = c;
  const long long nynz = (long long)e->ny * e->nz;
  if (axis == 0) {
    pl.n0 = e->ny; pl.n1 = e->nz; pl.base = index * nynz; pl.sa = e->nz; pl.sb = 1;
  } else if (axis == 1) {
    pl.n0 = e->nx; pl.n1 = e->nz; pl.base = (long long)index * e->nz; pl.sa = nynz; pl.sb = 1;
  } else {
    pl.n0 = e->nx; pl.n1 = e->ny; pl.base = index; pl.sa = nynz; pl.sb = e->nz;
  }
  d->P = (long long)pl.ncomp * pl.n0 * pl.n1;
  d->Ppad = (d->P + 31) / 32 * 32;
  d->F = nfreq;
  d->K = K;
  d->mode = mode;
  d->accum = accum;
  d->acc_elem = (accum == HFDTD_ACC_FP64 || e->fp64) ? 8 : 4;
  if (e->fp64 && accum != HFDTD_ACC_FP64) d->accum = HFDTD_ACC_FP64;  // FP64 engines sum in FP64
  const size_t nacc = (size_t)d->P * 2 * nfreq;
  hipError_t err = hipSetDevice(e->device);
  if (err == hipSuccess) err = hipMalloc(&d->acc, nacc * d->acc_elem);
  if (err == hipSuccess) err = hipMemsetAsync(d->acc, 0, nacc * d->acc_elem, e->stream);
  if (err == hipSuccess && d->accum == HFDTD_ACC_KAHAN) {
    err = hipMalloc((void **)&d->cmp, nacc * sizeof(float));
    if (err == hipSuccess) err = hipMemsetAsync(d->cmp, 0, nacc * sizeof(float), e->stream);
  }
  if (err == hipSuccess && mode == HFDTD_DFT_BUFFERED) {
    err = hipMalloc((void **)&d->buf, (size_t)K * d->Ppad * sizeof(float));
    if (err == hipSuccess)
      err = hipMemsetAsync(d->buf, 0, (size_t)K * d->Ppad * sizeof(float), e->stream);
    if (err == hipSuccess) err = hipMalloc((void **)&d->W, (size_t)K * 32 * sizeof(float));
    if (err == hipSuccess) err = hipMemsetAsync(d->W, 0, (size_t)K * 32 * sizeof(float), e->stream);
  }
  if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
  if (err != hipSuccess) {
    const int rc = hip_fail(err, "hfdtd_dft_plane_create");
    hfdtd_dft_destroy(d);
    return rc;
  }
  *out = d;
  return HFDTD_OK;
}

int hfdtd_dft_destroy(hfdtd_dft *d) {
  if (!d) return HFDTD_OK;
  (void)hipSetDevice(d->e->device);
  (void)hipStreamSynchronize(d->e->stream);
  for (void *p : {d->acc, (void *)d->cmp, (void *)d->buf, (void *)d->W})
    if (p) (void)hipFree(p);
  delete d;
  return HFDTD_OK;
}

int hfdtd_dft_accumulate(hfdtd_dft *d, const double *phase) {
  if (!d || !phase) return fail(HFDTD_EINVAL, "NULL argument");
  Engine *e = d->e;
  HF_SETTLE(e);
  HF_HIP(hipSetDevice(e->device));
  if (d->mode == HFDTD_DFT_DIRECT) {
    if (e->fp64) return direct_t<double, double, false>(d, phase);
    if (d->accum == HFDTD_ACC_FP64) return direct_t<float, double, false>(d, phase);
    if (d->accum == HFDTD_ACC_KAHAN) return direct_t<float, float, true>(d, phase);
    return direct_t<float, float, false>(d, phase);
  }
  PhaseRow ph;
  for (int j = 0; j < 2 * kDftMaxFreq; ++j) ph.v[j] = j < 2 * d->F ? (float)phase[j] : 0.0f;
  hipLaunchKernelGGL(k_dft_gather<float>, dim3((unsigned)std::max<long long>(1, ceil_div(d->P, 256))),
                     dim3(256), 0, e->stream, (const float *)e->field(0),
                     (long long)e->vstride, d->pl, d->P, d->buf + (size_t)d->k * d->Ppad, ph,
                     d->W + (size_t)d->k * 32);
  HF_HIP(hipGetLastError());
  if (++d->k == d->K) return dft_flush(d);
  return HFDTD_OK;
}

int hfdtd_dft_flush(hfdtd_dft *d) {
  if (!d) return fail(HFDTD_EINVAL, "NULL argument");
  return dft_flush(d);
}

int hfdtd_dft_reset(hfdtd_dft *d) {
  if (!d) return fail(HFDTD_EINVAL, "NULL argument");
  HF_HIP(hipSetDevice(d->e->device));
  const size_t nacc = (size_t)d->P * 2 * d->F;
  HF_HIP(hipMemsetAsync(d->acc, 0, nacc * d->acc_elem, d->e->stream));
  if (d->cmp) HF_HIP(hipMemsetAsync(d->cmp, 0, nacc * sizeof(float), d->e->stream));
  d->k = 0;
  return HFDTD_OK;
}

int hfdtd_dft_shape(const hfdtd_dft *d, int *ncomp, int *n0, int *n1, int *nfreq) {
  if (!d) return fail(HFDTD_EINVAL, "NULL argument");
  if (ncomp) *ncomp = d->pl.ncomp;
  if (n0) *n0 = d->pl.n0;
  if (n1) *n1 = d->pl.n1;
  if (nfreq) *nfreq = d->F;
  return HFDTD_OK;
}

int hfdtd_dft_read(hfdtd_dft *d, double *out) {
  if (!d || !out) return fail(HFDTD_EINVAL, "NULL argument");
  int rc = dft_flush(d);
  if (rc) return rc;
  HF_HIP(hipSetDevice(d->e->device));
  const size_t nacc = (size_t)d->P * 2 * d->F;
  if (d->acc_elem == 8) {
    HF_HIP(hipMemcpyAsync(out, d->acc, nacc * 8, hipMemcpyDeviceToHost, d->e->stream));
    HF_HIP(hipStreamSynchronize(d->e->stream));
    return HFDTD_OK;
  }
  std::vector<float> tmp(nacc), cmp;
  HF_HIP(hipMemcpyAsync(tmp.data(), d->acc, nacc * 4, hipMemcpyDeviceToHost, d->e->stream));
  if (d->cmp) {
    cmp.resize(nacc);
    HF_HIP(hipMemcpyAsync(cmp.data(), d->cmp, nacc * 4, hipMemcpyDeviceToHost, d->e->stream));
  }
  HF_HIP(hipStreamSynchronize(d->e->stream));
  // compensated sums: the true sum is acc - cmp
  for (size_t q = 0; q < nacc; ++q) out[q] = (double)tmp[q] - (cmp.empty() ? 0.0 : (double)cmp[q]);
  return HFDTD_OK;
}

}  // extern "C"
