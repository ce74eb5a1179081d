// Engine lifetime, device memory, materials and host-side CPML profiles.
//
// Replaces the allocation / upload half of OpenCLFDTD (engine.py:99-334) and
// the host coefficient helpers cpml.py:84-124 and materials.py:28-89.  All
// state lives in device memory owned by the engine; the host only borrows
// pointers for the duration of a call.

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "common.h"
#include "fdtd_core.h"

namespace hfdtd {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

int hip_fail(hipError_t err, const char *what) {
  g_last_error = std::string("HIP error ") + hipGetErrorName(err) + " (" +
                 hipGetErrorString(err) + ") in " + what;
  return err == hipErrorOutOfMemory || err == hipErrorMemoryAllocation ? HFDTD_ENOMEM
                                                                          : HFDTD_EHIP;
}

// ── host-side CPML profile (cpml.py:84-124 restated) ─────────────────────
//
// Depth xi into the layer is 1 at the outer wall: lo side (npml-i-off)/npml,
// hi side (i+1-off)/npml; sigma = sigma_max xi^3, kappa = 1 + 14 xi^3,
// alpha = alpha_max (1 - xi); b = exp(-(sigma/kappa + alpha) dt/eps0) is
// rounded to T before c = sigma/kappa (b-1)/(sigma + kappa alpha)/dl uses it.
template <typename T>
static void cpml_profile(int n, int npml, double dl, double dt, double off, T *b, T *c, T *kap) {
  for (int i = 0; i < n; ++i) {
    b[i] = T(1);
    c[i] = T(0);
    kap[i] = T(1);
  }
  if (npml <= 0 || n <= 0) return;
  const double eta0 = std::sqrt(kMU0 / kEPS0);
  const int m = 3;
  const double sig_max = 0.8 * (m + 1) / (eta0 * dl);
  const double alpha_max = 0.05 / eta0;
  const double kappa_max = 15.0;
  for (int i = 0; i < npml; ++i) {
    for (int side = 0; side < 2; ++side) {
      const bool lo = side == 0;
      const int idx = lo ? i : n - npml + i;
      if (idx < 0 || idx >= n) continue;
      double xi = lo ? (npml - i - off) / npml : (i + 1 - off) / npml;
      if (xi < 0.0) xi = 0.0;
      const double xm = std::pow(xi, double(m));
      const double sig = sig_max * xm;
      const double kp = 1.0 + (kappa_max - 1.0) * xm;
      const double al = alpha_max * (1.0 - xi);
      const double decay = (sig / kp + al) * dt / kEPS0;
      const T bt = T(std::exp(-decay));
      b[idx] = bt;
      const double den = sig + kp * al;
      c[idx] = den == 0.0 ? T(0) : T(sig / kp * (double(bt) - 1.0) / den / dl);
      kap[idx] = T(kp);
    }
  }
}

// ── device helpers ────────────────────────────────────────────────────────

template <typename T>
__global__ void k_fill(T *p, size_t n, T v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

template <typename T>
static int fill(Engine *e, void *p, size_t n, T v) {
  if (n == 0) return HFDTD_OK;
  unsigned blocks = ceil_div(n, 256);
  if (blocks > 65536u * 4u) blocks = 65536u * 4u;
  hipLaunchKernelGGL(k_fill<T>, dim3(blocks), dim3(256), 0, e->stream, (T *)p, n, v);
  HF_HIP(hipGetLastError());
  return HFDTD_OK;
}

static int fill_value(Engine *e, void *p, size_t n, double v) {
  return e->fp64 ? fill<double>(e, p, n, v) : fill<float>(e, p, n, (float)v);
}

// Edge-averaged material coefficients (materials.py:28-89 restated), one
// thread per cell, FP64 arithmetic, stored in T.  The edge of component c
// averages cells idx and idx+stride_c; the last plane keeps the cell value.
template <typename T, typename S>
__global__ void k_materials(int nx, int ny, int nz, const S *__restrict__ eps,
                            const S *__restrict__ sig, double dt, double eps0, T *__restrict__ ce_x,
                            T *__restrict__ ce_y, T *__restrict__ ce_z, T *__restrict__ ca_x,
                            T *__restrict__ ca_y, T *__restrict__ ca_z) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  const int i = blockIdx.z;
  if (k >= nz) return;
  const size_t sxy = (size_t)ny * nz;
  const size_t idx = (size_t)i * sxy + (size_t)j * nz + k;
  const double e0 = (double)eps[idx];
  const double ex = i < nx - 1 ? 0.5 * (e0 + (double)eps[idx + sxy]) : e0;
  const double ey = j < ny - 1 ? 0.5 * (e0 + (double)eps[idx + nz]) : e0;
  const double ez = k < nz - 1 ? 0.5 * (e0 + (double)eps[idx + 1]) : e0;
  if (sig == nullptr) {
    const double scale = dt / eps0;
    ce_x[idx] = T(scale / ex);
    ce_y[idx] = T(scale / ey);
    ce_z[idx] = T(scale / ez);
    ca_x[idx] = T(1);
    ca_y[idx] = T(1);
    ca_z[idx] = T(1);
    return;
  }
  const double s0 = (double)sig[idx];
  const double sx = i < nx - 1 ? 0.5 * (s0 + (double)sig[idx + sxy]) : s0;
  const double sy = j < ny - 1 ? 0.5 * (s0 + (double)sig[idx + nz]) : s0;
  const double sz = k < nz - 1 ? 0.5 * (s0 + (double)sig[idx + 1]) : s0;
  const double ee[3] = {ex, ey, ez};
  const double ss[3] = {sx, sy, sz};
  T *cb_out[3] = {ce_x, ce_y, ce_z};
  T *ca_out[3] = {ca_x, ca_y, ca_z};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double loss = ss[c] * dt / (2.0 * eps0 * ee[c]);
    const double den = 1.0 + loss;
    ca_out[c][idx] = T((1.0 - loss) / den);
    cb_out[c][idx] = T((dt / (eps0 * ee[c])) / den);
  }
}

static int dmalloc(Engine *e, void **p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) return HFDTD_OK;
  hipError_t err = hipMalloc(p, bytes);
  if (err != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    char buf[160];
    std::snprintf(buf, sizeof buf, "device allocation of %.3f GB failed (%s)",
                  bytes / 1073741824.0, hipGetErrorString(err));
    return fail(HFDTD_ENOMEM, buf);
  }
  e->bytes += bytes;
  return HFDTD_OK;
}

static void release(Engine *e) {
  if (e->device >= 0) (void)hipSetDevice(e->device);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  free_gather(e);
  if (e->prun.live) {  // the engine's own slabs back in place (their contents no longer matter)
    for (int p = 0; p < 2; ++p) {
      e->state[p] = e->prun.state[p];
      e->psi[p] = e->prun.psi[p];
    }
    e->coef = e->prun.coef;
    e->mat = e->prun.mat;
    e->epsv = e->prun.epsv;
    e->vstride = e->prun.vstride;
    e->sk = e->nz;
    e->prun.live = false;
  }
  if (e->prun.buf) (void)hipFree(e->prun.buf);
  if (e->prun.d_jobs) (void)hipFree(e->prun.d_jobs);
  e->prun.buf = nullptr;
  e->prun.d_jobs = nullptr;
  auto fr = [](void *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
  };
  fr(e->state[0]);
  fr(e->state[1]);
  fr(e->psi[0]);
  fr(e->psi[1]);
  fr(e->coef);
  fr(e->prof);
  fr(e->shadow);
  fr(e->mat);
  {
    void *t = e->mtab;
    fr(t);
    e->mtab = nullptr;
  }
  e->mat_ok = false;
  fr(e->epsv);
  e->eps_ok = false;
  {
    void *t = e->zrec.buf;
    fr(t);
    e->zrec = ZRec{};
  }
  for (hipEvent_t ev : e->events) (void)hipEventDestroy(ev);
  e->events.clear();
  e->events_used = 0;
  if (e->stream) (void)hipStreamDestroy(e->stream);
  e->stream = nullptr;
}

int timing_events(Engine *e, hipEvent_t &start, hipEvent_t &stop) {
  start = stop = nullptr;
  if (!e->timing) return HFDTD_OK;
  if (e->timing_launches++ % e->timing_every != 0) return HFDTD_OK;
  while (e->events.size() < e->events_used + 2) {
    hipEvent_t ev;
    HF_HIP(hipEventCreate(&ev));
    e->events.push_back(ev);
  }
  start = e->events[e->events_used++];
  stop = e->events[e->events_used++];
  return HFDTD_OK;
}

int ensure_alt(Engine *e) {
  const int other = e->cur ^ 1;
  if (e->state[other]) return HFDTD_OK;
  int rc = dmalloc(e, &e->state[other], 6 * e->vstride * e->elem);
  if (rc) return rc;
  return dmalloc(e, &e->psi[other], (e->psi_off[PSI_COUNT] + 64) * e->elem);
}

// Material index of a lossless FP32 grid (march sweep): the distinct eps_r
// values, at most kMatMax of them, a byte per cell naming its value, and
// for every pair (a, b) the ce of an edge between cells of eps a and b,
// computed as k_materials computes it (FP64, then rounded), so the table
// holds exactly the values of the ce volumes.  Grids with more values, FP64
// or lossy grids keep mat_ok false and the sweep reads the ce volumes.
template <typename S>
static int build_material_index(Engine *e, const S *eps) {
  e->mat_ok = false;
  // values keyed by their bit patterns (+0 and -0 stay apart), NaN refused
  auto bits = [](double v) {
    uint64_t u;
    std::memcpy(&u, &v, sizeof u);
    return u;
  };
  std::vector<uint64_t> keys;  // sorted distinct patterns
  uint64_t last = 0;
  bool have = false;
  for (size_t c = 0; c < e->ncell; ++c) {  // the distinct values (runs are common)
    const double v = (double)eps[c];
    const uint64_t u = bits(v);
    if (have && u == last) continue;
    // NaN, or a value whose doubling overflows (the last plane's ce uses the
    // cell value itself, the table 0.5 * (v + v)): no exact index
    if (v != v || std::fabs(v) > 8.0e307) return HFDTD_OK;
    auto it = std::lower_bound(keys.begin(), keys.end(), u);
    if (it == keys.end() || *it != u) {
      if ((int)keys.size() == kMatMax) return HFDTD_OK;  // too many materials
      keys.insert(it, u);
    }
    last = u;
    have = true;
  }
  std::vector<unsigned char> idx(e->ncell);
  int last_i = 0;
  have = false;
  for (size_t c = 0; c < e->ncell; ++c) {
    const uint64_t u = bits((double)eps[c]);
    if (!have || u != last) {
      last_i = (int)(std::lower_bound(keys.begin(), keys.end(), u) - keys.begin());
      last = u;
      have = true;
    }
    idx[c] = (unsigned char)last_i;
  }
  const int m = (int)keys.size();
  std::vector<double> vals(m);
  for (int a = 0; a < m; ++a) std::memcpy(&vals[a], &keys[a], sizeof(double));
  std::vector<float> tab((size_t)m * m);
  const double scale = e->dt / kEPS0;
  for (int a = 0; a < m; ++a)
    for (int b = 0; b < m; ++b) tab[(size_t)a * m + b] = (float)(scale / (0.5 * (vals[a] + vals[b])));
  if (!e->mat) {
    int rc = dmalloc(e, &e->mat, e->ncell + 64);
    if (rc) return rc;
  }
  if (!e->mtab) {
    void *t = nullptr;
    int rc = dmalloc(e, &t, (size_t)kMatMax * kMatMax * sizeof(float));
    if (rc) return rc;
    e->mtab = (float *)t;
  }
  HF_HIP(hipMemcpyAsync(e->mat, idx.data(), e->ncell, hipMemcpyHostToDevice, e->stream));
  HF_HIP(hipMemcpyAsync(e->mtab, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice,
                        e->stream));
  HF_HIP(hipStreamSynchronize(e->stream));
  e->nmat = m;
  e->mat_ok = true;
  return HFDTD_OK;
}

// The index of the default medium (eps_r = 1 everywhere): every cell is
// material 0, whose table entry dt/eps0 / (0.5 (1 + 1)) is the default ce
// volumes' value dt/eps0 exactly, so the march sweep reads one byte per cell
// before any set_epsilon (FP32 engines).
static int uniform_material_index(Engine *e) {
  if (!e->mat) {
    int rc = dmalloc(e, &e->mat, e->ncell + 64);
    if (rc) return rc;
  }
  if (!e->mtab) {
    void *t = nullptr;
    int rc = dmalloc(e, &t, (size_t)kMatMax * kMatMax * sizeof(float));
    if (rc) return rc;
    e->mtab = (float *)t;
  }
  const float ce = (float)(e->dt / kEPS0 / (0.5 * (1.0 + 1.0)));
  HF_HIP(hipMemsetAsync(e->mat, 0, e->ncell + 64, e->stream));
  HF_HIP(hipMemcpyAsync(e->mtab, &ce, sizeof(float), hipMemcpyHostToDevice, e->stream));
  HF_HIP(hipStreamSynchronize(e->stream));
  e->nmat = 1;
  e->mat_ok = true;
  return HFDTD_OK;
}

// The eps_r volume of a lossless FP32 grid without a material index: the
// march sweep recomputes ce from it.  Only when every value is exact in
// FP32 (always for FP32 input), so ce from it equals the ce volumes, which
// were made from the values as given; d_eps holds them on the device.  Best
// effort: without room for the volume the sweep reads the ce volumes.
template <typename S>
static void build_eps_volume(Engine *e, const S *host, const void *d_eps) {
  e->eps_ok = false;
  std::vector<float> conv;
  if constexpr (sizeof(S) == 8) {
    conv.resize(e->ncell);
    for (size_t c = 0; c < e->ncell; ++c) {
      const float f = (float)host[c];
      if ((double)f != host[c]) return;  // not exact in FP32 (NaN included)
      conv[c] = f;
    }
  }
  if (((double)e->ncell + 64.0) * 4.0 >= 2147483648.0) return;  // one buffer resource reaches it
  if (!e->epsv && dmalloc(e, &e->epsv, (e->ncell + 64) * sizeof(float)) != HFDTD_OK) {
    (void)hipGetLastError();
    return;
  }
  hipError_t err =
      sizeof(S) == 8
          ? hipMemcpyAsync(e->epsv, conv.data(), e->ncell * 4, hipMemcpyHostToDevice, e->stream)
          : hipMemcpyAsync(e->epsv, d_eps, e->ncell * 4, hipMemcpyDeviceToDevice, e->stream);
  if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
  if (err != hipSuccess) {
    (void)hipGetLastError();
    return;
  }
  e->eps_ok = true;
}

// ── row-padded runs ────────────────────────────────────────────────────────
// Where nz does not fill whole 128-byte lines, every row of a march tile
// starts mid-line and its 64-column segment touches three lines instead of
// two.  A row-padded run steps a copy whose rows are padded (the sweep
// reads the row stride from Geo::sk); the engine's state[], psi[], the ce
// source the sweep reads, vstride and sk point at the copy until it ends.

// Row copies of the state and psi of parity p between the engine's own
// slabs (dense rows) and the copy of geometry gp at buf (padded rows).
static void padded_jobs(const Engine *e, const Geo &gp, char *buf, size_t slab, size_t psil, int p,
                        bool to_copy, std::vector<PadJob> &jobs) {
  const int F = (int)(e->elem / 4);  // floats per element
  auto add = [&](const void *own, void *cp, long long rows, long long w, long long ow,
                 long long cw) {
    if (rows <= 0 || w <= 0) return;
    float *o = (float *)const_cast<void *>(own), *c = (float *)cp;
    jobs.push_back(to_copy ? PadJob{o, c, rows, (int)(w * F), (int)(ow * F), (int)(cw * F)}
                           : PadJob{c, o, rows, (int)(w * F), (int)(cw * F), (int)(ow * F)});
  };
  const long long rows = (long long)e->nx * e->ny, n2 = 2ll * e->npml;
  const void *own_state = e->prun.live ? e->prun.state[p] : e->state[p];
  const void *own_psi = e->prun.live ? e->prun.psi[p] : e->psi[p];
  const size_t ovs = e->prun.live ? e->prun.vstride : e->vstride;
  for (int c = 0; c < 6; ++c)
    add((const char *)own_state + (size_t)c * ovs * e->elem,
        buf + ((size_t)p * slab + (size_t)c * gp.vs) * e->elem, rows, e->nz, e->nz, gp.sk);
  for (int t = 0; t < PSI_COUNT; ++t) {
    const void *o = (const char *)own_psi + e->psi_off[t] * e->elem;
    void *c = buf + (2 * slab + (size_t)p * psil + psi_slot(gp, t)) * e->elem;
    const int u = t % 6;
    if (u < 2) add(o, c, n2 * e->ny, e->nz, e->nz, gp.sk);
    else if (u < 4) add(o, c, (long long)e->nx * n2, e->nz, e->nz, gp.sk);
    else add(o, c, 1, rows * n2, rows * n2, rows * n2);
  }
}

static size_t padded_psi_len(const Geo &gp) {
  return ((size_t)psi_slot(gp, PSI_COUNT) + 64 + 63) / 64 * 64;
}

int end_padded_run(Engine *e) {
  if (!e->prun.live) return HFDTD_OK;
  HF_HIP(hipSetDevice(e->device));
  const Geo gp = geo_of(e);
  const size_t slab = 6 * e->vstride, psil = padded_psi_len(gp);
  std::vector<PadJob> jobs;
  padded_jobs(e, gp, (char *)e->prun.buf, slab, psil, e->cur, false, jobs);
  int rc = run_pad_jobs(e->stream, jobs, e->prun.d_jobs, e->prun.jobs_cap);
  if (rc) return rc;
  for (int p = 0; p < 2; ++p) {
    e->state[p] = e->prun.state[p];
    e->psi[p] = e->prun.psi[p];
  }
  e->coef = e->prun.coef;
  e->mat = e->prun.mat;
  e->epsv = e->prun.epsv;
  e->vstride = e->prun.vstride;
  e->sk = e->nz;
  e->prun.live = false;
  e->touch();
  HF_HIP(hipStreamSynchronize(e->stream));
  return HFDTD_OK;
}

int settle(Engine *e) {
  int rc = flush_gather(e);
  if (rc) return rc;
  return end_padded_run(e);
}

// Start a row-padded run (see above); a no-op where the rows already fill
// whole lines.
static int begin_padded_run(Engine *e) {
  if (e->prun.live) return HFDTD_OK;
  const size_t per_line = 128 / e->elem;
  if (e->nz % per_line == 0) return HFDTD_OK;
  if (e->view_shadow)
    return fail(HFDTD_EUNSUPPORTED, "a row-padded run does not take the pair sweep's view");
  HF_HIP(hipSetDevice(e->device));
  Geo gp = geo_of(e);
  gp.sk = (int)((e->nz + per_line - 1) / per_line * per_line);
  gp.vs = ((long long)e->nx * e->ny * gp.sk + 64 + 63) / 64 * 64;
  const size_t slab = 6 * (size_t)gp.vs, psil = padded_psi_len(gp);
  const int ce = e->ce_source();
  const size_t cells_p = (size_t)e->nx * e->ny * gp.sk + 64;
  const size_t ce_elems = ce == kCeVolumes ? slab
                          : ce == kCeEps   ? (cells_p * 4 + e->elem - 1) / e->elem
                                           : (cells_p + e->elem - 1) / e->elem;
  // the march sweep's 32-bit offsets: slabs below 4 GB, or volumes below
  // 2 GB on per-volume addressing (FP32); psi slabs below 2 GB
  const double E = (double)e->elem;
  if ((6.0 * gp.vs * E >= 4294967296.0 && (e->fp64 || gp.vs * E >= 2147483648.0)) ||
      psil * E >= 2147483648.0)
    return fail(HFDTD_EUNSUPPORTED, "the padded copy exceeds the march sweep's slab limits");
  const size_t bytes = (2 * slab + 2 * psil + (ce_elems + 63) / 64 * 64) * e->elem;
  if (bytes > e->prun.bytes) {
    if (e->prun.buf) HF_HIP(hipFree(e->prun.buf));
    e->prun.buf = nullptr;
    e->prun.bytes = 0;
    const hipError_t err = hipMalloc(&e->prun.buf, bytes);
    if (err != hipSuccess) {
      (void)hipGetLastError();
      e->prun.buf = nullptr;
      return fail(HFDTD_ENOMEM, "no device memory for the row-padded copy");
    }
    HF_HIP(hipMemsetAsync(e->prun.buf, 0, bytes, e->stream));  // pad columns stay zero
    e->prun.bytes = bytes;
  }
  int rc = ensure_alt(e);
  if (rc) return rc;
  char *buf = (char *)e->prun.buf;
  char *cep = buf + (2 * slab + 2 * psil) * e->elem;
  std::vector<PadJob> jobs;
  padded_jobs(e, gp, buf, slab, psil, e->cur, true, jobs);
  const long long rows = (long long)e->nx * e->ny;
  if (ce == kCeVolumes) {
    const int F = (int)(e->elem / 4);
    for (int c = 0; c < 6; ++c)
      jobs.push_back(PadJob{(const float *)e->coef_vol(c), (float *)(cep + (size_t)c * gp.vs * e->elem),
                            rows, e->nz * F, e->nz * F, gp.sk * F});
  } else if (ce == kCeEps) {
    jobs.push_back(PadJob{(const float *)e->epsv, (float *)cep, rows, e->nz, e->nz, gp.sk});
  } else {
    HF_HIP(hipMemcpy2DAsync(cep, (size_t)gp.sk, e->mat, (size_t)e->nz, (size_t)e->nz, (size_t)rows,
                            hipMemcpyDeviceToDevice, e->stream));
  }
  if ((rc = run_pad_jobs(e->stream, jobs, e->prun.d_jobs, e->prun.jobs_cap))) return rc;
  for (int p = 0; p < 2; ++p) {
    e->prun.state[p] = e->state[p];
    e->prun.psi[p] = e->psi[p];
    e->state[p] = buf + (size_t)p * slab * e->elem;
    e->psi[p] = buf + (2 * slab + (size_t)p * psil) * e->elem;
  }
  e->prun.coef = e->coef;
  e->prun.mat = e->mat;
  e->prun.epsv = e->epsv;
  e->prun.vstride = e->vstride;
  if (ce == kCeVolumes) e->coef = cep;
  else if (ce == kCeEps) e->epsv = cep;
  else e->mat = cep;
  e->vstride = (size_t)gp.vs;
  e->sk = gp.sk;
  e->prun.live = true;
  e->touch();
  return HFDTD_OK;
}

}  // namespace hfdtd

using namespace hfdtd;

extern "C" {

const char *hfdtd_last_error(void) { return g_last_error.c_str(); }

int hfdtd_version(void) { return HFDTD_ABI_VERSION; }

int hfdtd_device_count(int *count) {
  if (!count) return fail(HFDTD_EINVAL, "count is NULL");
  *count = 0;
  hipError_t err = hipGetDeviceCount(count);
  if (err != hipSuccess) {
    (void)hipGetLastError();
    *count = 0;
    return fail(HFDTD_ENODEV, std::string("no HIP device: ") + hipGetErrorString(err));
  }
  return HFDTD_OK;
}

int hfdtd_device_info(int device, uint64_t *total_bytes, uint64_t *free_bytes, char *name,
                      int name_len) {
  int n = 0;
  int rc = hfdtd_device_count(&n);
  if (rc) return rc;
  if (device < 0 || device >= n) return fail(HFDTD_ENODEV, "device index out of range");
  HF_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  HF_HIP(hipGetDeviceProperties(&prop, device));
  size_t fr = 0, tot = 0;
  HF_HIP(hipMemGetInfo(&fr, &tot));
  if (total_bytes) *total_bytes = tot;
  if (free_bytes) *free_bytes = fr;
  if (name && name_len > 0) {
    // the marketing name can be empty under some drivers: fall back to the
    // ISA name and compute-unit count
    if (prop.name[0])
      std::snprintf(name, (size_t)name_len, "%s (%s)", prop.name, prop.gcnArchName);
    else
      std::snprintf(name, (size_t)name_len, "%s, %d CUs", prop.gcnArchName,
                    prop.multiProcessorCount);
  }
  return HFDTD_OK;
}

int hfdtd_cpml_profile(int n, int npml, double dl, double dt, double offset, int fp64, void *b,
                       void *c, void *kappa) {
  if (n < 0 || npml < 0 || !(dl > 0) || !(dt > 0) || !b || !c || !kappa)
    return fail(HFDTD_EINVAL, "hfdtd_cpml_profile: bad argument");
  if (fp64)
    cpml_profile<double>(n, npml, dl, dt, offset, (double *)b, (double *)c, (double *)kappa);
  else
    cpml_profile<float>(n, npml, dl, dt, offset, (float *)b, (float *)c, (float *)kappa);
  return HFDTD_OK;
}

int hfdtd_create(int nx, int ny, int nz, int npml, double dl, double dt, int fp64, int device,
                 hfdtd_engine **out) {
  if (!out) return fail(HFDTD_EINVAL, "out is NULL");
  *out = nullptr;
  if (nx < 1 || ny < 1 || nz < 1 || npml < 0)
    return fail(HFDTD_EINVAL, "grid extents must be >= 1 and npml >= 0");
  if (!(dl > 0.0) || !(dt > 0.0)) return fail(HFDTD_EINVAL, "dl and dt must be positive");
  if ((double)nx * ny * nz >= 2147483647.0)
    return fail(HFDTD_EUNSUPPORTED, "grid exceeds 2^31 cells");
  if (npml > 0 && 8.0 * npml * ((double)ny * nz + (double)nx * nz + (double)nx * ny) >= 2147483647.0)
    return fail(HFDTD_EUNSUPPORTED, "psi stripes exceed 2^31 entries");
  int ndev = 0;
  int rc = hfdtd_device_count(&ndev);
  if (rc) return rc;
  if (device < 0 || device >= ndev) return fail(HFDTD_ENODEV, "device index out of range");

  hfdtd_engine *h = new hfdtd_engine();
  Engine *e = &h->s;
  e->nx = nx;
  e->ny = ny;
  e->nz = nz;
  e->sk = nz;
  e->npml = npml;
  e->fp64 = fp64 ? 1 : 0;
  e->device = device;
  e->dl = dl;
  e->dt = dt;
  e->elem = fp64 ? 8 : 4;
  e->ncell = (size_t)nx * ny * nz;
  // >= 64 elements of padding after every volume: the fused sweep's lanes
  // outside the grid store into it
  e->vstride = (e->ncell + 64 + 63) / 64 * 64;
  // folded constants, evaluated in FP64 as engine.py:543-544, 634 do
  e->dtm_dl = dt / (kMU0 * dl);
  e->dtm = dt / kMU0;
  e->inv_dl = 1.0 / dl;
  {
    size_t len[3] = {0, 0, 0};
    if (npml > 0) {
      len[0] = (size_t)2 * npml * ny * nz;
      len[1] = (size_t)nx * 2 * npml * nz;
      len[2] = (size_t)nx * ny * 2 * npml;
    }
    e->psi_off[0] = 0;
    for (int q = 0; q < PSI_COUNT; ++q) e->psi_off[q + 1] = e->psi_off[q] + len[kPsiAxis[q]];
  }

  auto bail = [&](int code) {
    release(e);
    delete h;
    return code;
  };
  hipError_t err = hipSetDevice(device);
  if (err != hipSuccess) return bail(hip_fail(err, "hipSetDevice"));
  err = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
  if (err != hipSuccess) return bail(hip_fail(err, "hipStreamCreateWithFlags"));
  {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
        cus > 0)
      e->cus = cus;
  }

  if ((rc = dmalloc(e, &e->state[0], 6 * e->vstride * e->elem))) return bail(rc);
  if ((rc = dmalloc(e, &e->coef, 6 * e->vstride * e->elem))) return bail(rc);
  // + 64 elements of padding: the fused sweep's lanes outside a stripe
  // store their psi there
  if ((rc = dmalloc(e, &e->psi[0], (e->psi_off[PSI_COUNT] + 64) * e->elem))) return bail(rc);
  if ((rc = dmalloc(e, &e->prof, 2 * 3 * e->prof_len() * e->elem))) return bail(rc);

  // CPML profiles, H nodes offset 1/2 and E nodes offset 0 (engine.py:262-283),
  // laid out [stagger][b, c, ik][x | y | z]
  {
    const int dims[3] = {nx, ny, nz};
    const size_t L = e->prof_len();
    std::vector<unsigned char> host(2 * 3 * L * e->elem);
    size_t axoff = 0;
    for (int a = 0; a < 3; ++a) {
      const int n = dims[a];
      for (int stag = 0; stag < 2; ++stag) {
        const double off = stag == 0 ? 0.0 : 0.5;
        unsigned char *pb = host.data() + ((size_t)(stag * 3 + 0) * L + axoff) * e->elem;
        unsigned char *pc = host.data() + ((size_t)(stag * 3 + 1) * L + axoff) * e->elem;
        unsigned char *pk = host.data() + ((size_t)(stag * 3 + 2) * L + axoff) * e->elem;
        if (fp64) {
          std::vector<double> kd(n);
          cpml_profile<double>(n, npml, dl, dt, off, (double *)pb, (double *)pc, kd.data());
          for (int i = 0; i < n; ++i) ((double *)pk)[i] = 1.0 / (kd[i] * dl);
        } else {
          std::vector<float> kf(n);
          cpml_profile<float>(n, npml, dl, dt, off, (float *)pb, (float *)pc, kf.data());
          // 1/(kappa*dl) in FP64 from the stored kappa, then rounded (engine.py:264-266)
          for (int i = 0; i < n; ++i) ((float *)pk)[i] = (float)(1.0 / ((double)kf[i] * dl));
        }
      }
      axoff += n;
    }
    err = hipMemcpy(e->prof, host.data(), host.size(), hipMemcpyHostToDevice);
    if (err != hipSuccess) return bail(hip_fail(err, "hipMemcpy(cpml)"));
  }

  // defaults: fields 0, ce = dt/eps0, ca = 1 (engine.py:189-204)
  if ((rc = fill_value(e, e->state[0], 6 * e->vstride, 0.0))) return bail(rc);
  if ((rc = fill_value(e, e->coef, 3 * e->vstride, dt / kEPS0))) return bail(rc);
  if ((rc = fill_value(e, e->coef_vol(3), 3 * e->vstride, 1.0))) return bail(rc);
  if ((rc = fill_value(e, e->psi[0], e->psi_off[PSI_COUNT], 0.0))) return bail(rc);
  err = hipStreamSynchronize(e->stream);
  if (err != hipSuccess) return bail(hip_fail(err, "hipStreamSynchronize(create)"));
  if (!fp64 && (rc = uniform_material_index(e))) return bail(rc);
  *out = h;
  return HFDTD_OK;
}

int hfdtd_destroy(hfdtd_engine *h) {
  if (!h) return HFDTD_OK;
  release(&h->s);
  delete h;
  return HFDTD_OK;
}

int hfdtd_memory_bytes(const hfdtd_engine *h, uint64_t *bytes) {
  if (!h || !bytes) return fail(HFDTD_EINVAL, "NULL argument");
  *bytes = h->s.bytes;
  return HFDTD_OK;
}

static void *volume_ptr(Engine *e, int which) {
  if (which >= HFDTD_EX && which <= HFDTD_HZ) return e->field(which);
  if (which >= HFDTD_CEX && which <= HFDTD_CAZ) return e->coef_vol(which - HFDTD_CEX);
  return nullptr;
}

int hfdtd_set_materials(hfdtd_engine *h, const void *eps_r, const void *sigma, int src_fp64) {
  if (!h || !eps_r) return fail(HFDTD_EINVAL, "NULL argument");
  Engine *e = &h->s;
  HF_SETTLE(e);
  HF_HIP(hipSetDevice(e->device));
  const size_t sb = e->ncell * (src_fp64 ? 8 : 4);
  void *d_eps = nullptr, *d_sig = nullptr;
  HF_HIP(hipMalloc(&d_eps, sb));
  if (sigma) {
    hipError_t err = hipMalloc(&d_sig, sb);
    if (err != hipSuccess) {
      (void)hipFree(d_eps);
      return hip_fail(err, "hipMalloc(sigma)");
    }
  }
  int rc = HFDTD_OK;
  hipError_t err = hipMemcpyAsync(d_eps, eps_r, sb, hipMemcpyHostToDevice, e->stream);
  if (err == hipSuccess && sigma)
    err = hipMemcpyAsync(d_sig, sigma, sb, hipMemcpyHostToDevice, e->stream);
  if (err == hipSuccess) {
    dim3 grid(ceil_div(e->nz, 64), e->ny, e->nx);
#define HF_MAT(T, S)                                                                         \
  hipLaunchKernelGGL((k_materials<T, S>), grid, dim3(64), 0, e->stream, e->nx, e->ny, e->nz, \
                     (const S *)d_eps, (const S *)d_sig, e->dt, kEPS0, (T *)e->coef_vol(0),   \
                     (T *)e->coef_vol(1), (T *)e->coef_vol(2), (T *)e->coef_vol(3),            \
                     (T *)e->coef_vol(4), (T *)e->coef_vol(5))
    if (e->fp64) {
      if (src_fp64) HF_MAT(double, double); else HF_MAT(double, float);
    } else {
      if (src_fp64) HF_MAT(float, double); else HF_MAT(float, float);
    }
#undef HF_MAT
    err = hipGetLastError();
  }
  if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
  if (err != hipSuccess) rc = hip_fail(err, "set_materials");
  e->lossless = sigma == nullptr;
  e->mat_ok = false;
  e->eps_ok = false;
  if (rc == HFDTD_OK && !e->fp64 && sigma == nullptr)
    rc = src_fp64 ? build_material_index(e, (const double *)eps_r)
                  : build_material_index(e, (const float *)eps_r);
  if (rc == HFDTD_OK && !e->fp64 && sigma == nullptr && !e->mat_ok) {
    if (src_fp64) build_eps_volume(e, (const double *)eps_r, d_eps);
    else build_eps_volume(e, (const float *)eps_r, d_eps);
  }
  (void)hipFree(d_eps);
  if (d_sig) (void)hipFree(d_sig);
  return rc;
}

int hfdtd_write_volume(hfdtd_engine *h, int which, const void *host) {
  if (!h || !host) return fail(HFDTD_EINVAL, "NULL argument");
  Engine *e = &h->s;
  HF_SETTLE(e);
  void *slot = volume_ptr(e, which);
  if (!slot) return fail(HFDTD_EINVAL, "volume id out of range");
  if (which >= HFDTD_CEX) e->mat_ok = e->eps_ok = false;  // coefficients no longer from eps_r
  HF_HIP(hipSetDevice(e->device));
  if (which <= HFDTD_HZ) e->touch();
  HF_HIP(hipMemcpyAsync(slot, host, e->ncell * e->elem, hipMemcpyHostToDevice, e->stream));
  HF_HIP(hipStreamSynchronize(e->stream));
  if (which >= HFDTD_CAX) e->lossless = false;
  return HFDTD_OK;
}

int hfdtd_read_volume(hfdtd_engine *h, int which, void *host) {
  if (!h || !host) return fail(HFDTD_EINVAL, "NULL argument");
  Engine *e = &h->s;
  HF_SETTLE(e);
  void *slot = volume_ptr(e, which);
  if (!slot) return fail(HFDTD_EINVAL, "volume id out of range");
  HF_HIP(hipSetDevice(e->device));
  HF_HIP(hipMemcpyAsync(host, slot, e->ncell * e->elem, hipMemcpyDeviceToHost, e->stream));
  HF_HIP(hipStreamSynchronize(e->stream));
  return HFDTD_OK;
}

int hfdtd_read_point(hfdtd_engine *h, int which, int64_t index, double *out) {
  if (!h || !out) return fail(HFDTD_EINVAL, "NULL argument");
  Engine *e = &h->s;
  HF_SETTLE(e);
  void *slot = volume_ptr(e, which);
  if (!slot) return fail(HFDTD_EINVAL, "volume id out of range");
  if (index < 0 || (size_t)index >= e->ncell) return fail(HFDTD_EINVAL, "index out of range");
  HF_HIP(hipSetDevice(e->device));
  unsigned char buf[8];
  HF_HIP(hipMemcpyAsync(buf, (const char *)slot + (size_t)index * e->elem, e->elem,
                        hipMemcpyDeviceToHost, e->stream));
  HF_HIP(hipStreamSynchronize(e->stream));
  if (e->fp64) {
    double v;
    std::memcpy(&v, buf, 8);
    *out = v;
  } else {
    float v;
    std::memcpy(&v, buf, 4);
    *out = v;
  }
  return HFDTD_OK;
}

int hfdtd_reset_fields(hfdtd_engine *h) {
  if (!h) return fail(HFDTD_EINVAL, "NULL argument");
  Engine *e = &h->s;
  HF_SETTLE(e);
  HF_HIP(hipSetDevice(e->device));
  e->touch();
  int rc;
  if ((rc = fill_value(e, e->state[e->cur], 6 * e->vstride, 0.0))) return rc;
  if ((rc = fill_value(e, e->psi[e->cur], e->psi_off[PSI_COUNT], 0.0))) return rc;
  HF_HIP(hipStreamSynchronize(e->stream));
  return HFDTD_OK;
}

int hfdtd_update_H(hfdtd_engine *h) {
  if (!h) return fail(HFDTD_EINVAL, "NULL argument");
  HF_SETTLE(&h->s);
  return launch_update_H(&h->s);
}

int hfdtd_update_E(hfdtd_engine *h) {
  if (!h) return fail(HFDTD_EINVAL, "NULL argument");
  HF_SETTLE(&h->s);
  return launch_update_E(&h->s);
}

int hfdtd_add_sources(hfdtd_engine *h, const hfdtd_source *src, int n) {
  if (!h || (n > 0 && !src) || n < 0) return fail(HFDTD_EINVAL, "bad source list");
  HF_SETTLE(&h->s);
  return launch_sources(&h->s, src, n);
}

int hfdtd_step(hfdtd_engine *h, const hfdtd_source *src, int n, int path) {
  if (!h || (n > 0 && !src) || n < 0) return fail(HFDTD_EINVAL, "bad source list");
  Engine *e = &h->s;
  // a row-padded run takes march steps whose sources the sweep applies, and
  // such a step takes a held gather into its launch
  if (path != HFDTD_PATH_MARCH || n > kMarchSources) HF_SETTLE(e);
  if (path == HFDTD_PATH_FUSED) return launch_fused_step(e, src, n);
  if (path == HFDTD_PATH_MARCH) return launch_march_step(e, src, n);
  if (path != HFDTD_PATH_SPLIT) return fail(HFDTD_EINVAL, "unknown step path");
  int rc = launch_update_H(e);
  if (rc) return rc;
  if (n > 0 && (rc = launch_sources(e, src, n))) return rc;
  return launch_update_E(e);
}

int hfdtd_step_pair(hfdtd_engine *h, const hfdtd_source *src_a, int na, const hfdtd_source *src_b,
                    int nb) {
  if (!h || na < 0 || nb < 0 || (na > 0 && !src_a) || (nb > 0 && !src_b))
    return fail(HFDTD_EINVAL, "bad source list");
  HF_SETTLE(&h->s);
  return launch_pair_steps(&h->s, src_a, na, src_b, nb);
}

int hfdtd_padded_begin(hfdtd_engine *h) {
  if (!h) return fail(HFDTD_EINVAL, "NULL argument");
  int rc = flush_gather(&h->s);
  if (rc) return rc;
  return begin_padded_run(&h->s);
}

int hfdtd_padded_end(hfdtd_engine *h) {
  if (!h) return fail(HFDTD_EINVAL, "NULL argument");
  return settle(&h->s);
}

int hfdtd_padded_live(const hfdtd_engine *h, int *live) {
  if (!h || !live) return fail(HFDTD_EINVAL, "NULL argument");
  *live = h->s.prun.live ? 1 : 0;
  return HFDTD_OK;
}

int hfdtd_set_watch(hfdtd_engine *h, const int *boxes, int n) {
  if (!h || (n > 0 && !boxes)) return fail(HFDTD_EINVAL, "NULL argument");
  if (n < 0 || n > 8) return fail(HFDTD_EINVAL, "at most 8 watched boxes");
  Engine *e = &h->s;
  HF_SETTLE(e);
  HF_HIP(hipSetDevice(e->device));
  const int hi[3] = {e->nx, e->ny, e->nz};
  for (int b = 0; b < n; ++b) {
    for (int a = 0; a < 3; ++a) {
      const int lo = std::min(std::max(boxes[7 * b + 2 * a], 0), hi[a]);
      e->watch_b[b][2 * a] = lo;
      e->watch_b[b][2 * a + 1] = std::min(std::max(boxes[7 * b + 2 * a + 1], lo), hi[a]);
    }
    const int c = boxes[7 * b + 6];
    e->watch_c[b] = c ? c & 63 : 63;
  }
  if (n > 0 && !e->shadow) {
    int rc = dmalloc(e, &e->shadow, 6 * e->vstride * e->elem);
    if (rc) return rc;
  }
  e->watch_n = n;
  return HFDTD_OK;
}

int hfdtd_set_view(hfdtd_engine *h, int shadow) {
  if (!h) return fail(HFDTD_EINVAL, "NULL argument");
  if (shadow && !h->s.shadow) return fail(HFDTD_EINVAL, "no watched boxes: no shadow slab");
  h->s.view_shadow = shadow != 0;
  h->s.touch();  // field() now reads another slab
  return HFDTD_OK;
}

int hfdtd_material_count(const hfdtd_engine *h, int *nmat) {
  if (!h || !nmat) return fail(HFDTD_EINVAL, "NULL argument");
  *nmat = h->s.mat_ok ? h->s.nmat : 0;
  return HFDTD_OK;
}

int hfdtd_ce_source(const hfdtd_engine *h, int *source) {
  if (!h || !source) return fail(HFDTD_EINVAL, "NULL argument");
  *source = h->s.ce_source();
  return HFDTD_OK;
}

int hfdtd_set_pair_planes(hfdtd_engine *h, int planes_i) {
  if (!h) return fail(HFDTD_EINVAL, "NULL argument");
  if (planes_i < 0 || planes_i > 1024) return fail(HFDTD_EINVAL, "planes_i must be 0..1024");
  h->s.pair_planes = planes_i ? planes_i : 24;
  return HFDTD_OK;
}

int hfdtd_synchronize(hfdtd_engine *h) {
  if (!h) return fail(HFDTD_EINVAL, "NULL argument");
  HF_HIP(hipSetDevice(h->s.device));
  int rc = flush_gather(&h->s);
  if (rc) return rc;
  HF_HIP(hipStreamSynchronize(h->s.stream));
  return HFDTD_OK;
}

int hfdtd_timing(hfdtd_engine *h, int enable) {
  if (!h) return fail(HFDTD_EINVAL, "NULL argument");
  if (enable < 0) return fail(HFDTD_EINVAL, "enable must be 0 (off) or n >= 1 (every n-th launch)");
  Engine *e = &h->s;
  e->timing = enable != 0;
  if (enable) {
    e->events_used = 0;
    e->timing_every = enable;
    e->timing_launches = 0;
  }
  return HFDTD_OK;
}

int hfdtd_timing_launches(const hfdtd_engine *h, long long *launches) {
  if (!h || !launches) return fail(HFDTD_EINVAL, "NULL argument");
  *launches = h->s.timing_launches;
  return HFDTD_OK;
}

int hfdtd_timing_read(hfdtd_engine *h, double *total_ms, int *launches) {
  if (!h || !total_ms || !launches) return fail(HFDTD_EINVAL, "NULL argument");
  Engine *e = &h->s;
  HF_HIP(hipSetDevice(e->device));
  HF_HIP(hipStreamSynchronize(e->stream));
  double tot = 0.0;
  const size_t n = e->events_used / 2;
  for (size_t q = 0; q < n; ++q) {
    float ms = 0.f;
    HF_HIP(hipEventElapsedTime(&ms, e->events[2 * q], e->events[2 * q + 1]));
    tot += ms;
  }
  *total_ms = tot;
  *launches = (int)n;
  return HFDTD_OK;
}

int hfdtd_set_split_schedule(hfdtd_engine *h, int run_blocks) {
  if (!h) return fail(HFDTD_EINVAL, "NULL argument");
  if (run_blocks < 0 || run_blocks > 4096) return fail(HFDTD_EINVAL, "run_blocks must be 0..4096");
  if (run_blocks) h->s.split_run = run_blocks;
  return HFDTD_OK;
}

int hfdtd_set_march_tiling(hfdtd_engine *h, int rows_per_wave, int planes_i, int run_blocks,
                           int nontemporal) {
  if (!h) return fail(HFDTD_EINVAL, "NULL argument");
  if (rows_per_wave < 0 || rows_per_wave > 2) return fail(HFDTD_EINVAL, "rows_per_wave must be 0..2");
  if (planes_i < -1 || planes_i > 4096) return fail(HFDTD_EINVAL, "planes_i must be -1..4096");
  if (run_blocks < 0 || run_blocks > 4096) return fail(HFDTD_EINVAL, "run_blocks must be 0..4096");
  if (nontemporal < -2 || nontemporal > 1) return fail(HFDTD_EINVAL, "nontemporal must be -2..1");
  Engine *e = &h->s;
  if (rows_per_wave) e->march_rows = rows_per_wave;
  if (planes_i) e->march_planes = planes_i < 0 ? 0 : planes_i;
  if (run_blocks) e->march_run = run_blocks;
  if (nontemporal != -2) e->march_nt = nontemporal;
  return HFDTD_OK;
}

int hfdtd_set_march_width(hfdtd_engine *h, int width) {
  if (!h) return fail(HFDTD_EINVAL, "NULL argument");
  if (width != 0 && width != 64 && width != 32)
    return fail(HFDTD_EINVAL, "march width must be 64 or 32 (0 = 64)");
  h->s.march_width = width ? width : 64;
  return HFDTD_OK;
}

int hfdtd_set_overlap(hfdtd_engine *h, int on) {
  if (!h) return fail(HFDTD_EINVAL, "NULL argument");
  if (on != 0 && on != 1) return fail(HFDTD_EINVAL, "on must be 0 or 1");
  HF_HIP(hipSetDevice(h->s.device));
  int rc = flush_gather(&h->s);
  if (rc) return rc;
  h->s.fuse_gather = on != 0;
  return HFDTD_OK;
}

int hfdtd_set_march_addressing(hfdtd_engine *h, int per_volume) {
  if (!h) return fail(HFDTD_EINVAL, "NULL argument");
  if (per_volume < 0 || per_volume > 1) return fail(HFDTD_EINVAL, "per_volume must be 0 or 1");
  h->s.march_per_volume = per_volume != 0;
  return HFDTD_OK;
}

int hfdtd_set_fused_tiling(hfdtd_engine *h, int rows_j, int planes_i, int xcd_groups) {
  if (!h) return fail(HFDTD_EINVAL, "NULL argument");
  if ((rows_j != 0 && rows_j != 1 && rows_j != 2 && rows_j != 4 && rows_j != 6 &&
       rows_j != 8 && rows_j != 10 && rows_j != 12) ||
      planes_i < -1)
    return fail(HFDTD_EINVAL, "rows_j must be 1, 2, 4, 6, 8, 10 or 12 and planes_i >= -1");
  if (xcd_groups < 0 || xcd_groups > 64) return fail(HFDTD_EINVAL, "xcd_groups must be 0..64");
  if (rows_j) h->s.tile_rows = rows_j;
  if (planes_i) h->s.tile_planes = planes_i < 0 ? 0 : planes_i;
  if (xcd_groups) h->s.xcd = xcd_groups;
  return HFDTD_OK;
}

}  // extern "C"
