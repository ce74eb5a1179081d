// Streaming-bandwidth probe: what plain streaming kernels reach on this
// device, the practical ceiling against which a sweep's measured traffic is
// read (the 8 TB/s figure is the HBM peak; no kernel that both reads and
// writes gets near it).  Each kernel walks its arrays with 16-byte loads and
// stores in a grid-stride loop; the caller picks the grid.
#include "common.h"

namespace hfdtd {
namespace {

typedef float sf4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_stream_copy(const sf4 *__restrict__ a,
                                                     sf4 *__restrict__ c, size_t n) {
  for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    c[i] = a[i];
}

__global__ __launch_bounds__(256) void k_stream_triad(const sf4 *__restrict__ a,
                                                      const sf4 *__restrict__ b,
                                                      sf4 *__restrict__ c, size_t n) {
  for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    c[i] = a[i] + 1.5f * b[i];
}

// six arrays in, six out (the balance of a Yee step's field traffic)
__global__ __launch_bounds__(256) void k_stream_six(const sf4 *__restrict__ a,
                                                    sf4 *__restrict__ c, size_t n) {
  for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    sf4 v[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) v[q] = a[i + q * n];
#pragma unroll
    for (int q = 0; q < 6; ++q) c[i + q * n] = v[q] * 1.01f + v[(q + 1) % 6];
  }
}

__global__ __launch_bounds__(256) void k_stream_read(const sf4 *__restrict__ a,
                                                     float *__restrict__ out, size_t n) {
  sf4 s = {0.f, 0.f, 0.f, 0.f};
  for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    s += a[i];
  if (s.x + s.y + s.z + s.w == 12345.0f) out[0] = s.x;  // keeps the loads; never true here
}

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

struct Ev {
  hipEvent_t e = nullptr;
  ~Ev() {
    if (e) (void)hipEventDestroy(e);
  }
};

}  // namespace
}  // namespace hfdtd

using namespace hfdtd;

extern "C" int hfdtd_stream_probe(int device, int kind, uint64_t bytes, int blocks, int reps,
                                  double *seconds) {
  if (!seconds) return fail(HFDTD_EINVAL, "NULL argument");
  if (kind < 0 || kind > 3) return fail(HFDTD_EINVAL, "kind must be 0..3");
  if (bytes < 4096 || bytes % 64 || bytes > (16ull << 30))
    return fail(HFDTD_EINVAL, "bytes must be a multiple of 64 in [4 KiB, 16 GiB]");
  if (blocks < 1 || blocks > (1 << 20) || reps < 1 || reps > 10000)
    return fail(HFDTD_EINVAL, "blocks must be 1..2^20 and reps 1..10000");
  int n_dev = 0;
  int rc = hfdtd_device_count(&n_dev);
  if (rc) return rc;
  if (device < 0 || device >= n_dev) return fail(HFDTD_ENODEV, "device index out of range");
  HF_HIP(hipSetDevice(device));
  const size_t n = bytes / 16;
  const int arrays = kind == 2 ? 6 : 1;
  DevBuf a, b, c;
  HF_HIP(hipMalloc(&a.p, bytes * arrays));
  HF_HIP(hipMalloc(&c.p, kind == 3 ? 64 : bytes * arrays));
  if (kind == 1) HF_HIP(hipMalloc(&b.p, bytes));
  HF_HIP(hipMemset(a.p, 0, bytes * arrays));
  if (b.p) HF_HIP(hipMemset(b.p, 0, bytes));
  auto launch = [&]() {
    const dim3 g((unsigned)blocks), t(256);
    if (kind == 0)
      hipLaunchKernelGGL(k_stream_copy, g, t, 0, 0, (const sf4 *)a.p, (sf4 *)c.p, n);
    else if (kind == 1)
      hipLaunchKernelGGL(k_stream_triad, g, t, 0, 0, (const sf4 *)a.p, (const sf4 *)b.p,
                         (sf4 *)c.p, n);
    else if (kind == 2)
      hipLaunchKernelGGL(k_stream_six, g, t, 0, 0, (const sf4 *)a.p, (sf4 *)c.p, n);
    else
      hipLaunchKernelGGL(k_stream_read, g, t, 0, 0, (const sf4 *)a.p, (float *)c.p, n);
    return hipGetLastError();
  };
  HF_HIP(launch());  // warm-up
  HF_HIP(hipDeviceSynchronize());
  Ev e0, e1;
  HF_HIP(hipEventCreate(&e0.e));
  HF_HIP(hipEventCreate(&e1.e));
  HF_HIP(hipEventRecord(e0.e, 0));
  for (int r = 0; r < reps; ++r) HF_HIP(launch());
  HF_HIP(hipEventRecord(e1.e, 0));
  HF_HIP(hipEventSynchronize(e1.e));
  float ms = 0.f;
  HF_HIP(hipEventElapsedTime(&ms, e0.e, e1.e));
  *seconds = (double)ms * 1e-3 / reps;
  return HFDTD_OK;
}
