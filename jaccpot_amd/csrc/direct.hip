// Float64 direct summation at a set of targets -- the error reference every
// measurement is reported against (jaccpot/autodiff.py:12-98: softened
// Newtonian sum, self term masked).
//
// Grid: (target blocks of 256) x (source chunks).  Each workgroup streams its
// source chunk through LDS; per-chunk partials go to scratch (the caller's, or
// for jfm_direct_sum_f64_static a fixed per-GPU array) and a second kernel
// reduces them in chunk order, so the result is deterministic.

#include <hip/hip_runtime.h>

#include <mutex>

#include "launch.hpp"
#include "rows.hpp"

namespace jfm {
namespace {

constexpr int kThreads = 256;
constexpr int kChunks = 256;

__global__ __launch_bounds__(kThreads) void direct_partial_kernel(const double* __restrict__ pos4, int32_t n,
                                                                  const int32_t* __restrict__ targets, int32_t nt,
                                                                  double eps2, double* __restrict__ scratch) {
  __shared__ V4<double> tile[kThreads];
  const int t = blockIdx.x * kThreads + threadIdx.x;
  const int chunk = blockIdx.y;
  const int64_t cb = static_cast<int64_t>(n) * chunk / kChunks;
  const int64_t ce = static_cast<int64_t>(n) * (chunk + 1) / kChunks;
  const int ti = t < nt ? targets[t] : 0;
  const V4<double> xi = load4(pos4, ti);
  double ax = 0, ay = 0, az = 0, ap = 0;
  for (int64_t base = cb; base < ce; base += kThreads) {
    __syncthreads();
    const int64_t j = base + threadIdx.x;
    tile[threadIdx.x] = j < ce ? load4(pos4, static_cast<int>(j)) : V4<double>{0.0, 0.0, 0.0, 0.0};
    __syncthreads();
    const int lim = static_cast<int>(min<int64_t>(kThreads, ce - base));
    for (int k = 0; k < lim; ++k) {
      if (base + k == ti) continue;
      const V4<double> s = tile[k];
      const double dx = s.x - xi.x, dy = s.y - xi.y, dz = s.z - xi.z;
      const double r2 = dx * dx + dy * dy + dz * dz + eps2;
      const double rinv = r2 > 0.0 ? 1.0 / sqrt(r2) : 0.0;
      const double mr = s.w * rinv;
      const double mr3 = mr * rinv * rinv;
      ax += mr3 * dx;
      ay += mr3 * dy;
      az += mr3 * dz;
      ap += mr;
    }
  }
  if (t < nt) store4(scratch, chunk * nt + t, V4<double>{ax, ay, az, ap});
}

__global__ __launch_bounds__(kThreads) void direct_reduce_kernel(const double* __restrict__ scratch, int32_t nt,
                                                                 double G, double* __restrict__ acc,
                                                                 double* __restrict__ pot) {
  const int t = blockIdx.x * kThreads + threadIdx.x;
  if (t >= nt) return;
  double ax = 0, ay = 0, az = 0, ap = 0;
  for (int c = 0; c < kChunks; ++c) {
    const V4<double> v = load4(scratch, c * nt + t);
    ax += v.x;
    ay += v.y;
    az += v.z;
    ap += v.w;
  }
  acc[3 * t + 0] = G * ax;
  acc[3 * t + 1] = G * ay;
  acc[3 * t + 2] = G * az;
  if (pot) pot[t] = -G * ap;
}

// Scratch of jfm_direct_sum_f64_static: the per-chunk partials of up to
// kStaticTargets targets (kChunks * 4 * kStaticTargets doubles = 32 MiB),
// one copy per GPU, held until the process ends.
constexpr int32_t kStaticTargets = 4096;
__device__ __align__(256) double scratch[static_cast<int64_t>(kChunks) * 4 * kStaticTargets];

// Device address of `scratch` on each GPU, looked up once per device with that
// device current (host code never passes the symbol itself to a kernel).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

int32_t static_scratch(double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(JFM_ERR_HIP, std::string("jfm_direct_sum_f64_static: hipGetDevice: ") +
                                                   hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return fail(JFM_ERR_ARG, "jfm_direct_sum_f64_static: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(JFM_ERR_HIP, std::string("jfm_direct_sum_f64_static: hipGetSymbolAddress: ") +
                                   hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return JFM_OK;
}

int32_t launch_direct(hipStream_t s, const double* pos4, int32_t n, const int32_t* targets, int32_t nt, double eps2,
                      double G, double* part, double* acc, double* pot, const char* what) {
  dim3 grid((nt + kThreads - 1) / kThreads, kChunks);
  direct_partial_kernel<<<grid, kThreads, 0, s>>>(pos4, n, targets, nt, eps2, part);
  int32_t rc = check_launch(what);
  if (rc != JFM_OK) return rc;
  direct_reduce_kernel<<<(nt + kThreads - 1) / kThreads, kThreads, 0, s>>>(part, nt, G, acc, pot);
  return check_launch(what);
}

}  // namespace
}  // namespace jfm

using namespace jfm;

extern "C" {

int64_t jfm_direct_scratch_doubles(int32_t nt) { return static_cast<int64_t>(kChunks) * 4 * (nt > 0 ? nt : 0); }

int32_t jfm_direct_sum_f64(void* stream, const double* pos4, int32_t n, const int32_t* targets, int32_t nt,
                           double eps2, double G, double* scratch, double* acc, double* pot) {
  if (n < 1 || nt < 0) return fail(JFM_ERR_ARG, "jfm_direct_sum_f64: bad sizes");
  if (nt == 0) return JFM_OK;
  if (!pos4 || !targets || !scratch || !acc) return fail(JFM_ERR_ARG, "jfm_direct_sum_f64: NULL buffer");
  if (static_cast<int64_t>(kChunks) * nt > (int64_t(1) << 31) / 4)
    return fail(JFM_ERR_ARG, "jfm_direct_sum_f64: too many targets");
  return launch_direct(as_stream(stream), pos4, n, targets, nt, eps2, G, scratch, acc, pot, "jfm_direct_sum_f64");
}

int32_t jfm_direct_sum_f64_static_max_targets(void) { return kStaticTargets; }

int32_t jfm_direct_sum_f64_static(void* stream, const double* pos4, int32_t n, const int32_t* targets, int32_t nt,
                                  double eps2, double G, double* acc, double* pot) {
  if (n < 1 || nt < 0) return fail(JFM_ERR_ARG, "jfm_direct_sum_f64_static: bad sizes");
  if (nt > kStaticTargets)
    return fail(JFM_ERR_ARG, "jfm_direct_sum_f64_static: " + std::to_string(nt) + " targets exceed the maximum of " +
                                 std::to_string(kStaticTargets));
  if (nt == 0) return JFM_OK;
  if (!pos4 || !targets || !acc) return fail(JFM_ERR_ARG, "jfm_direct_sum_f64_static: NULL buffer");
  double* part = nullptr;
  const int32_t rc = static_scratch(&part);
  if (rc != JFM_OK) return rc;
  // every partial the reduction reads is written by this call's first kernel,
  // so the previous call's contents need no clearing
  return launch_direct(as_stream(stream), pos4, n, targets, nt, eps2, G, part, acc, pot,
                       "jfm_direct_sum_f64_static");
}

}  // extern "C"
