// M2L over the far interaction list: rotate the source multipole onto the
// pair axis, apply the z-axis M2L core, rotate the local back and sum per
// target node.
//
// Mapping: one wave per target node; lane l takes the target's far pairs
// l, l+64, ...  Every lane runs the whole per-pair cascade in registers:
// the rotation blocks are pair-independent constants folded into the
// instruction stream (tables_gen.h), so the per-pair state is just the
// coefficient vector and the cos/sin(m az), cos/sin(m ax) ladders.  The lane
// partial sums are reduced through LDS in a fixed order, so the result is
// deterministic and written once per target (no atomics).

#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "launch.hpp"
#include "rows.hpp"

namespace jfm {
namespace {

constexpr int kWave = 64;

// G <= P is the pair's expansion order (gear): the source multipole is read
// to degree G, translated at order G and added into degrees <= G of the
// target local.  ACC adds into the existing rows instead of overwriting them
// (used when a target's pairs are split across several gears).
template <typename T, int P, int G, bool ACC>
__global__ __launch_bounds__(kWave) void m2l_kernel(const int32_t* __restrict__ targets,
                                                    const int32_t* __restrict__ far_off,
                                                    const int32_t* __restrict__ far_src, const T* __restrict__ mult,
                                                    const T* __restrict__ centers4, T* __restrict__ locals,
                                                    int32_t cs) {
  constexpr int C = ncoef(G);
  __shared__ T red[kWave][C + 1];
  const int tgt = targets[blockIdx.x];
  const int lane = threadIdx.x;
  const int b = far_off[tgt], e = far_off[tgt + 1];
  const V4<T> ct = load4(centers4, tgt);
  T acc[C];
#pragma unroll
  for (int j = 0; j < C; ++j) acc[j] = T(0);
  for (int k = b + lane; k < e; k += kWave) {
    const int src = far_src[k];
    const V4<T> cs4 = load4(centers4, src);
    T v[C];
    load_row<T, C>(mult + static_cast<int64_t>(src) * cs, v);
    Align<T, G> a;
    a.init(ct.x - cs4.x, ct.y - cs4.y, ct.z - cs4.z);  // target - source
    rotate_all<T, G, Rot::ToZMultipole>(v, a);
    T w[C];
    zm2l<T, G>(v, w, a.r);
    rotate_all<T, G, Rot::FromZLocal>(w, a);
#pragma unroll
    for (int j = 0; j < C; ++j) acc[j] += w[j];
  }
#pragma unroll
  for (int j = 0; j < C; ++j) red[lane][j] = acc[j];
  __syncthreads();
  const int lim = ACC ? C : cs;
  for (int j = lane; j < lim; j += kWave) {
    T s = T(0);
    if (j < C)
      for (int l = 0; l < kWave; ++l) s += red[l][j];
    T* dst = locals + static_cast<int64_t>(tgt) * cs + j;
    if (ACC)
      *dst += s;
    else
      *dst = s;
  }
}

}  // namespace
}  // namespace jfm

using namespace jfm;

extern "C" int32_t jfm_m2l(void* stream, int32_t dtype, int32_t p, const int32_t* targets, int32_t num_targets,
                           const int32_t* far_off, const int32_t* far_src, const void* multipoles,
                           const void* centers4, void* locals) {
  if (num_targets < 0) return fail(JFM_ERR_ARG, "jfm_m2l: negative target count");
  if (num_targets == 0) return JFM_OK;
  if (!targets || !far_off || !far_src || !multipoles || !centers4 || !locals)
    return fail(JFM_ERR_ARG, "jfm_m2l: NULL buffer");
  hipStream_t s = as_stream(stream);
  JFM_FOR_DTYPE(dtype, JFM_FOR_ORDER(p, {
    const int cs = coef_stride(P, sizeof(T));
    m2l_kernel<T, P, P, false><<<num_targets, kWave, 0, s>>>(targets, far_off, far_src, (const T*)multipoles,
                                                             (const T*)centers4, (T*)locals, cs);
  }));
  return check_launch("jfm_m2l");
}

namespace {
template <typename T, int P>
int32_t launch_gear(hipStream_t s, int32_t gear, bool acc, int32_t nt, const int32_t* targets, const int32_t* off,
                    const int32_t* src, const void* mult, const void* centers4, void* locals) {
  const int cs = coef_stride(P, sizeof(T));
#define JFM_GEAR(G)                                                                                              \
  if (gear == G) {                                                                                               \
    if constexpr (G <= P) {                                                                                      \
      if (acc)                                                                                                   \
        m2l_kernel<T, P, G, true><<<nt, kWave, 0, s>>>(targets, off, src, (const T*)mult, (const T*)centers4,    \
                                                       (T*)locals, cs);                                          \
      else                                                                                                       \
        m2l_kernel<T, P, G, false><<<nt, kWave, 0, s>>>(targets, off, src, (const T*)mult, (const T*)centers4,   \
                                                        (T*)locals, cs);                                         \
      return check_launch("jfm_m2l_gear");                                                                       \
    }                                                                                                            \
  }
  JFM_GEAR(1) JFM_GEAR(2) JFM_GEAR(3) JFM_GEAR(4) JFM_GEAR(5) JFM_GEAR(6) JFM_GEAR(7) JFM_GEAR(8)
#undef JFM_GEAR
  return fail(JFM_ERR_UNSUPPORTED, "jfm_m2l_gear: gear must be in [1, p]");
}
}  // namespace

extern "C" int32_t jfm_m2l_gear(void* stream, int32_t dtype, int32_t p, int32_t gear, int32_t accumulate,
                                const int32_t* targets, int32_t num_targets, const int32_t* far_off,
                                const int32_t* far_src, const void* multipoles, const void* centers4,
                                void* locals) {
  if (num_targets < 0) return fail(JFM_ERR_ARG, "jfm_m2l_gear: negative target count");
  if (gear < 1 || gear > p) return fail(JFM_ERR_ARG, "jfm_m2l_gear: gear must be in [1, p]");
  if (num_targets == 0) return JFM_OK;
  if (!targets || !far_off || !far_src || !multipoles || !centers4 || !locals)
    return fail(JFM_ERR_ARG, "jfm_m2l_gear: NULL buffer");
  hipStream_t s = as_stream(stream);
  JFM_FOR_DTYPE(dtype, JFM_FOR_ORDER(p, {
    return launch_gear<T, P>(s, gear, accumulate != 0, num_targets, targets, far_off, far_src, multipoles,
                             centers4, locals);
  }));
  return JFM_OK;
}
