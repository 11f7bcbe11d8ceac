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

template <typename T, int P>
__global__ __launch_bounds__(kWave) void m2l_kernel(const int32_t* __restrict__ targets,
                                                    const int32_t* __restrict__ far_off,
                                                    const int32_t* __restrict__ far_src, const T* __restrict__ mult,
                                                    const T* __restrict__ centers4, T* __restrict__ locals,
                                                    int32_t cs) {
  constexpr int C = ncoef(P);
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
    Align<T, P> a;
    a.init(ct.x - cs4.x, ct.y - cs4.y, ct.z - cs4.z);  // target - source
    rotate_all<T, P, Rot::ToZMultipole>(v, a);
    T w[C];
    zm2l<T, P>(v, w, a.r);
    rotate_all<T, P, Rot::FromZLocal>(w, a);
#pragma unroll
    for (int j = 0; j < C; ++j) acc[j] += w[j];
  }
#pragma unroll
  for (int j = 0; j < C; ++j) red[lane][j] = acc[j];
  __syncthreads();
  for (int j = lane; j < cs; j += kWave) {
    T s = T(0);
    if (j < C)
      for (int l = 0; l < kWave; ++l) s += red[l][j];
    locals[static_cast<int64_t>(tgt) * cs + j] = s;
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
    m2l_kernel<T, P><<<num_targets, kWave, 0, s>>>(targets, far_off, far_src, (const T*)multipoles,
                                                   (const T*)centers4, (T*)locals, cs);
  }));
  return check_launch("jfm_m2l");
}
