// Host (CPU) twins of the device operators: the same templates from
// device_math.hpp instantiated for the host, one pair at a time.  Validation
// only -- the CPU suite compares them with the oracle so the kernels' math is
// checked without a GPU.  No product path calls them.

#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "launch.hpp"

namespace jfm {
namespace {

template <typename T, int P>
void host_op(int32_t op, const double* in, const double* delta, double* out) {
  constexpr int C = ncoef(P);
  T v[C], w[C];
  const T dx = static_cast<T>(delta[0]), dy = static_cast<T>(delta[1]), dz = static_cast<T>(delta[2]);
  if (op == 3) {  // p2m of one particle: in[0] = mass, delta = particle - centre
    T re[ntri(P)], im[ntri(P)];
    solid_harmonics<T, P>(dx, dy, dz, re, im);
    const T mass = static_cast<T>(in[0]);
    for (int n = 0; n <= P; ++n) {
      out[sh_idx(n, 0)] = static_cast<double>(mass * re[tri(n, 0)]);
      for (int m = 1; m <= n; ++m) {
        out[sh_idx(n, m)] = static_cast<double>(mass * re[tri(n, m)]);
        out[sh_idx(n, -m)] = static_cast<double>(mass * im[tri(n, m)]);
      }
    }
    return;
  }
  for (int j = 0; j < C; ++j) v[j] = static_cast<T>(in[j]);
  if (op == 4) {  // l2p: out = phi, d phi / d delta
    T phi, gx, gy, gz;
    l2p_eval<T, P>([&v](int j) { return v[j]; }, dx, dy, dz, phi, gx, gy, gz);
    out[0] = phi;
    out[1] = gx;
    out[2] = gy;
    out[3] = gz;
    return;
  }
  Align<T, P> a;
  a.init(dx, dy, dz);
  if (op == 0) {  // m2m, delta = source - destination
    rotate_all<T, P, Rot::ToZMultipole>(v, a);
    zshift_m2m<T, P>(v, a.r);
    rotate_all<T, P, Rot::FromZMultipole>(v, a);
    for (int j = 0; j < C; ++j) out[j] = v[j];
  } else if (op == 1) {  // m2l, delta = target - source
    rotate_all<T, P, Rot::ToZMultipole>(v, a);
    zm2l<T, P>(v, w, a.r);
    rotate_all<T, P, Rot::FromZLocal>(w, a);
    for (int j = 0; j < C; ++j) out[j] = w[j];
  } else {  // l2l, delta = parent - child
    rotate_all<T, P, Rot::ToZLocal>(v, a);
    zshift_l2l<T, P>(v, a.r);
    rotate_all<T, P, Rot::FromZLocal>(v, a);
    for (int j = 0; j < C; ++j) out[j] = v[j];
  }
}

}  // namespace
}  // namespace jfm

extern "C" int32_t jfm_host_operator(int32_t op, int32_t p, int32_t dtype, const double* in, const double* delta,
                                     double* out) {
  if (op < 0 || op > 4) return jfm::fail(JFM_ERR_ARG, "jfm_host_operator: op must be 0..4");
  if (!in || !delta || !out) return jfm::fail(JFM_ERR_ARG, "jfm_host_operator: NULL buffer");
  JFM_FOR_DTYPE(dtype, JFM_FOR_ORDER(p, { jfm::host_op<T, P>(op, in, delta, out); }));
  return JFM_OK;
}
