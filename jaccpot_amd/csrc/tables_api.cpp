// Host export of the B_U constants the kernels are compiled with.
#include <cstdint>
#include <vector>

#include "common.hpp"
#include "tables.hpp"

extern "C" int32_t jfm_tables_bu(int32_t p, double* out, int64_t capacity) {
  if (p < 0 || p > JFM_MAX_ORDER) return jfm::fail(JFM_ERR_ARG, "jfm_tables_bu: order out of range");
  int64_t need = 0;
  for (int l = 0; l <= p; ++l) need += static_cast<int64_t>(2 * l + 1) * (2 * l + 1);
  if (!out || capacity < need) return jfm::fail(JFM_ERR_ARG, "jfm_tables_bu: buffer too small");
  int64_t pos = 0;
  for (int l = 0; l <= p; ++l) {
    std::vector<double> b = jfm::dehnen_bu(l);
    for (double v : b) out[pos++] = v;
  }
  return JFM_OK;
}
