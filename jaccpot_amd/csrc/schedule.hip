// Prepare-time list surgery on the device: splitting the far CSR by per-pair
// gear and laying a target-grouped pair list out for the lane-per-pair M2L.
//
// The reference builds its per-order pair groups on the host
// (runtime/_interaction_cache.py, p_gears bucketing before the M2L launches);
// here every step is a small stream-ordered kernel so prepare_state never
// walks the (tens of millions of) far pairs on the CPU:
//   row_of[k]            target of pair k of a CSR (binary search in off)
//   gear keys            (gear slot << 32 | k): one radix sort groups the pairs
//                        by gear and keeps each gear's CSR order
//   gather               target and source of the pairs of one gear
//   pair keys            (target << 32 | source) -> CSR by jfm_keys_to_csr
//   non-empty rows       targets that have pairs (appended, then sorted)
//   flat spans           targets whose pairs cross a 64-pair block boundary,
//                        one per such target, found at its first crossing

#include <hip/hip_runtime.h>

#include <cstdint>

#include "launch.hpp"

namespace jfm {
namespace {

constexpr int kFlatBlock = 64;  // pairs per wave of the lane-per-pair M2L

__global__ void __launch_bounds__(256) pair_rows_kernel(int32_t rows, const int32_t* __restrict__ off, int64_t K,
                                                        int32_t* __restrict__ row_of) {
  const int64_t k = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (k >= K) return;
  int32_t lo = 0, hi = rows;  // last row r with off[r] <= k
  while (hi - lo > 1) {
    const int32_t mid = (lo + hi) >> 1;
    if (off[mid] <= k) lo = mid;
    else hi = mid;
  }
  row_of[k] = lo;
}

__global__ void __launch_bounds__(256) gear_keys_kernel(int64_t K, const int8_t* __restrict__ gear_of_pair,
                                                        int32_t g0, int32_t g1, int32_t g2,
                                                        uint64_t* __restrict__ keys, int32_t* __restrict__ bad) {
  const int64_t k = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (k >= K) return;
  const int32_t g = gear_of_pair[k];
  uint64_t slot = 3;
  if (g == g0) slot = 0;
  else if (g == g1) slot = 1;
  else if (g == g2) slot = 2;
  if (slot == 3) atomicOr(bad, 1);
  keys[k] = (slot << 32) | static_cast<uint32_t>(k);
}

__global__ void __launch_bounds__(256) gather_pairs_kernel(int64_t n, const int32_t* __restrict__ index,
                                                           const int32_t* __restrict__ row_of,
                                                           const int32_t* __restrict__ src,
                                                           int32_t* __restrict__ out_tgt,
                                                           int32_t* __restrict__ out_src) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const int32_t k = index[i];
  out_tgt[i] = row_of[k];
  out_src[i] = src[k];
}

__global__ void __launch_bounds__(256) pair_keys_kernel(int64_t n, const int32_t* __restrict__ tgt,
                                                        const int32_t* __restrict__ src, uint64_t* __restrict__ keys) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) keys[i] = (uint64_t(uint32_t(tgt[i])) << 32) | uint32_t(src[i]);
}

__global__ void __launch_bounds__(256) nonempty_rows_kernel(int32_t rows, const int32_t* __restrict__ off,
                                                            uint64_t* __restrict__ out,
                                                            unsigned long long* __restrict__ count) {
  const int32_t r = blockIdx.x * 256 + threadIdx.x;
  if (r >= rows || off[r + 1] <= off[r]) return;
  out[atomicAdd(count, 1ull)] = uint64_t(r);
}

// boundary w (pairs 64w-1 | 64w): a target present on both sides spans
// blocks; it is recorded at its first crossing, with the block of its last pair
__global__ void __launch_bounds__(256) flat_spans_kernel(int64_t K, const int32_t* __restrict__ tgt,
                                                         int32_t* __restrict__ span_tgt, int32_t* __restrict__ span_w0,
                                                         int32_t* __restrict__ span_w1,
                                                         unsigned long long* __restrict__ count) {
  const int64_t w = int64_t(blockIdx.x) * 256 + threadIdx.x + 1;
  const int64_t at = w * kFlatBlock;
  if (at >= K) return;
  const int32_t t = tgt[at];
  if (tgt[at - 1] != t) return;
  const int64_t prev = (w - 1) * kFlatBlock;  // first pair of block w-1
  if (prev > 0 && tgt[prev - 1] == t) return;  // started before block w-1: not its first crossing
  int64_t w1 = w;
  while ((w1 + 1) * kFlatBlock < K && tgt[(w1 + 1) * kFlatBlock] == t) ++w1;
  const unsigned long long i = atomicAdd(count, 1ull);
  span_tgt[i] = t;
  span_w0[i] = static_cast<int32_t>(w - 1);
  span_w1[i] = static_cast<int32_t>(w1);
}

inline unsigned grid_for(int64_t n) { return static_cast<unsigned>((n + 255) / 256); }

}  // namespace
}  // namespace jfm

using jfm::fail;

extern "C" {

int32_t jfm_pair_rows(void* stream, int32_t rows, const int32_t* off, int64_t K, int32_t* row_of) {
  if (rows < 1 || K < 0 || !off || (K > 0 && !row_of)) return fail(JFM_ERR_ARG, "jfm_pair_rows: bad argument");
  if (K == 0) return JFM_OK;
  hipLaunchKernelGGL(jfm::pair_rows_kernel, dim3(jfm::grid_for(K)), dim3(256), 0, jfm::as_stream(stream), rows, off,
                     K, row_of);
  return jfm::check_launch("jfm_pair_rows");
}

int32_t jfm_gear_keys(void* stream, int64_t K, const int8_t* gear_of_pair, const int32_t* gears3, uint64_t* keys,
                      int32_t* flags) {
  if (K < 0 || K > INT32_MAX || !gears3 || !flags || (K > 0 && (!gear_of_pair || !keys)))
    return fail(JFM_ERR_ARG, "jfm_gear_keys: bad argument");
  if (K == 0) return JFM_OK;
  hipLaunchKernelGGL(jfm::gear_keys_kernel, dim3(jfm::grid_for(K)), dim3(256), 0, jfm::as_stream(stream), K,
                     gear_of_pair, gears3[0], gears3[1], gears3[2], keys, flags);
  return jfm::check_launch("jfm_gear_keys");
}

int32_t jfm_gather_pairs(void* stream, int64_t n, const int32_t* index, const int32_t* row_of, const int32_t* src,
                         int32_t* out_tgt, int32_t* out_src) {
  if (n < 0 || (n > 0 && (!index || !row_of || !src || !out_tgt || !out_src)))
    return fail(JFM_ERR_ARG, "jfm_gather_pairs: bad argument");
  if (n == 0) return JFM_OK;
  hipLaunchKernelGGL(jfm::gather_pairs_kernel, dim3(jfm::grid_for(n)), dim3(256), 0, jfm::as_stream(stream), n, index,
                     row_of, src, out_tgt, out_src);
  return jfm::check_launch("jfm_gather_pairs");
}

int32_t jfm_pair_keys(void* stream, int64_t n, const int32_t* tgt, const int32_t* src, uint64_t* keys) {
  if (n < 0 || (n > 0 && (!tgt || !src || !keys))) return fail(JFM_ERR_ARG, "jfm_pair_keys: bad argument");
  if (n == 0) return JFM_OK;
  hipLaunchKernelGGL(jfm::pair_keys_kernel, dim3(jfm::grid_for(n)), dim3(256), 0, jfm::as_stream(stream), n, tgt, src,
                     keys);
  return jfm::check_launch("jfm_pair_keys");
}

int32_t jfm_nonempty_rows(void* stream, int32_t rows, const int32_t* off, uint64_t* out, uint64_t* count) {
  if (rows < 0 || !off || !count || (rows > 0 && !out)) return fail(JFM_ERR_ARG, "jfm_nonempty_rows: bad argument");
  if (rows == 0) return JFM_OK;
  hipLaunchKernelGGL(jfm::nonempty_rows_kernel, dim3(jfm::grid_for(rows)), dim3(256), 0, jfm::as_stream(stream), rows,
                     off, out, reinterpret_cast<unsigned long long*>(count));
  return jfm::check_launch("jfm_nonempty_rows");
}

int32_t jfm_flat_spans(void* stream, int64_t K, const int32_t* pair_tgt, int32_t* span_tgt, int32_t* span_w0,
                       int32_t* span_w1, uint64_t* count) {
  if (K < 0 || !count || (K > 0 && (!pair_tgt || !span_tgt || !span_w0 || !span_w1)))
    return fail(JFM_ERR_ARG, "jfm_flat_spans: bad argument");
  const int64_t boundaries = K > 0 ? (K - 1) / jfm::kFlatBlock : 0;
  if (boundaries == 0) return JFM_OK;
  hipLaunchKernelGGL(jfm::flat_spans_kernel, dim3(jfm::grid_for(boundaries)), dim3(256), 0, jfm::as_stream(stream), K,
                     pair_tgt, span_tgt, span_w0, span_w1, reinterpret_cast<unsigned long long*>(count));
  return jfm::check_launch("jfm_flat_spans");
}

}  // extern "C"
