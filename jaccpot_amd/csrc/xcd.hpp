// Launch-order -> logical block mapping for the eight XCDs of an MI355X.
#pragma once

#include <hip/hip_runtime.h>

namespace jfm {

constexpr int kXcds = 8;

// Workgroups are dealt round-robin to the eight XCDs; this maps launch order
// to a logical block so that each XCD takes one contiguous run of blocks
// (work items in tree order: neighbouring targets share most of their
// sources, which then stay in that XCD's L2).  Bijective: XCD x < r owns
// q + 1 blocks, the others q, where nblk = 8 q + r.
__device__ __forceinline__ int xcd_contiguous(int b, int nblk) {
  const int q = nblk / kXcds, r = nblk % kXcds;
  const int x = b % kXcds, slot = b / kXcds;
  return x * q + min(x, r) + slot;
}

}  // namespace jfm
