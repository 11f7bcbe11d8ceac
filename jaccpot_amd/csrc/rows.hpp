// Vectorised row access for 16-byte-padded coefficient rows and pos4/centers4 records.
#pragma once

#include <hip/hip_runtime.h>

namespace jfm {

template <typename T>
struct alignas(4 * sizeof(T)) V4 {
  T x, y, z, w;
};

template <typename T>
__device__ __forceinline__ V4<T> load4(const T* __restrict__ p, int idx) {
  return reinterpret_cast<const V4<T>*>(p)[idx];
}

template <typename T>
__device__ __forceinline__ void store4(T* __restrict__ p, int idx, V4<T> v) {
  reinterpret_cast<V4<T>*>(p)[idx] = v;
}

// Loads the first C entries of a 16-byte-aligned row.
template <typename T, int C>
__device__ __forceinline__ void load_row(const T* __restrict__ row, T* v) {
  if constexpr (sizeof(T) == 4) {
    constexpr int NV = (C + 3) / 4;
    const float4* r4 = reinterpret_cast<const float4*>(row);
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const float4 q = r4[k];
      if (4 * k + 0 < C) v[4 * k + 0] = q.x;
      if (4 * k + 1 < C) v[4 * k + 1] = q.y;
      if (4 * k + 2 < C) v[4 * k + 2] = q.z;
      if (4 * k + 3 < C) v[4 * k + 3] = q.w;
    }
  } else {
    constexpr int NV = (C + 1) / 2;
    const double2* r2 = reinterpret_cast<const double2*>(row);
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const double2 q = r2[k];
      if (2 * k + 0 < C) v[2 * k + 0] = q.x;
      if (2 * k + 1 < C) v[2 * k + 1] = q.y;
    }
  }
}

// Stores C entries (zero-filling the row padding) into a 16-byte-aligned row.
template <typename T, int C>
__device__ __forceinline__ void store_row(T* __restrict__ row, const T* v) {
  if constexpr (sizeof(T) == 4) {
    constexpr int NV = (C + 3) / 4;
    float4* r4 = reinterpret_cast<float4*>(row);
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      float4 q;
      q.x = (4 * k + 0 < C) ? v[4 * k + 0] : 0.0f;
      q.y = (4 * k + 1 < C) ? v[4 * k + 1] : 0.0f;
      q.z = (4 * k + 2 < C) ? v[4 * k + 2] : 0.0f;
      q.w = (4 * k + 3 < C) ? v[4 * k + 3] : 0.0f;
      r4[k] = q;
    }
  } else {
    constexpr int NV = (C + 1) / 2;
    double2* r2 = reinterpret_cast<double2*>(row);
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double2 q;
      q.x = (2 * k + 0 < C) ? v[2 * k + 0] : 0.0;
      q.y = (2 * k + 1 < C) ? v[2 * k + 1] : 0.0;
      r2[k] = q;
    }
  }
}

}  // namespace jfm
