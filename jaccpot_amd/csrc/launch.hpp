// Launch-side helpers shared by the .hip translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "common.hpp"

namespace jfm {

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int32_t check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(JFM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return JFM_OK;
}

inline bool valid_dtype(int32_t d) { return d == JFM_F32 || d == JFM_F64; }

}  // namespace jfm

// Expands the body with a constexpr int P for every order compiled into the library.
#define JFM_FOR_ORDER(p, ...)                                                   \
  switch (p) {                                                                   \
    case 1: { constexpr int P = 1; __VA_ARGS__; } break;                                \
    case 2: { constexpr int P = 2; __VA_ARGS__; } break;                                \
    case 3: { constexpr int P = 3; __VA_ARGS__; } break;                                \
    case 4: { constexpr int P = 4; __VA_ARGS__; } break;                                \
    case 5: { constexpr int P = 5; __VA_ARGS__; } break;                                \
    case 6: { constexpr int P = 6; __VA_ARGS__; } break;                                \
    case 7: { constexpr int P = 7; __VA_ARGS__; } break;                                \
    case 8: { constexpr int P = 8; __VA_ARGS__; } break;                                \
    default:                                                                     \
      return ::jfm::fail(JFM_ERR_UNSUPPORTED, "order " + std::to_string(p) +     \
                                                  " is not compiled (1..8)");    \
  }

#define JFM_FOR_DTYPE(dtype, ...)                 \
  if ((dtype) == JFM_F32) {                        \
    using T = float;                               \
    __VA_ARGS__;                                   \
  } else if ((dtype) == JFM_F64) {                 \
    using T = double;                              \
    __VA_ARGS__;                                   \
  } else {                                         \
    return ::jfm::fail(JFM_ERR_ARG, "bad dtype");  \
  }
