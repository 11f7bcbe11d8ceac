// Learned per-pair M2L order: pair features, order-8 truncation-error labels,
// and the 8 -> 64 -> 64 -> 3 MLP (ReLU, bfloat16 weights, float32
// accumulation) with its forward, backward and Adam kernels.
//
// Features of a far pair (target t, source s), all dimensionless:
//   f0 = log2(d / (r_t + r_s))     separation over the summed bounding radii
//   f1 = log10(m_s / m_total), f2 = log10(m_t / m_total)
//   f3..f7 = log10(|M_s,n| / (m_s r_s^n) + 1e-30), n = 2..6 (|.| = L2 over m)
// Label of a pair for gear g (any three increasing gears in 1..8; gear 8 is
// the reference, label = floor): the relative size of the order-g truncation
// error of the pair's force against the order-8 translation over the target
// sphere,
//   e_g = sqrt(sum_n w_n^2 N_n(L8 - Lg)^2) / sqrt(sum_n w_n^2 N_n(L8)^2),
//   w_n = n r_t^(n-1),  N_n(L)^2 = sum_m (L_n^m)^2 / ((n+|m|)! (n-|m|)!) x (1 if m = 0 else 1/2),
// where N_n is the rotation-invariant size of a degree-n local term in this
// real basis and w_n the size of a degree-n gradient at radius r_t; stored as
// log10(max(e_g, 1e-7)).
//
// MLP parameters (float32 masters, flat): W1[64][8] b1[64] W2[64][64] b2[64]
// W3[3][64] b3[3] = 4931 values; the kernels read a bfloat16 copy and
// accumulate in float32.  Backward: per workgroup the per-pair activations
// and deltas of 32 pairs at a time are staged in LDS and every thread owns a
// fixed set of parameters, so the per-block partial gradients and their
// final block-ordered reduction are deterministic.

#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "launch.hpp"
#include "rows.hpp"

namespace jfm {
namespace {

constexpr int kIn = 8, kHid = 64, kOut = 3;
constexpr int oW1 = 0, oB1 = oW1 + kHid * kIn, oW2 = oB1 + kHid, oB2 = oW2 + kHid * kHid, oW3 = oB2 + kHid,
              oB3 = oW3 + kOut * kHid, kParams = oB3 + kOut;
static_assert(kParams == 4931, "parameter layout");

constexpr int kThreads = 256;
constexpr int kWave = 64;
constexpr int kSub = 32;  // pairs staged per backward sub-batch

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

__device__ __forceinline__ uint16_t f32_to_bf16(float f) {  // round to nearest even (finite inputs)
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}

// ---------------------------------------------------------------- features
template <int P>
__global__ __launch_bounds__(kThreads) void features_kernel(int32_t K, const int32_t* __restrict__ tgt_of,
                                                            const int32_t* __restrict__ far_src,
                                                            const float* __restrict__ mult, int32_t cs,
                                                            const float* __restrict__ centers4,
                                                            const float* __restrict__ radius, float inv_total,
                                                            float* __restrict__ feat) {
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= K) return;
  const int t = tgt_of[k], s = far_src[k];
  const V4<float> ct = load4(centers4, t), cs4 = load4(centers4, s);
  const float dx = ct.x - cs4.x, dy = ct.y - cs4.y, dz = ct.z - cs4.z;
  const float d = sqrtf(dx * dx + dy * dy + dz * dz);
  const float rt = radius[t], rs = radius[s];
  float f[kIn];
  f[0] = log2f(d / fmaxf(rt + rs, 1e-30f));
  f[1] = log10f(fmaxf(cs4.w * inv_total, 1e-30f));
  f[2] = log10f(fmaxf(ct.w * inv_total, 1e-30f));
  const float* row = mult + static_cast<int64_t>(s) * cs;
  float rn = rs * rs;
  for (int n = 2; n <= 6; ++n) {
    float sq = 0.0f;
    if (n <= P)
      for (int m = -n; m <= n; ++m) {
        const float v = row[sh_idx(n, m)];
        sq += v * v;
      }
    f[n + 1] = log10f(sqrtf(sq) / fmaxf(cs4.w * rn, 1e-30f) + 1e-30f);
    rn *= rs;
  }
  for (int j = 0; j < kIn; ++j) feat[static_cast<int64_t>(k) * kIn + j] = f[j];
}

// ------------------------------------------------------------------ labels
// sqrt(c_nm), c_nm = 1 / ((n+|m|)! (n-|m|)!) x (1 if m = 0 else 1/2)
__device__ __forceinline__ float degree_norm_root(int n, int m) {
  const int a = m < 0 ? -m : m;
  const double c = (a == 0 ? 1.0 : 0.5) / (kFact[n + a] * kFact[n - a]);
  return static_cast<float>(sqrt(c));
}

template <int G>
__device__ __forceinline__ void m2l_order(const float* M, const Align<float, G>& a, float* L) {
  constexpr int C = ncoef(G);
  float v[C];
  for (int j = 0; j < C; ++j) v[j] = M[j];
  rotate_all<float, G, Rot::ToZMultipole>(v, a);
  zm2l<float, G>(v, L, a.r);
  rotate_all<float, G, Rot::FromZLocal>(L, a);
}

// Weighted truncation error of gear G against the order-8 local L8 of one pair.
template <int G>
__device__ __forceinline__ float label_error(const float* M, float dx, float dy, float dz, const float* L8,
                                             const float* w) {
  float Lg[ncoef(G)];
  Align<float, G> a;
  a.init(dx, dy, dz);
  m2l_order<G>(M, a, Lg);
  float e = 0.0f;
  for (int n = 0; n <= 8; ++n)
    for (int m = -n; m <= n; ++m) {
      const float lg = (n <= G) ? Lg[sh_idx(n, m)] : 0.0f;
      const float q = w[n] * degree_norm_root(n, m) * (L8[sh_idx(n, m)] - lg);
      e += q * q;
    }
  return e;
}

// labels[k][j] for gears g0 < g1 < g2 (each 1..8; gear 8 is the reference
// itself, so its label is the floor)
__global__ __launch_bounds__(64) void labels_kernel(int32_t K, const int32_t* __restrict__ tgt_of,
                                                    const int32_t* __restrict__ far_src,
                                                    const float* __restrict__ mult8, int32_t cs8,
                                                    const float* __restrict__ centers4,
                                                    const float* __restrict__ radius, int32_t g0, int32_t g1,
                                                    int32_t g2, float* __restrict__ labels) {
  const int k = blockIdx.x * 64 + threadIdx.x;
  if (k >= K) return;
  const int t = tgt_of[k], s = far_src[k];
  const V4<float> ct = load4(centers4, t), cs4 = load4(centers4, s);
  const float dx = ct.x - cs4.x, dy = ct.y - cs4.y, dz = ct.z - cs4.z;
  float M[ncoef(8)];
  load_row<float, ncoef(8)>(mult8 + static_cast<int64_t>(s) * cs8, M);
  float L8[ncoef(8)];
  {
    Align<float, 8> a;
    a.init(dx, dy, dz);
    m2l_order<8>(M, a, L8);
  }
  const float rt = radius[t];
  float w[9];  // n r_t^(n-1)
  float rp = 1.0f;
  w[0] = 0.0f;
  for (int n = 1; n <= 8; ++n) {
    w[n] = static_cast<float>(n) * rp;
    rp *= rt;
  }
  float ref = 0.0f;
  for (int n = 0; n <= 8; ++n)
    for (int m = -n; m <= n; ++m) {
      const float q = w[n] * degree_norm_root(n, m) * L8[sh_idx(n, m)];
      ref += q * q;
    }
  ref = sqrtf(fmaxf(ref, 1e-37f));
  const int gs[3] = {g0, g1, g2};
  for (int j = 0; j < 3; ++j) {
    float e = 0.0f;
    switch (gs[j]) {  // uniform
      case 1: e = label_error<1>(M, dx, dy, dz, L8, w); break;
      case 2: e = label_error<2>(M, dx, dy, dz, L8, w); break;
      case 3: e = label_error<3>(M, dx, dy, dz, L8, w); break;
      case 4: e = label_error<4>(M, dx, dy, dz, L8, w); break;
      case 5: e = label_error<5>(M, dx, dy, dz, L8, w); break;
      case 6: e = label_error<6>(M, dx, dy, dz, L8, w); break;
      case 7: e = label_error<7>(M, dx, dy, dz, L8, w); break;
      default: e = 0.0f; break;
    }
    labels[static_cast<int64_t>(k) * 3 + j] = log10f(fmaxf(sqrtf(e) / ref, 1e-7f));
  }
}

// --------------------------------------------------------------------- MLP
__device__ __forceinline__ void stage_weights(const uint16_t* __restrict__ wbf, float* w) {
  for (int j = threadIdx.x; j < kParams; j += blockDim.x) w[j] = bf16_to_f32(wbf[j]);
}

__device__ __forceinline__ void mlp_forward_one(const float* w, const float* x, float* h1, float* h2, float* y) {
#pragma unroll 4
  for (int i = 0; i < kHid; ++i) {
    float a = w[oB1 + i];
#pragma unroll
    for (int j = 0; j < kIn; ++j) a = fmaf(w[oW1 + i * kIn + j], x[j], a);
    h1[i] = fmaxf(a, 0.0f);
  }
#pragma unroll 2
  for (int i = 0; i < kHid; ++i) {
    float a = w[oB2 + i];
#pragma unroll
    for (int j = 0; j < kHid; ++j) a = fmaf(w[oW2 + i * kHid + j], h1[j], a);
    h2[i] = fmaxf(a, 0.0f);
  }
#pragma unroll
  for (int i = 0; i < kOut; ++i) {
    float a = w[oB3 + i];
#pragma unroll
    for (int j = 0; j < kHid; ++j) a = fmaf(w[oW3 + i * kHid + j], h2[j], a);
    y[i] = a;
  }
}

// mean / inv_std (optional, [kIn]): features are normalised on the fly,
// x_j = (X_j - mean_j) * inv_std_j.
__global__ __launch_bounds__(kThreads) void mlp_forward_kernel(int32_t K, const float* __restrict__ X,
                                                               const float* __restrict__ mean,
                                                               const float* __restrict__ inv_std,
                                                               const uint16_t* __restrict__ wbf,
                                                               float* __restrict__ Y) {
  __shared__ float w[kParams];
  stage_weights(wbf, w);
  __syncthreads();
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= K) return;
  float x[kIn], h1[kHid], h2[kHid], y[kOut];
  for (int j = 0; j < kIn; ++j) x[j] = X[static_cast<int64_t>(k) * kIn + j];
  if (mean)
    for (int j = 0; j < kIn; ++j) x[j] = (x[j] - mean[j]) * inv_std[j];
  mlp_forward_one(w, x, h1, h2, y);
  for (int j = 0; j < kOut; ++j) Y[static_cast<int64_t>(k) * kOut + j] = y[j];
}

// Forward on the matrix cores (bf16 MFMA, fp32 accumulation), 32 pairs per
// wave tile, layers computed transposed: H1^T = relu(W1 X^T + b1),
// H2^T = relu(W2 H1^T + b2), Y^T = W3 H2^T + b3 with v_mfma_f32_32x32x8bf16.
// Operand layout of that instruction (64 lanes, h = lane / 32):
//   A[i][k]: lane i + 32 h holds k = 4h .. 4h+3;  B[k][j]: lane j + 32 h holds
//   k = 4h .. 4h+3;  C[i][j]: lane j + 32 h, element v holds i = 8(v/4) + 4h + v%4,
// so the 16 outputs a lane holds for pair j are exactly the k-rows the next
// layer's B operand needs from it (rows 8r + 4h .. +3 for k-block r): the
// activations never leave registers.  Activations are rounded to bf16 for the
// next product (the weights are bf16 in every path).
typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

__device__ __forceinline__ bf16x4_t pack_bf16(float a, float b, float c, float d) {
  bf16x4_t r;
  r[0] = static_cast<short>(f32_to_bf16(a));
  r[1] = static_cast<short>(f32_to_bf16(b));
  r[2] = static_cast<short>(f32_to_bf16(c));
  r[3] = static_cast<short>(f32_to_bf16(d));
  return r;
}

// 4 consecutive bf16 weights W[row][col .. col+3] of a row-major block, or
// zeros for rows >= nrows (padding of the 3-output layer).
__device__ __forceinline__ bf16x4_t weight_frag(const uint16_t* __restrict__ wbf, int off, int ncols, int row,
                                                int col, int nrows) {
  bf16x4_t r = {0, 0, 0, 0};
  if (row < nrows)
    for (int q = 0; q < 4; ++q) r[q] = static_cast<short>(wbf[off + row * ncols + col + q]);
  return r;
}

__global__ __launch_bounds__(kThreads) void mlp_forward_mfma_kernel(int32_t K, const float* __restrict__ X,
                                                                    const float* __restrict__ mean,
                                                                    const float* __restrict__ inv_std,
                                                                    const uint16_t* __restrict__ wbf,
                                                                    float* __restrict__ Y) {
  const int lane = threadIdx.x % kWave;
  const int h = lane / 32;
  const int c32 = lane % 32;
  // weight fragments (A operands), loaded once per wave
  bf16x4_t a1[2], a2[2][8], a3[8];
  for (int ob = 0; ob < 2; ++ob) a1[ob] = weight_frag(wbf, oW1, kIn, ob * 32 + c32, 4 * h, kHid);
  for (int ob = 0; ob < 2; ++ob)
    for (int kb = 0; kb < 8; ++kb) a2[ob][kb] = weight_frag(wbf, oW2, kHid, ob * 32 + c32, kb * 8 + 4 * h, kHid);
  for (int kb = 0; kb < 8; ++kb) a3[kb] = weight_frag(wbf, oW3, kHid, c32, kb * 8 + 4 * h, kOut);
  // biases in the C layout: element v of block ob is row ob*32 + 8(v/4) + 4h + v%4
  float bias1[2][16], bias2[2][16], bias3[16];
  for (int v = 0; v < 16; ++v) {
    const int r = 8 * (v / 4) + 4 * h + (v % 4);
    for (int ob = 0; ob < 2; ++ob) {
      bias1[ob][v] = bf16_to_f32(wbf[oB1 + ob * 32 + r]);
      bias2[ob][v] = bf16_to_f32(wbf[oB2 + ob * 32 + r]);
    }
    bias3[v] = r < kOut ? bf16_to_f32(wbf[oB3 + r]) : 0.0f;
  }
  float mu[4] = {0, 0, 0, 0}, is[4] = {1, 1, 1, 1};
  if (mean)
    for (int q = 0; q < 4; ++q) {
      mu[q] = mean[4 * h + q];
      is[q] = inv_std[4 * h + q];
    }

  const int tiles = (K + 31) / 32;
  const int waves = gridDim.x * (kThreads / kWave);
  for (int t = blockIdx.x * (kThreads / kWave) + static_cast<int>(threadIdx.x) / kWave; t < tiles; t += waves) {
    const int pair = t * 32 + c32;
    // B operand of layer 1: features 4h .. 4h+3 of this lane's pair
    float x[4] = {0, 0, 0, 0};
    if (pair < K)
      for (int q = 0; q < 4; ++q) x[q] = (X[static_cast<int64_t>(pair) * kIn + 4 * h + q] - mu[q]) * is[q];
    const bf16x4_t bx = pack_bf16(x[0], x[1], x[2], x[3]);
    // layer 1
    bf16x4_t b2op[8];
    for (int ob = 0; ob < 2; ++ob) {
      f32x16_t c;
      for (int v = 0; v < 16; ++v) c[v] = bias1[ob][v];
      c = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a1[ob], bx, c, 0, 0, 0);
      for (int r = 0; r < 4; ++r)
        b2op[ob * 4 + r] = pack_bf16(fmaxf(c[4 * r], 0.0f), fmaxf(c[4 * r + 1], 0.0f), fmaxf(c[4 * r + 2], 0.0f),
                                     fmaxf(c[4 * r + 3], 0.0f));
    }
    // layer 2
    bf16x4_t b3op[8];
    for (int ob = 0; ob < 2; ++ob) {
      f32x16_t c;
      for (int v = 0; v < 16; ++v) c[v] = bias2[ob][v];
      for (int kb = 0; kb < 8; ++kb) c = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a2[ob][kb], b2op[kb], c, 0, 0, 0);
      for (int r = 0; r < 4; ++r)
        b3op[ob * 4 + r] = pack_bf16(fmaxf(c[4 * r], 0.0f), fmaxf(c[4 * r + 1], 0.0f), fmaxf(c[4 * r + 2], 0.0f),
                                     fmaxf(c[4 * r + 3], 0.0f));
    }
    // layer 3 (rows 0..2 of the output block; rows 3..31 are padding)
    f32x16_t c;
    for (int v = 0; v < 16; ++v) c[v] = bias3[v];
    for (int kb = 0; kb < 8; ++kb) c = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a3[kb], b3op[kb], c, 0, 0, 0);
    if (h == 0 && pair < K)
      for (int v = 0; v < kOut; ++v) Y[static_cast<int64_t>(pair) * kOut + v] = c[v];
  }
}

// Per-workgroup partial gradient of loss = mean_k mean_j (y_kj - t_kj)^2 over
// pairs [blockIdx.x * per_block, ...).  partial[block][kParams + 1] (last = loss sum).
__global__ __launch_bounds__(kThreads) void mlp_backward_kernel(int32_t K, int32_t per_block,
                                                                const float* __restrict__ X,
                                                                const float* __restrict__ Tgt,
                                                                const uint16_t* __restrict__ wbf,
                                                                float* __restrict__ partial) {
  __shared__ float w[kParams];
  __shared__ float s_x[kSub][kIn], s_h1[kSub][kHid], s_h2[kSub][kHid], s_d1[kSub][kHid], s_d2[kSub][kHid],
      s_dy[kSub][kOut], s_loss[kSub];
  stage_weights(wbf, w);
  __syncthreads();
  const float scale = 2.0f / (static_cast<float>(K) * kOut);
  // each thread owns parameters j = threadIdx.x + kThreads * q
  constexpr int kOwn = (kParams + kThreads - 1) / kThreads;
  float g[kOwn];
  for (int q = 0; q < kOwn; ++q) g[q] = 0.0f;
  float loss = 0.0f;
  const int b0 = blockIdx.x * per_block;
  const int b1 = min(K, b0 + per_block);
  for (int base = b0; base < b1; base += kSub) {
    const int cnt = min(kSub, b1 - base);
    if (threadIdx.x < kSub) {
      const int r = threadIdx.x;
      if (r < cnt) {
        const int k = base + r;
        float x[kIn], h1[kHid], h2[kHid], y[kOut];
        for (int j = 0; j < kIn; ++j) x[j] = X[static_cast<int64_t>(k) * kIn + j];
        mlp_forward_one(w, x, h1, h2, y);
        float dy[kOut], l = 0.0f;
        for (int j = 0; j < kOut; ++j) {
          const float e = y[j] - Tgt[static_cast<int64_t>(k) * kOut + j];
          l += e * e;
          dy[j] = scale * e;
        }
        float d2[kHid], d1[kHid];
        for (int i = 0; i < kHid; ++i) {
          float a = 0.0f;
          for (int j = 0; j < kOut; ++j) a = fmaf(w[oW3 + j * kHid + i], dy[j], a);
          d2[i] = h2[i] > 0.0f ? a : 0.0f;
        }
        for (int i = 0; i < kHid; ++i) {
          float a = 0.0f;
          for (int j = 0; j < kHid; ++j) a = fmaf(w[oW2 + j * kHid + i], d2[j], a);
          d1[i] = h1[i] > 0.0f ? a : 0.0f;
        }
        for (int j = 0; j < kIn; ++j) s_x[r][j] = x[j];
        for (int i = 0; i < kHid; ++i) {
          s_h1[r][i] = h1[i];
          s_h2[r][i] = h2[i];
          s_d1[r][i] = d1[i];
          s_d2[r][i] = d2[i];
        }
        for (int j = 0; j < kOut; ++j) s_dy[r][j] = dy[j];
        s_loss[r] = l;
      } else {
        for (int j = 0; j < kIn; ++j) s_x[r][j] = 0.0f;
        for (int i = 0; i < kHid; ++i) s_h1[r][i] = s_h2[r][i] = s_d1[r][i] = s_d2[r][i] = 0.0f;
        for (int j = 0; j < kOut; ++j) s_dy[r][j] = 0.0f;
        s_loss[r] = 0.0f;
      }
    }
    __syncthreads();
    for (int q = 0; q < kOwn; ++q) {
      const int j = threadIdx.x + kThreads * q;
      if (j >= kParams) break;
      float a = 0.0f;
      if (j < oB1) {
        const int i = j / kIn, c = j % kIn;
        for (int r = 0; r < kSub; ++r) a = fmaf(s_d1[r][i], s_x[r][c], a);
      } else if (j < oW2) {
        const int i = j - oB1;
        for (int r = 0; r < kSub; ++r) a += s_d1[r][i];
      } else if (j < oB2) {
        const int i = (j - oW2) / kHid, c = (j - oW2) % kHid;
        for (int r = 0; r < kSub; ++r) a = fmaf(s_d2[r][i], s_h1[r][c], a);
      } else if (j < oW3) {
        const int i = j - oB2;
        for (int r = 0; r < kSub; ++r) a += s_d2[r][i];
      } else if (j < oB3) {
        const int i = (j - oW3) / kHid, c = (j - oW3) % kHid;
        for (int r = 0; r < kSub; ++r) a = fmaf(s_dy[r][i], s_h2[r][c], a);
      } else {
        const int i = j - oB3;
        for (int r = 0; r < kSub; ++r) a += s_dy[r][i];
      }
      g[q] += a;
    }
    if (threadIdx.x == 0)
      for (int r = 0; r < kSub; ++r) loss += s_loss[r];
    __syncthreads();
  }
  float* out = partial + static_cast<int64_t>(blockIdx.x) * (kParams + 1);
  for (int q = 0; q < kOwn; ++q) {
    const int j = threadIdx.x + kThreads * q;
    if (j < kParams) out[j] = g[q];
  }
  if (threadIdx.x == 0) out[kParams] = loss / (static_cast<float>(K) * kOut);
}

__global__ __launch_bounds__(kThreads) void reduce_blocks_kernel(const float* __restrict__ partial, int32_t nblk,
                                                                 float* __restrict__ grad) {
  const int j = blockIdx.x * kThreads + threadIdx.x;
  if (j > kParams) return;
  float s = 0.0f;
  for (int b = 0; b < nblk; ++b) s += partial[static_cast<int64_t>(b) * (kParams + 1) + j];
  grad[j] = s;
}

__global__ __launch_bounds__(kThreads) void adam_kernel(int32_t n, float* __restrict__ master,
                                                        const float* __restrict__ grad, float* __restrict__ m1,
                                                        float* __restrict__ m2, float lr, float beta1, float beta2,
                                                        float eps, float bc1, float bc2, uint16_t* __restrict__ wbf) {
  const int j = blockIdx.x * kThreads + threadIdx.x;
  if (j >= n) return;
  const float g = grad[j];
  const float a = beta1 * m1[j] + (1.0f - beta1) * g;
  const float b = beta2 * m2[j] + (1.0f - beta2) * g * g;
  m1[j] = a;
  m2[j] = b;
  const float upd = lr * (a / bc1) / (sqrtf(b / bc2) + eps);
  const float w = master[j] - upd;
  master[j] = w;
  wbf[j] = f32_to_bf16(w);
}

__global__ __launch_bounds__(kThreads) void to_bf16_kernel(int32_t n, const float* __restrict__ src,
                                                           uint16_t* __restrict__ dst) {
  const int j = blockIdx.x * kThreads + threadIdx.x;
  if (j < n) dst[j] = f32_to_bf16(src[j]);
}

// Per target node t (one thread), over its far pairs k in [far_off[t], far_off[t+1]):
// w_k = log10(m_s / d^2) scales a pair's predicted relative error y_k,g to an
// absolute force-error proxy.  The budget of t is the largest proxy of the
// top gear over its pairs (the fixed-order method's dominant error term;
// for a top gear of 8, y_8 = y_g1 + (8 - g1) / (g1 - g0) (y_g1 - y_g0));
// pair k takes the smallest gear with y_k,g + w_k <= budget + tol, the top
// gear otherwise.
__global__ __launch_bounds__(kThreads) void select_gears_kernel(int32_t num_nodes, const int32_t* __restrict__ far_off,
                                                                const int32_t* __restrict__ far_src,
                                                                const float* __restrict__ Y,
                                                                const float* __restrict__ centers4, float tol,
                                                                int32_t g0, int32_t g1, int32_t g2,
                                                                int8_t* __restrict__ gears) {
  // one wave per target node, lanes over its far pairs (far lists range from
  // a few pairs to thousands; a thread per node left the long ones running
  // alone); the budget is a wave maximum, exact in any order
  const int t = blockIdx.x * (kThreads / kWave) + static_cast<int>(threadIdx.x) / kWave;
  const int lane = static_cast<int>(threadIdx.x) % kWave;
  if (t >= num_nodes) return;  // wave-uniform
  const int lo = far_off[t], hi = far_off[t + 1];
  if (lo == hi) return;  // wave-uniform
  const V4<float> ct = load4(centers4, t);
  auto weight = [&](int k) {
    const V4<float> cs4 = load4(centers4, far_src[k]);
    const float dx = ct.x - cs4.x, dy = ct.y - cs4.y, dz = ct.z - cs4.z;
    const float d2 = fmaxf(dx * dx + dy * dy + dz * dz, 1e-37f);
    return log10f(fmaxf(cs4.w, 1e-37f) / d2);
  };
  // the top gear's predicted error; a top gear of 8 is the label reference
  // (its label is the floor), so its error is extrapolated geometrically in
  // the order from the two lower gears
  const bool extrap = g2 >= 8;
  const float slope = static_cast<float>(g2 - g1) / static_cast<float>(g1 - g0);
  auto top = [&](int k) {
    const float* y = Y + static_cast<int64_t>(k) * kOut;
    return extrap ? y[1] + slope * fminf(y[1] - y[0], 0.0f) : y[2];
  };
  float budget = -INFINITY;
  for (int k = lo + lane; k < hi; k += kWave) budget = fmaxf(budget, top(k) + weight(k));
  for (int o = kWave / 2; o >= 1; o >>= 1) budget = fmaxf(budget, __shfl_xor(budget, o, kWave));
  for (int k = lo + lane; k < hi; k += kWave) {
    const float thr = budget + tol - weight(k);
    const float* y = Y + static_cast<int64_t>(k) * kOut;
    int g = g2;
    if (y[1] <= thr) g = g1;
    if (y[0] <= thr) g = g0;
    gears[k] = static_cast<int8_t>(g);
  }
}

// Particle-level budget (gear_budget="leaf"): the per-node maximum top-gear
// proxy (node_max_kernel), then along every root-to-leaf path the largest of
// them (the dominant truncation term a particle of that leaf receives from
// any level, path_max_kernel root first), then for every node the smallest of
// its leaves' path maxima (subtree_min_kernel, deepest first): a pair of
// target t may then err up to 10^tol of the dominant term of every particle
// under t.
__global__ __launch_bounds__(kThreads) void node_max_kernel(int32_t num_nodes, const int32_t* __restrict__ far_off,
                                                            const int32_t* __restrict__ far_src,
                                                            const float* __restrict__ Y,
                                                            const float* __restrict__ centers4, int32_t g0,
                                                            int32_t g1, int32_t g2, float* __restrict__ node_max) {
  const int t = blockIdx.x * (kThreads / kWave) + static_cast<int>(threadIdx.x) / kWave;
  const int lane = static_cast<int>(threadIdx.x) % kWave;
  if (t >= num_nodes) return;  // wave-uniform
  const int lo = far_off[t], hi = far_off[t + 1];
  const V4<float> ct = load4(centers4, t);
  const bool extrap = g2 >= 8;
  const float slope = static_cast<float>(g2 - g1) / static_cast<float>(g1 - g0);
  float m = -INFINITY;
  for (int k = lo + lane; k < hi; k += kWave) {
    const V4<float> cs4 = load4(centers4, far_src[k]);
    const float dx = ct.x - cs4.x, dy = ct.y - cs4.y, dz = ct.z - cs4.z;
    const float d2 = fmaxf(dx * dx + dy * dy + dz * dz, 1e-37f);
    const float* y = Y + static_cast<int64_t>(k) * kOut;
    const float top = extrap ? y[1] + slope * fminf(y[1] - y[0], 0.0f) : y[2];
    m = fmaxf(m, top + log10f(fmaxf(cs4.w, 1e-37f) / d2));
  }
  for (int o = kWave / 2; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, kWave));
  if (lane == 0) node_max[t] = m;
}

__global__ __launch_bounds__(kThreads) void path_max_kernel(const int32_t* __restrict__ nodes, int32_t count,
                                                            const int32_t* __restrict__ parent,
                                                            float* __restrict__ v) {
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= count) return;
  const int u = nodes[k];
  v[u] = fmaxf(v[u], v[parent[u]]);
}

__global__ __launch_bounds__(kThreads) void subtree_min_kernel(const int32_t* __restrict__ nodes, int32_t count,
                                                               const int32_t* __restrict__ left,
                                                               const int32_t* __restrict__ right,
                                                               float* __restrict__ v) {
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= count) return;
  const int u = nodes[k];
  v[u] = fminf(v[left[u]], v[right[u]]);
}

// select_gears_kernel with the budget of every target given
__global__ __launch_bounds__(kThreads) void select_budget_kernel(int32_t num_nodes, const int32_t* __restrict__ far_off,
                                                                 const int32_t* __restrict__ far_src,
                                                                 const float* __restrict__ Y,
                                                                 const float* __restrict__ centers4, float tol,
                                                                 int32_t g0, int32_t g1, int32_t g2,
                                                                 const float* __restrict__ budget,
                                                                 int8_t* __restrict__ gears) {
  const int t = blockIdx.x * (kThreads / kWave) + static_cast<int>(threadIdx.x) / kWave;
  const int lane = static_cast<int>(threadIdx.x) % kWave;
  if (t >= num_nodes) return;
  const int lo = far_off[t], hi = far_off[t + 1];
  const V4<float> ct = load4(centers4, t);
  const float b = budget[t];
  for (int k = lo + lane; k < hi; k += kWave) {
    const V4<float> cs4 = load4(centers4, far_src[k]);
    const float dx = ct.x - cs4.x, dy = ct.y - cs4.y, dz = ct.z - cs4.z;
    const float d2 = fmaxf(dx * dx + dy * dy + dz * dz, 1e-37f);
    const float thr = b + tol - log10f(fmaxf(cs4.w, 1e-37f) / d2);
    const float* y = Y + static_cast<int64_t>(k) * kOut;
    int g = g2;
    if (y[1] <= thr) g = g1;
    if (y[0] <= thr) g = g0;
    gears[k] = static_cast<int8_t>(g);
  }
}

inline int nblocks(int64_t n, int b) { return static_cast<int>((n + b - 1) / b); }

}  // namespace
}  // namespace jfm

using namespace jfm;

extern "C" {

int32_t jfm_mlp_num_params(void) { return kParams; }

int64_t jfm_mlp_backward_scratch_floats(int32_t K, int32_t per_block) {
  if (K <= 0 || per_block <= 0) return 0;
  return static_cast<int64_t>(nblocks(K, per_block)) * (kParams + 1);
}

int32_t jfm_pair_features(void* stream, int32_t p, int32_t K, const int32_t* tgt_of, const int32_t* far_src,
                          const void* multipoles, const void* centers4, const float* radius, double total_mass,
                          float* features) {
  if (K < 0 || !(total_mass > 0.0)) return fail(JFM_ERR_ARG, "jfm_pair_features: bad sizes");
  if (K == 0) return JFM_OK;
  if (!tgt_of || !far_src || !multipoles || !centers4 || !radius || !features)
    return fail(JFM_ERR_ARG, "jfm_pair_features: NULL buffer");
  hipStream_t s = as_stream(stream);
  JFM_FOR_ORDER(p, {
    features_kernel<P><<<nblocks(K, kThreads), kThreads, 0, s>>>(K, tgt_of, far_src, (const float*)multipoles,
                                                                 coef_stride(P, 4), (const float*)centers4, radius,
                                                                 static_cast<float>(1.0 / total_mass), features);
  });
  return check_launch("jfm_pair_features");
}

int32_t jfm_pair_labels(void* stream, int32_t K, const int32_t* tgt_of, const int32_t* far_src,
                        const void* multipoles8, const void* centers4, const float* radius, const int32_t* gears3,
                        float* labels) {
  if (K < 0 || !gears3) return fail(JFM_ERR_ARG, "jfm_pair_labels: bad arguments");
  if (!(1 <= gears3[0] && gears3[0] < gears3[1] && gears3[1] < gears3[2] && gears3[2] <= 8))
    return fail(JFM_ERR_ARG, "jfm_pair_labels: gears must be increasing within 1..8");
  if (K == 0) return JFM_OK;
  if (!tgt_of || !far_src || !multipoles8 || !centers4 || !radius || !labels)
    return fail(JFM_ERR_ARG, "jfm_pair_labels: NULL buffer");
  labels_kernel<<<nblocks(K, 64), 64, 0, as_stream(stream)>>>(K, tgt_of, far_src, (const float*)multipoles8,
                                                               coef_stride(8, 4), (const float*)centers4, radius,
                                                               gears3[0], gears3[1], gears3[2], labels);
  return check_launch("jfm_pair_labels");
}

int32_t jfm_mlp_forward(void* stream, int32_t K, const float* X, const float* mean, const float* inv_std,
                        const uint16_t* wbf16, float* Y) {
  if (K < 0) return fail(JFM_ERR_ARG, "jfm_mlp_forward: bad sizes");
  if (K == 0) return JFM_OK;
  if (!X || !wbf16 || !Y) return fail(JFM_ERR_ARG, "jfm_mlp_forward: NULL buffer");
  if ((mean == nullptr) != (inv_std == nullptr)) return fail(JFM_ERR_ARG, "jfm_mlp_forward: mean without inv_std");
  mlp_forward_kernel<<<nblocks(K, kThreads), kThreads, 0, as_stream(stream)>>>(K, X, mean, inv_std, wbf16, Y);
  return check_launch("jfm_mlp_forward");
}

int32_t jfm_mlp_forward_mfma(void* stream, int32_t K, const float* X, const float* mean, const float* inv_std,
                             const uint16_t* wbf16, float* Y) {
  if (K < 0) return fail(JFM_ERR_ARG, "jfm_mlp_forward_mfma: bad sizes");
  if (K == 0) return JFM_OK;
  if (!X || !wbf16 || !Y) return fail(JFM_ERR_ARG, "jfm_mlp_forward_mfma: NULL buffer");
  if ((mean == nullptr) != (inv_std == nullptr)) return fail(JFM_ERR_ARG, "jfm_mlp_forward_mfma: mean without inv_std");
  const int tiles = (K + 31) / 32;
  const int blocks = min((tiles + 3) / 4, 256 * 8);  // 4 waves per workgroup, tiles grid-strided
  mlp_forward_mfma_kernel<<<blocks, kThreads, 0, as_stream(stream)>>>(K, X, mean, inv_std, wbf16, Y);
  return check_launch("jfm_mlp_forward_mfma");
}

int32_t jfm_mlp_backward(void* stream, int32_t K, int32_t per_block, const float* X, const float* targets,
                         const uint16_t* wbf16, float* scratch, float* grad_and_loss) {
  if (K <= 0 || per_block <= 0) return fail(JFM_ERR_ARG, "jfm_mlp_backward: bad sizes");
  if (!X || !targets || !wbf16 || !scratch || !grad_and_loss) return fail(JFM_ERR_ARG, "jfm_mlp_backward: NULL");
  hipStream_t s = as_stream(stream);
  const int nb = nblocks(K, per_block);
  mlp_backward_kernel<<<nb, kThreads, 0, s>>>(K, per_block, X, targets, wbf16, scratch);
  int32_t rc = check_launch("jfm_mlp_backward");
  if (rc != JFM_OK) return rc;
  reduce_blocks_kernel<<<nblocks(kParams + 1, kThreads), kThreads, 0, s>>>(scratch, nb, grad_and_loss);
  return check_launch("jfm_mlp_backward");
}

int32_t jfm_mlp_adam(void* stream, int32_t step, float lr, float beta1, float beta2, float eps, float* master,
                     const float* grad, float* m1, float* m2, uint16_t* wbf16) {
  if (step < 1) return fail(JFM_ERR_ARG, "jfm_mlp_adam: step must be >= 1");
  if (!master || !grad || !m1 || !m2 || !wbf16) return fail(JFM_ERR_ARG, "jfm_mlp_adam: NULL buffer");
  const float bc1 = 1.0f - powf(beta1, static_cast<float>(step));
  const float bc2 = 1.0f - powf(beta2, static_cast<float>(step));
  adam_kernel<<<nblocks(kParams, kThreads), kThreads, 0, as_stream(stream)>>>(kParams, master, grad, m1, m2, lr,
                                                                             beta1, beta2, eps, bc1, bc2, wbf16);
  return check_launch("jfm_mlp_adam");
}

int32_t jfm_mlp_to_bf16(void* stream, const float* master, uint16_t* wbf16) {
  if (!master || !wbf16) return fail(JFM_ERR_ARG, "jfm_mlp_to_bf16: NULL buffer");
  to_bf16_kernel<<<nblocks(kParams, kThreads), kThreads, 0, as_stream(stream)>>>(kParams, master, wbf16);
  return check_launch("jfm_mlp_to_bf16");
}

int32_t jfm_select_gears(void* stream, int32_t num_nodes, const int32_t* far_off, const int32_t* far_src,
                         const float* predictions, const void* centers4, float tol_log10, const int32_t* gears3,
                         int8_t* gears) {
  if (num_nodes < 0 || !gears3) return fail(JFM_ERR_ARG, "jfm_select_gears: bad arguments");
  if (num_nodes == 0) return JFM_OK;
  if (!far_off || !far_src || !predictions || !centers4 || !gears)
    return fail(JFM_ERR_ARG, "jfm_select_gears: NULL buffer");
  select_gears_kernel<<<nblocks(num_nodes, kThreads / kWave), kThreads, 0, as_stream(stream)>>>(
      num_nodes, far_off, far_src, predictions, (const float*)centers4, tol_log10, gears3[0], gears3[1], gears3[2],
      gears);
  return check_launch("jfm_select_gears");
}

int32_t jfm_gear_node_max(void* stream, int32_t num_nodes, const int32_t* far_off, const int32_t* far_src,
                          const float* predictions, const void* centers4, const int32_t* gears3, float* node_max) {
  if (num_nodes < 1 || !far_off || !far_src || !predictions || !centers4 || !gears3 || !node_max)
    return fail(JFM_ERR_ARG, "jfm_gear_node_max: bad arguments");
  if (!(gears3[0] < gears3[1] && gears3[1] < gears3[2])) return fail(JFM_ERR_ARG, "jfm_gear_node_max: gears");
  node_max_kernel<<<nblocks(num_nodes, kThreads / kWave), kThreads, 0, as_stream(stream)>>>(
      num_nodes, far_off, far_src, predictions, (const float*)centers4, gears3[0], gears3[1], gears3[2], node_max);
  return check_launch("jfm_gear_node_max");
}

int32_t jfm_level_path_max(void* stream, const int32_t* nodes, int32_t count, const int32_t* parent, float* values) {
  if (count < 0) return fail(JFM_ERR_ARG, "jfm_level_path_max: negative count");
  if (count == 0) return JFM_OK;
  if (!nodes || !parent || !values) return fail(JFM_ERR_ARG, "jfm_level_path_max: NULL buffer");
  path_max_kernel<<<nblocks(count, kThreads), kThreads, 0, as_stream(stream)>>>(nodes, count, parent, values);
  return check_launch("jfm_level_path_max");
}

int32_t jfm_level_subtree_min(void* stream, const int32_t* nodes, int32_t count, const int32_t* left,
                              const int32_t* right, float* values) {
  if (count < 0) return fail(JFM_ERR_ARG, "jfm_level_subtree_min: negative count");
  if (count == 0) return JFM_OK;
  if (!nodes || !left || !right || !values) return fail(JFM_ERR_ARG, "jfm_level_subtree_min: NULL buffer");
  subtree_min_kernel<<<nblocks(count, kThreads), kThreads, 0, as_stream(stream)>>>(nodes, count, left, right, values);
  return check_launch("jfm_level_subtree_min");
}

int32_t jfm_select_gears_budget(void* stream, int32_t num_nodes, const int32_t* far_off, const int32_t* far_src,
                                const float* predictions, const void* centers4, float tol_log10,
                                const int32_t* gears3, const float* budget, int8_t* gears) {
  if (num_nodes < 1 || !far_off || !far_src || !predictions || !centers4 || !gears3 || !budget || !gears)
    return fail(JFM_ERR_ARG, "jfm_select_gears_budget: bad arguments");
  if (!(gears3[0] < gears3[1] && gears3[1] < gears3[2])) return fail(JFM_ERR_ARG, "jfm_select_gears_budget: gears");
  select_budget_kernel<<<nblocks(num_nodes, kThreads / kWave), kThreads, 0, as_stream(stream)>>>(
      num_nodes, far_off, far_src, predictions, (const float*)centers4, tol_log10, gears3[0], gears3[1], gears3[2],
      budget, gears);
  return check_launch("jfm_select_gears_budget");
}

}  // extern "C"
