// Convolutional subgrid closure for coarse 2D LES on the matrix cores.
//
// Network: 5x5 periodic convolutions (the domain is periodic), 32 channels,
// GELU between layers; BF16 weights and activations, FP32 accumulation on
// MFMA (v_mfma_f32_32x32x16_bf16 in the convolutions and the weight
// gradients), FP32 pre-activations, FP32 master weights with Adam.
//
// Layouts (channels last):
//   activations  x[B][H][W][C]  BF16, C = 8 or 32 (channel-padded with 0)
//   pre-act.     y[B][H][W][co] FP32
//   weights      w[32][25][C]   BF16 (output channel, tap dy*5+dx, input
//                channel), rows beyond the layer's outputs are zero
//   bias         b[32]          FP32
// Forward:  y[p][o] = b[o] + sum_{t, c} w[o][t][c] x[p + d(t)][c],
//           d(t) = (dy - 2, dx - 2) with periodic wrap.
// Input gradient: the same kernel on dy with the flipped, transposed
// weights wt[c][24 - t][o] = w[o][t][c].
// Weight gradient: dw[o][t][c] = sum_p dy[p][o] x[p + d(t)][c].
//
// MFMA 32x32x8 BF16 operand layout (64 lanes, m = lane % 32,
// g = lane / 32): lane supplies A[m][4g .. 4g+3] and B[4g .. 4g+3][m];
// it receives D[8 (i / 4) + 4 g + i % 4][m], i = 0..15.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fsx {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));  // 16-byte moves

constexpr int kCnnTaps = 25;
constexpr int kCnnRows = 4;  // output rows per workgroup of the convolutions

__device__ __forceinline__ uint16_t to_bf16(float f) {
  // round to nearest even (finite inputs)
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) {
  return __uint_as_float((uint32_t)h << 16);
}

// GELU, exact (erf) form
// erf x from e = exp(-x^2): Abramowitz & Stegun, Handbook of Mathematical
// Functions (1964), formula 7.1.26, erf x = 1 - (a1 t + a2 t^2 + a3 t^3 +
// a4 t^4 + a5 t^5) exp(-x^2) with t = 1 / (1 + p x) for x >= 0 (odd in
// x), absolute error at most 1.5e-7: one reciprocal and one exponential,
// which the GELU derivative shares with the Gaussian density.  Far below
// the BF16 rounding of the activations; the convolution epilogues that
// evaluate it 16 times per lane per tile were VALU-bound with erff.
__device__ __forceinline__ float erf_as(float x, float e) {
  const float t = __builtin_amdgcn_rcpf(fmaf(0.3275911f, fabsf(x), 1.0f));
  float p = fmaf(1.061405429f, t, -1.453152027f);
  p = fmaf(p, t, 1.421413741f);
  p = fmaf(p, t, -0.284496736f);
  p = fmaf(p, t, 0.254829592f);
  return copysignf(fmaf(-p * t, e, 1.0f), x);
}
__device__ __forceinline__ float gelu(float v) {
  const float x = v * 0.70710678118654752f;
  return 0.5f * v * (1.0f + erf_as(x, __expf(-x * x)));
}
__device__ __forceinline__ float gelu_grad(float v) {
  const float x = v * 0.70710678118654752f;
  const float e = __expf(-x * x);
  const float cdf = 0.5f * (1.0f + erf_as(x, e));
  return cdf + v * (0.39894228040143268f * e);
}

__device__ __forceinline__ f32x16 mfma_bf16(bf16x4 a, bf16x4 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
}

// D (32 x 32, row-major FP32) = A (32 x 8 BF16, row-major) x B (8 x 32 BF16,
// row-major): one wave; checks the operand layout above.
static __global__ __launch_bounds__(64) void k_mfma_probe(const uint16_t* A, const uint16_t* B,
                                                   float* D) {
  const int lane = threadIdx.x, m = lane & 31, g = lane >> 5;
  bf16x4 a, b;
  for (int j = 0; j < 4; ++j) {
    a[j] = (short)A[m * 8 + 4 * g + j];
    b[j] = (short)B[(4 * g + j) * 32 + m];
  }
  f32x16 acc = {};
  acc = mfma_bf16(a, b, acc);
  for (int i = 0; i < 16; ++i) D[(8 * (i >> 2) + 4 * g + (i & 3)) * 32 + m] = acc[i];
}

// v_mfma_f32_32x32x16_bf16 (gfx950): lane supplies A[m][8g .. 8g+7] and
// B[8g .. 8g+7][m]; the result layout is the one of the 32x32x8 form.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ f32x16 mfma16_bf16(bf16x8 a, bf16x8 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// D (32 x 32, row-major FP32) = A (32 x 16 BF16, row-major) x B (16 x 32
// BF16, row-major): one wave; checks the operand layout of the K = 16 form.
static __global__ __launch_bounds__(64) void k_mfma16_probe(const uint16_t* A, const uint16_t* B,
                                                     float* D) {
  const int lane = threadIdx.x, m = lane & 31, g = lane >> 5;
  uint16_t ha[8], hb[8];
  for (int j = 0; j < 8; ++j) {
    ha[j] = A[m * 16 + 8 * g + j];
    hb[j] = B[(8 * g + j) * 32 + m];
  }
  bf16x8 a, b;
  __builtin_memcpy(&a, ha, 16);
  __builtin_memcpy(&b, hb, 16);
  f32x16 acc = {};
  acc = mfma16_bf16(a, b, acc);
  for (int i = 0; i < 16; ++i) D[(8 * (i >> 2) + 4 * g + (i & 3)) * 32 + m] = acc[i];
}

// LDS pitches of k_conv_fwd in BF16 elements: odd multiples of 8 (16 bytes),
// so that the 16-byte reads of 16 consecutive lanes hit distinct banks.
__host__ __device__ constexpr int odd8(int n) {
  return ((n + 7) / 8) % 2 ? (n + 7) / 8 * 8 : (n + 7) / 8 * 8 + 8;
}
__host__ __device__ constexpr int cnn_kp(int ci) { return odd8(kCnnTaps * ci); }
__host__ __device__ constexpr int cnn_ps(int ci) { return odd8(ci); }

// LDS bytes of k_conv_fwd: weights [32][KP] + patch [R + 4][W + 4][PS]
__host__ __device__ inline size_t conv_fwd_lds(int ci, int W, int R) {
  return (size_t)2 * (32 * cnn_kp(ci) + (size_t)(R + 4) * (W + 4) * cnn_ps(ci));
}

// Epilogue of FUSE = 3 (input gradient of a hidden layer through the
// activation): a lane's 16 outputs (channels 8 (i / 4) + 4 g + i % 4 of
// pixel pix) become d = BF16((acc + bias) GELU'(pre)) with the layer's
// pre-activations pre[pix][32], the value k_cnn_act_bwd computes from the
// FP32 gradient this replaces.
// FP32 outputs acc + bias of one lane's pixel: channels 8 j + 4 g .. + 3
// (j < 4) are contiguous, one 16-byte store each where co_out is a multiple
// of 4 and y 16-byte aligned, else one store per channel.
__device__ __forceinline__ void conv_store_pre(const f32x16& acc, const float* bias, float* y,
                                               int co_out, size_t pix, int g) {
  float* yo = y + pix * co_out;
  if ((co_out & 3) == 0 && (reinterpret_cast<uintptr_t>(y) & 15) == 0) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int co = 8 * j + 4 * g;
      if (co < co_out) {
        float4 v;
        v.x = acc[4 * j] + bias[co];
        v.y = acc[4 * j + 1] + bias[co + 1];
        v.z = acc[4 * j + 2] + bias[co + 2];
        v.w = acc[4 * j + 3] + bias[co + 3];
        *reinterpret_cast<float4*>(&yo[co]) = v;
      }
    }
  } else {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int co = 8 * (i >> 2) + 4 * g + (i & 3);
      if (co < co_out) yo[co] = acc[i] + bias[co];
    }
  }
}

__device__ __forceinline__ void conv_bwd_act_store(const f32x16& acc, const float* bias,
                                                   const float* pre, uint16_t* d, size_t pix,
                                                   int g) {
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int co = 8 * j + 4 * g;
    const float4 p = *reinterpret_cast<const float4*>(&pre[pix * 32 + co]);
    const float pv[4] = {p.x, p.y, p.z, p.w};
    uint16_t h[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) h[q] = to_bf16((acc[4 * j + q] + bias[co + q]) * gelu_grad(pv[q]));
    uint2 v;
    v.x = (uint32_t)h[0] | ((uint32_t)h[1] << 16);
    v.y = (uint32_t)h[2] | ((uint32_t)h[3] << 16);
    *reinterpret_cast<uint2*>(&d[pix * 32 + co]) = v;
  }
}

// One workgroup (4 waves) per (image, R output rows); each wave computes
// 32 output channels x 32 pixels tiles.
// Epilogue: FUSE = 0 writes the FP32 pre-activations y (co_out channels);
// FUSE = 1 writes the BF16 GELU activations a[p][32] of a hidden layer
// instead (co_out = 32); FUSE = 2 writes both; FUSE = 3 (input gradient)
// reads the pre-activations from y and writes BF16 (acc + bias) GELU'(y)
// to a (co_out = 32).
// XF32 (CI = 8 only): x is one FP32 channel x[B][H][W], rounded to BF16 and
// zero-padded to 8 channels while it is staged.
// NT threads: 256 (4 waves) for one-row blocks, 512 (8 waves, two tiles
// each per row of 128 columns) for blocks of 2-4 rows
template <int CI, int FUSE = 0, bool XF32 = false, int NT = 256>
static __global__ __launch_bounds__(NT) void k_conv_fwd(const uint16_t* __restrict__ x,
                                                  const uint16_t* __restrict__ w,
                                                  const float* __restrict__ bias,
                                                  float* __restrict__ y, int H, int W,
                                                  int co_out, int R,
                                                  uint16_t* __restrict__ a_out = nullptr) {
  constexpr int K = kCnnTaps * CI;
  constexpr int KP = cnn_kp(CI);
  constexpr int PS = cnn_ps(CI);
  extern __shared__ __align__(16) uint16_t cnn_sm[];
  uint16_t* sw = cnn_sm;
  uint16_t* sp = cnn_sm + 32 * KP;
  const int WP = W + 4;
  const int nyb = H / R;
  const int b = blockIdx.x / nyb, y0 = (blockIdx.x % nyb) * R;
  const int nt = blockDim.x;
  // Staging: 16-byte global loads and LDS stores (K, CI and both pitches
  // are multiples of 8), kLoads weight and kLoads patch loads in flight per
  // thread before the stores. Weights: 32 rows of K. Patch: rows
  // y0 - 2 .. y0 + R + 1, columns -2 .. W + 1 (periodic), pixels of PS;
  // element t = (r WP + c) NV + q. A thread's elements are t0 + j nt: q is
  // fixed (nt is a multiple of NV) and (r, c) advance by nt / NV columns
  // with a row carry, so no division runs per element.
  constexpr int kLoads = 16;
  constexpr int NW16 = 32 * K / 8;
  constexpr int NV = CI / 8;  // 16-byte vectors per pixel
  const int NP16 = (R + 4) * WP * NV;
  const int q = threadIdx.x % NV, cstep = nt / NV;
  int pc = (threadIdx.x / NV) % WP, pr = (threadIdx.x / NV) / WP;
  for (int t0 = threadIdx.x; t0 < NW16 || t0 < NP16; t0 += kLoads * nt) {
    u32x4 vw[kLoads], vp[kLoads];
    int off[kLoads];
#pragma unroll
    for (int j = 0; j < kLoads; ++j) {
      const int t = t0 + j * nt;
      if (t < NW16) vw[j] = reinterpret_cast<const u32x4*>(w)[t];
      off[j] = (pr * WP + pc) * PS + 8 * q;
      if (t < NP16) {
        int gy = y0 - 2 + pr, gx = pc - 2;
        gy += gy < 0 ? H : (gy >= H ? -H : 0);
        gx += gx < 0 ? W : (gx >= W ? -W : 0);
        if constexpr (XF32) {
          static_assert(CI == 8, "one FP32 input channel pads to 8");
          const float v = reinterpret_cast<const float*>(x)[((size_t)b * H + gy) * W + gx];
          vp[j] = u32x4{to_bf16(v), 0u, 0u, 0u};
        } else {
          vp[j] = *reinterpret_cast<const u32x4*>(
              &x[(((size_t)b * H + gy) * W + gx) * CI + 8 * q]);
        }
      }
      pc += cstep;
      while (pc >= WP) {
        pc -= WP;
        ++pr;
      }
    }
#pragma unroll
    for (int j = 0; j < kLoads; ++j) {
      const int t = t0 + j * nt;
      if (t < NW16) {
        const int co = (8 * t) / K, k = (8 * t) % K;
        *reinterpret_cast<u32x4*>(&sw[co * KP + k]) = vw[j];
      }
      if (t < NP16) *reinterpret_cast<u32x4*>(&sp[off[j]]) = vp[j];
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int m = lane & 31, g = lane >> 5;
  const int nchunk = W / 32;
  for (int tile = wid; tile < R * nchunk; tile += nw) {
    const int r = tile / nchunk, n0 = (tile % nchunk) * 32;
    f32x16 acc = {};
    // k = (tap, ci): 8 consecutive k of a lane share one tap (CI % 8 == 0)
#pragma unroll 5
    for (int k0 = 0; k0 + 16 <= K; k0 += 16) {
      const int k = k0 + 8 * g;
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(&sw[m * KP + k]);
      const int tap = k / CI, ci = k % CI;
      const int dy = tap / 5, dx = tap % 5;
      const bf16x8 bb =
          *reinterpret_cast<const bf16x8*>(&sp[((r + dy) * WP + n0 + m + dx) * PS + ci]);
      acc = mfma16_bf16(a, bb, acc);
    }
    if constexpr (K % 16 != 0) {
      // last 8 k: the upper half-wave supplies zeros
      const int k = K - 8;
      bf16x8 a = {}, bb = {};
      if (g == 0) {
        const int tap = k / CI, ci = k % CI;
        a = *reinterpret_cast<const bf16x8*>(&sw[m * KP + k]);
        bb = *reinterpret_cast<const bf16x8*>(
            &sp[((r + tap / 5) * WP + n0 + m + tap % 5) * PS + ci]);
      }
      acc = mfma16_bf16(a, bb, acc);
    }
    const size_t pix = ((size_t)b * H + y0 + r) * W + n0 + m;
    if constexpr (FUSE == 3) {
      conv_bwd_act_store(acc, bias, y, a_out, pix, g);
      continue;
    }
    if (FUSE != 1) {
      conv_store_pre(acc, bias, y, co_out, pix, g);
    }
    if (FUSE != 0) {
      // channels 8 j + 4 g .. + 3 of this lane are contiguous: one 8-byte store
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int co = 8 * j + 4 * g;
        uint16_t h[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) h[q] = to_bf16(gelu(acc[4 * j + q] + bias[co + q]));
        uint2 v;
        v.x = (uint32_t)h[0] | ((uint32_t)h[1] << 16);
        v.y = (uint32_t)h[2] | ((uint32_t)h[3] << 16);
        *reinterpret_cast<uint2*>(&a_out[pix * 32 + co]) = v;
      }
    }
  }
}

// Persistent form of k_conv_fwd for batches (row blocks of kCnnRows = 4
// rows, 8 waves): workgroup g takes a contiguous run of row blocks
// (image-major) of its XCD's contiguous share, stages the 32 x K weights
// once, and keeps the patch rows in a ring of 8 slots (unwrapped image row
// yu >= -2 in slot (yu + 8) mod 8), so the next row block of the same
// image stages its 4 new rows only.  Same arithmetic per output as
// k_conv_fwd.
template <int CI, int FUSE, bool XF32>
static __global__ __launch_bounds__(512) void k_conv_fwd_run(
    const uint16_t* __restrict__ x, const uint16_t* __restrict__ w,
    const float* __restrict__ bias, float* __restrict__ y, int B, int H, int W, int co_out,
    uint16_t* __restrict__ a_out) {
  constexpr int R = kCnnRows;
  constexpr int NSLOT = R + 4;
  static_assert(NSLOT == 8, "ring of 8 patch rows");
  constexpr int K = kCnnTaps * CI;
  constexpr int KP = cnn_kp(CI);
  constexpr int PS = cnn_ps(CI);
  constexpr int NV = CI / 8;  // 16-byte vectors per pixel
  constexpr int kLoads = 16;
  extern __shared__ __align__(16) uint16_t cnn_sm[];
  uint16_t* sw = cnn_sm;
  uint16_t* sp = cnn_sm + 32 * KP;
  const int WP = W + 4;
  const int nt = blockDim.x;
  const int nyb = H / R;
  const int nblk = B * nyb;
  const int G = gridDim.x;
  int first, last;
  if (G % 8 == 0) {
    const int gx = G / 8, xcd = blockIdx.x % 8, i = blockIdx.x / 8;
    const int share = (nblk + 7) / 8, per = (share + gx - 1) / gx;
    first = xcd * share + i * per;
    last = min(min(first + per, (xcd + 1) * share), nblk);
  } else {
    const int per = (nblk + G - 1) / G;
    first = blockIdx.x * per;
    last = min(first + per, nblk);
  }
  // weights once: 32 rows of K, 16-byte vectors
  {
    constexpr int NW16 = 32 * K / 8;
    for (int t0 = threadIdx.x; t0 < NW16; t0 += kLoads * nt) {
      u32x4 vw[kLoads];
#pragma unroll
      for (int j = 0; j < kLoads; ++j) {
        const int t = t0 + j * nt;
        if (t < NW16) vw[j] = reinterpret_cast<const u32x4*>(w)[t];
      }
#pragma unroll
      for (int j = 0; j < kLoads; ++j) {
        const int t = t0 + j * nt;
        if (t < NW16) {
          const int co = (8 * t) / K, k = (8 * t) % K;
          *reinterpret_cast<u32x4*>(&sw[co * KP + k]) = vw[j];
        }
      }
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int m = lane & 31, g = lane >> 5;
  const int nchunk = W / 32;
  int have_b = -1, have_y0 = 0;
  for (int blk = first; blk < last; ++blk) {
    const int b = blk / nyb, y0 = (blk - b * nyb) * R;
    // patch rows pr0 .. NSLOT - 1 (rows y0 - 2 + pr) are staged
    const int pr0 = (b == have_b && y0 == have_y0 + R) ? NSLOT - R : 0;
    have_b = b;
    have_y0 = y0;
    __syncthreads();  // the previous block's MFMA reads of the ring are done
    {
      // element t = ((pr - pr0) WP + c) NV + q as in k_conv_fwd, the row
      // going to its ring slot
      const int NP16 = (NSLOT - pr0) * WP * NV;
      const int q = threadIdx.x % NV, cstep = nt / NV;
      int pc = (threadIdx.x / NV) % WP, pr = pr0 + (threadIdx.x / NV) / WP;
      for (int t0 = threadIdx.x; t0 < NP16; t0 += kLoads * nt) {
        u32x4 vp[kLoads];
        int off[kLoads];
#pragma unroll
        for (int j = 0; j < kLoads; ++j) {
          const int t = t0 + j * nt;
          const int yu = y0 - 2 + pr;
          off[j] = (((yu + NSLOT) & (NSLOT - 1)) * WP + pc) * PS + 8 * q;
          if (t < NP16) {
            int gy = yu, gx = pc - 2;
            gy += gy < 0 ? H : (gy >= H ? -H : 0);
            gx += gx < 0 ? W : (gx >= W ? -W : 0);
            if constexpr (XF32) {
              static_assert(CI == 8, "one FP32 input channel pads to 8");
              const float v = reinterpret_cast<const float*>(x)[((size_t)b * H + gy) * W + gx];
              vp[j] = u32x4{to_bf16(v), 0u, 0u, 0u};
            } else {
              vp[j] = *reinterpret_cast<const u32x4*>(
                  &x[(((size_t)b * H + gy) * W + gx) * CI + 8 * q]);
            }
          }
          pc += cstep;
          while (pc >= WP) {
            pc -= WP;
            ++pr;
          }
        }
#pragma unroll
        for (int j = 0; j < kLoads; ++j) {
          const int t = t0 + j * nt;
          if (t < NP16) *reinterpret_cast<u32x4*>(&sp[off[j]]) = vp[j];
        }
      }
    }
    __syncthreads();
    const int ybase = (y0 - 2 + NSLOT) & (NSLOT - 1);  // slot of patch row 0
    // B operand of tile (r, n0) at k
    auto bop = [&](int r, int n0, int k) {
      const int tap = k / CI, ci = k % CI;
      const int slot = (ybase + r + tap / 5) & (NSLOT - 1);
      return *reinterpret_cast<const bf16x8*>(&sp[(slot * WP + n0 + m + tap % 5) * PS + ci]);
    };
    auto epilogue = [&](const f32x16& acc, int r, int n0) {
      const size_t pix = ((size_t)b * H + y0 + r) * W + n0 + m;
      if constexpr (FUSE == 3) {
        conv_bwd_act_store(acc, bias, y, a_out, pix, g);
        return;
      }
      if (FUSE != 1) {
        conv_store_pre(acc, bias, y, co_out, pix, g);
      }
      if (FUSE != 0) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int co = 8 * j + 4 * g;
          uint16_t h[4];
#pragma unroll
          for (int qq = 0; qq < 4; ++qq) h[qq] = to_bf16(gelu(acc[4 * j + qq] + bias[co + qq]));
          uint2 v;
          v.x = (uint32_t)h[0] | ((uint32_t)h[1] << 16);
          v.y = (uint32_t)h[2] | ((uint32_t)h[3] << 16);
          *reinterpret_cast<uint2*>(&a_out[pix * 32 + co]) = v;
        }
      }
    };
    // A wave takes its tiles two at a time (tile and tile + nw), which
    // share the weight fragment of every k step: 3 LDS reads per 2 MFMAs
    // instead of 4, and two independent accumulation chains.  Each
    // accumulator runs the same MFMA sequence as a lone tile.
    const int ntile = R * nchunk;
    for (int tile = wid; tile < ntile; tile += 2 * nw) {
      const int r = tile / nchunk, n0 = (tile % nchunk) * 32;
      const int tile2 = tile + nw;
      f32x16 acc = {};
      if (tile2 < ntile) {
        const int r2 = tile2 / nchunk, n2 = (tile2 % nchunk) * 32;
        f32x16 acc2 = {};
#pragma unroll 5
        for (int k0 = 0; k0 + 16 <= K; k0 += 16) {
          const int k = k0 + 8 * g;
          const bf16x8 a = *reinterpret_cast<const bf16x8*>(&sw[m * KP + k]);
          acc = mfma16_bf16(a, bop(r, n0, k), acc);
          acc2 = mfma16_bf16(a, bop(r2, n2, k), acc2);
        }
        if constexpr (K % 16 != 0) {
          const int k = K - 8;
          bf16x8 a = {}, bb = {}, bb2 = {};
          if (g == 0) {
            a = *reinterpret_cast<const bf16x8*>(&sw[m * KP + k]);
            bb = bop(r, n0, k);
            bb2 = bop(r2, n2, k);
          }
          acc = mfma16_bf16(a, bb, acc);
          acc2 = mfma16_bf16(a, bb2, acc2);
        }
        epilogue(acc, r, n0);
        epilogue(acc2, r2, n2);
      } else {
#pragma unroll 5
        for (int k0 = 0; k0 + 16 <= K; k0 += 16) {
          const int k = k0 + 8 * g;
          const bf16x8 a = *reinterpret_cast<const bf16x8*>(&sw[m * KP + k]);
          acc = mfma16_bf16(a, bop(r, n0, k), acc);
        }
        if constexpr (K % 16 != 0) {
          const int k = K - 8;
          bf16x8 a = {}, bb = {};
          if (g == 0) {
            a = *reinterpret_cast<const bf16x8*>(&sw[m * KP + k]);
            bb = bop(r, n0, k);
          }
          acc = mfma16_bf16(a, bb, acc);
        }
        epilogue(acc, r, n0);
      }
    }
  }
}

// wt[c][24 - t][o] = w[o][t][c] for o < co, c < ci; wt has 32 rows and
// CO_PAD input channels (zero-padded).
static __global__ void k_conv_flip(const uint16_t* w, int ci, int co, int co_pad, uint16_t* wt) {
  const int n = 32 * kCnnTaps * co_pad;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int o = i % co_pad, t = (i / co_pad) % kCnnTaps, c = i / (co_pad * kCnnTaps);
    uint16_t v = 0;
    if (c < ci && o < co) v = w[((size_t)o * kCnnTaps + (24 - t)) * ci + c];
    wt[i] = v;
  }
}

// Weight gradient dw[o][t][c] = sum_p dy[p][o] x[p + d(t)][c] and bias
// gradient db[o] = sum_p dy[p][o], as a GEMM over the pixels p (the MFMA K
// dimension) on v_mfma_f32_32x32x16_bf16: M = 32 output channels o,
// N = 32 input channels c of one tap.
//
// Row blocks of ROWS output rows.  Staging transposes to pixels-fast in
// LDS: dy as [ROWS][32 o][DP], the x patch (rows y0 - 2 .. y0 + ROWS + 1,
// patch column j = image column j - 2, periodic) as [ROWS + 4][32 c][XP].  Channel row c starts wg_pad(c) elements into
// its slot and each patch row is 32 elements longer than its 32 channel
// rows, so the transposed 8-byte writes of the lanes of one pixel run (4
// channel vectors) spread over the banks; channel rows beyond CI / CO_IN
// are zero (written once), so the MMA loop reads every lane unmasked.
// A thread loads 8 pixels x 8 channels with 16-byte loads (two halves of 4
// pixels; lanes of one pixel run are consecutive, 16 cache lines per wave
// instruction), transposes each 4 x 8 block of BF16 in registers and writes
// 8 rows of 8 bytes.  Both pitches are odd multiples of 8 elements, so the
// 16-byte reads of 8 consecutive lanes (channels) fall in distinct banks.
//
// Wave w (5 waves) owns the taps of row dy = w.  Per 16 pixels it reads the
// dy operand once and two aligned 16-byte x runs per lane, one chunk ahead
// of the MFMAs; the operands of taps dx = 0..4 (x columns shifted by dx) are
// the two runs shifted by dx elements (v_alignbit), so all LDS reads stay
// 16-byte aligned.  Wave 0 also multiplies dy by a column of ones: the bias
// gradient.
//
// Persistent: workgroup g (of G) takes a contiguous run of row blocks of
// its XCD's contiguous share and keeps the x patch rows in a ring of
// ROWS + 4 slots, so consecutive row blocks of an image stage ROWS new x
// rows instead of ROWS + 4; it writes its partials part[g][32][25 CI]
// and, after all G weight partials, part[G 32 25 CI + 32 g + o];
// k_cnn_reduce sums them in a fixed order (deterministic).
// Rows per block ROWS: kWgRowsMax where the block's LDS fits (fewer x rows
// staged per output row and half the blocks: batch 32 of 128^2, 63 -> 55 us
// with the partials' reduction), else kWgRows, which also bounds the
// workgroup count of the workspace.
constexpr int kWgRowsMax = 4;
constexpr int kWgRows = 2;
constexpr int kWgThreads = 320; // 5 waves, one per tap row
constexpr int kWgGroups = 256;  // persistent workgroups (one per CU; halves the partials)

__host__ __device__ inline int wg_dp(int W) { return odd8(W); }
__host__ __device__ inline int wg_xp(int W) { return odd8(W + 8); }
__host__ __device__ inline int wg_pad(int c) { return 8 * ((c >> 3) & 3); }
// LDS elements of one dy row / one x patch row (32 channel rows + pads)
__host__ __device__ inline int wg_drow(int W) { return 32 * wg_dp(W) + 32; }
__host__ __device__ inline int wg_xrow(int W) { return 32 * wg_xp(W) + 32; }
// LDS bytes of k_conv_wgrad with row blocks of `rows` rows
__host__ __device__ inline size_t conv_wgrad_lds(int ci, int W, int rows = kWgRows) {
  (void)ci;
  return (size_t)2 * ((size_t)rows * wg_drow(W) + (size_t)(rows + 4) * wg_xrow(W));
}
// Workgroups of k_conv_wgrad for B images of H rows (at most: row blocks of
// kWgRows rows; blocks of more rows give at most as many)
inline int wgrad_groups(long long B, int H, int rows = kWgRows) {
  const long long nblk = B * (H / rows);
  if (nblk >= kWgGroups) return kWgGroups;
  return (int)nblk;
}

// Transposes 4 pixels x 8 channels of BF16 (v[j]: channels of pixel j, two
// per dword, low half first) into t[c]: the 4 pixels of channel c.
__device__ __forceinline__ void bf16_transpose4x8(const u32x4 (&v)[4], uint2 (&t)[8]) {
#pragma unroll
  for (int d = 0; d < 4; ++d) {
    t[2 * d].x = (v[0][d] & 0xffffu) | (v[1][d] << 16);
    t[2 * d].y = (v[2][d] & 0xffffu) | (v[3][d] << 16);
    t[2 * d + 1].x = (v[0][d] >> 16) | (v[1][d] & 0xffff0000u);
    t[2 * d + 1].y = (v[2][d] >> 16) | (v[3][d] & 0xffff0000u);
  }
}

__device__ __forceinline__ bf16x8 as_bf16x8(u32x4 v) { return __builtin_bit_cast(bf16x8, v); }

// MFMAs of one 16-pixel chunk: dy operand a, x runs x0 (columns px .. px + 7
// of the patch) and x1 (px + 8 .. px + 15); tap dx reads px + dx .. + 7.
template <bool BIAS>
__device__ __forceinline__ void wg_chunk(const u32x4& a, const u32x4& x0, const u32x4& x1,
                                         const u32x4& ones, f32x16 (&acc)[5], f32x16& accb) {
  const bf16x8 A = as_bf16x8(a);
  const u32x4 b1 = {__builtin_amdgcn_alignbit(x0[1], x0[0], 16),
                    __builtin_amdgcn_alignbit(x0[2], x0[1], 16),
                    __builtin_amdgcn_alignbit(x0[3], x0[2], 16),
                    __builtin_amdgcn_alignbit(x1[0], x0[3], 16)};
  const u32x4 b2 = {x0[1], x0[2], x0[3], x1[0]};
  const u32x4 b3 = {__builtin_amdgcn_alignbit(x0[2], x0[1], 16),
                    __builtin_amdgcn_alignbit(x0[3], x0[2], 16),
                    __builtin_amdgcn_alignbit(x1[0], x0[3], 16),
                    __builtin_amdgcn_alignbit(x1[1], x1[0], 16)};
  const u32x4 b4 = {x0[2], x0[3], x1[0], x1[1]};
  acc[0] = mfma16_bf16(A, as_bf16x8(x0), acc[0]);
  acc[1] = mfma16_bf16(A, as_bf16x8(b1), acc[1]);
  acc[2] = mfma16_bf16(A, as_bf16x8(b2), acc[2]);
  acc[3] = mfma16_bf16(A, as_bf16x8(b3), acc[3]);
  acc[4] = mfma16_bf16(A, as_bf16x8(b4), acc[4]);
  if constexpr (BIAS) accb = mfma16_bf16(A, as_bf16x8(ones), accb);
}

// MFMA loop over the staged row block: full 16-pixel chunks with the loads
// of the next chunk issued before the MFMAs of the current one, then, for
// W % 16 == 8, the last 8 pixels (upper half-wave supplies zeros).
template <bool BIAS, int ROWS>
__device__ __forceinline__ void wg_rows(const uint16_t* pa0, const uint16_t* const (&pxrow)[ROWS],
                                        int W, int DROW, int g, f32x16 (&acc)[5],
                                        f32x16& accb) {
  const u32x4 ones = {0x3f803f80u, 0x3f803f80u, 0x3f803f80u, 0x3f803f80u};
  const int wfull = W & ~15;
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    const uint16_t* pa = pa0 + r * DROW;
    const uint16_t* pxr = pxrow[r];
    u32x4 a = *reinterpret_cast<const u32x4*>(pa);
    u32x4 x0 = *reinterpret_cast<const u32x4*>(pxr);
    u32x4 x1 = *reinterpret_cast<const u32x4*>(pxr + 8);
    for (int px0 = 16; px0 < wfull; px0 += 16) {
      const u32x4 an = *reinterpret_cast<const u32x4*>(pa + px0);
      const u32x4 x0n = *reinterpret_cast<const u32x4*>(pxr + px0);
      const u32x4 x1n = *reinterpret_cast<const u32x4*>(pxr + px0 + 8);
      wg_chunk<BIAS>(a, x0, x1, ones, acc, accb);
      a = an;
      x0 = x0n;
      x1 = x1n;
    }
    wg_chunk<BIAS>(a, x0, x1, ones, acc, accb);
    if (W & 8) {
      const u32x4 z = {0u, 0u, 0u, 0u};
      u32x4 at = z, x0t = z, x1t = z;
      if (g == 0) {
        at = *reinterpret_cast<const u32x4*>(pa + wfull);
        x0t = *reinterpret_cast<const u32x4*>(pxr + wfull);
        x1t = *reinterpret_cast<const u32x4*>(pxr + wfull + 8);
      }
      wg_chunk<BIAS>(at, x0t, x1t, g == 0 ? ones : z, acc, accb);
    }
  }
}

template <int CI, int CO_IN, int ROWS>
static __global__ __launch_bounds__(kWgThreads) __attribute__((amdgpu_waves_per_eu(3)))
void k_conv_wgrad(const uint16_t* __restrict__ x, const uint16_t* __restrict__ dy,
                  float* __restrict__ part, int B, int H, int W) {
  constexpr int N = kCnnTaps * CI;
  constexpr int NVI = CI / 8, NVO = CO_IN / 8;  // 16-byte vectors per pixel
  extern __shared__ __align__(16) uint16_t cnn_sm[];
  const int DP = wg_dp(W), XP = wg_xp(W);
  const int DROW = wg_drow(W), XROW = wg_xrow(W);
  uint16_t* sd = cnn_sm;                    // [ROWS][DROW]
  uint16_t* sx = cnn_sm + ROWS * DROW;   // [ROWS + 4][XROW]
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int m = lane & 31, g = lane >> 5;
  // channel rows that staging never writes: zero once
  if (CI < 32 || CO_IN < 32) {
    const u32x4 z = {0u, 0u, 0u, 0u};
    for (int r = 0; r < ROWS + 4; ++r) {
      if (r < ROWS)
        for (int i = threadIdx.x; i < (32 - CO_IN) * (DP / 8); i += kWgThreads) {
          const int o = CO_IN + i / (DP / 8);
          *reinterpret_cast<u32x4*>(sd + r * DROW + o * DP + wg_pad(o) + 8 * (i % (DP / 8))) = z;
        }
      for (int i = threadIdx.x; i < (32 - CI) * (XP / 8); i += kWgThreads) {
        const int c = CI + i / (XP / 8);
        *reinterpret_cast<u32x4*>(sx + r * XROW + c * XP + wg_pad(c) + 8 * (i % (XP / 8))) = z;
      }
    }
  }
  f32x16 acc[5], accb = {};
#pragma unroll
  for (int q = 0; q < 5; ++q) acc[q] = f32x16{};
  const int nyb = H / ROWS;
  const int nblk = B * nyb;
  const int G = gridDim.x;
  // a contiguous run of row blocks per workgroup, the runs of one XCD's
  // workgroups contiguous too (blocks go round-robin to the 8 XCDs; the
  // halo rows between runs are then shared in that XCD's L2)
  int first, last;
  if (G % 8 == 0) {
    const int gx = G / 8, xcd = blockIdx.x % 8, i = blockIdx.x / 8;
    const int share = (nblk + 7) / 8, per = (share + gx - 1) / gx;
    first = xcd * share + i * per;
    last = min(min(first + per, (xcd + 1) * share), nblk);
  } else {
    const int per = (nblk + G - 1) / G;
    first = blockIdx.x * per;
    last = min(first + per, nblk);
  }
  const int ncg = W / 8 + 1;  // 8-column groups of the x patch
  // this lane's dy operand row (pixels-fast): channel m
  const uint16_t* pa0 = sd + m * DP + wg_pad(m) + 8 * g;
  const int xoff = m * XP + wg_pad(m) + 8 * g;
  // x rows live in a ring of ROWS + 4 slots: image row y (unwrapped,
  // y >= -2) in slot (y + 6) mod 6; the next row block of the same image
  // keeps 4 rows and stages 2
  constexpr int NSLOT = ROWS + 4;
  int have_b = -1, have_y0 = 0;
  for (int blk = first; blk < last; ++blk) {
    const int b = blk / nyb, y0 = (blk - b * nyb) * ROWS;
    const int pr0 = (b == have_b && y0 == have_y0 + ROWS) ? NSLOT - ROWS : 0;
    have_b = b;
    have_y0 = y0;
    __syncthreads();
    // x patch rows pr0 .. NSLOT - 1: item = (patch row, column group,
    // channel vector), channel vector fastest
    for (int it = threadIdx.x; it < (NSLOT - pr0) * NVI * ncg; it += kWgThreads) {
      const int q = it % NVI, rk = it / NVI;
      const int k = rk % ncg, pr = pr0 + rk / ncg;
      const int yu = y0 - 2 + pr;  // unwrapped image row
      const int gy = yu < 0 ? yu + H : (yu >= H ? yu - H : yu);
      const uint16_t* row = x + ((size_t)b * H + gy) * W * CI + 8 * q;
      uint16_t* dst = sx + ((yu + NSLOT) % NSLOT) * XROW + 8 * q * XP + wg_pad(8 * q) + 8 * k;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        u32x4 v[4];
        uint2 t[8];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          int gx = 8 * k - 2 + 4 * h + j;
          gx += gx < 0 ? W : (gx >= W ? -W : 0);
          v[j] = *reinterpret_cast<const u32x4*>(row + (size_t)gx * CI);
        }
        bf16_transpose4x8(v, t);
#pragma unroll
        for (int c = 0; c < 8; ++c) *reinterpret_cast<uint2*>(dst + c * XP + 4 * h) = t[c];
      }
    }
    // dy rows: item = (row, column group, channel vector)
    for (int it = threadIdx.x; it < ROWS * NVO * (W / 8); it += kWgThreads) {
      const int q = it % NVO, rk = it / NVO;
      const int k = rk % (W / 8), r = rk / (W / 8);
      const uint16_t* src = dy + (((size_t)b * H + y0 + r) * W + 8 * k) * CO_IN + 8 * q;
      uint16_t* dst = sd + r * DROW + 8 * q * DP + wg_pad(8 * q) + 8 * k;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        u32x4 v[4];
        uint2 t[8];
#pragma unroll
        for (int j = 0; j < 4; ++j)
          v[j] = *reinterpret_cast<const u32x4*>(src + (4 * h + j) * CO_IN);
        bf16_transpose4x8(v, t);
#pragma unroll
        for (int c = 0; c < 8; ++c) *reinterpret_cast<uint2*>(dst + c * DP + 4 * h) = t[c];
      }
    }
    __syncthreads();
    // output row r, tap row wid: image row y0 + r + wid - 2
    const uint16_t* pxrow[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r)
      pxrow[r] = sx + ((y0 - 2 + r + wid + NSLOT) % NSLOT) * XROW + xoff;
    if (wid == 0)
      wg_rows<true, ROWS>(pa0, pxrow, W, DROW, g, acc, accb);
    else
      wg_rows<false, ROWS>(pa0, pxrow, W, DROW, g, acc, accb);
  }
  // partials: n = (5 w + dx) CI + c, lanes m < CI
  float* dst = part + (size_t)blockIdx.x * 32 * N;
  if (m < CI) {
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      const int n = (5 * wid + q) * CI + m;
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int o = 8 * (i >> 2) + 4 * g + (i & 3);
        dst[(size_t)o * N + n] = acc[q][i];
      }
    }
  }
  if (wid == 0 && m == 0) {
    float* bp = part + (size_t)G * 32 * N + (size_t)blockIdx.x * 32;
#pragma unroll
    for (int i = 0; i < 16; ++i) bp[8 * (i >> 2) + 4 * g + (i & 3)] = accb[i];
  }
}

// out[i] = sum_g part[g][i], i < n. A workgroup (4 waves) owns 64
// consecutive outputs; wave w sums the groups w, w + 4, ... into 8
// interleaved partial sums (8 loads in flight per lane), which are combined
// in a fixed order, then the 4 waves in a fixed order: deterministic, the
// same result for every launch. Two such sums run in one launch: workgroups
// below nb1 = (n + 63) / 64 sum part / n into out, the (n2 + 63) / 64
// after them part2 / n2 into out2 (the weight and the bias partials).
static __global__ __launch_bounds__(256) void k_cnn_reduce(const float* part, int ngroups, int n,
                                                    float* out, int nb1, const float* part2,
                                                    int n2, float* out2) {
  __shared__ float red[4][64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int blk = blockIdx.x;
  if (blk >= nb1) {
    blk -= nb1;
    part = part2;
    n = n2;
    out = out2;
  }
  const int i = blk * 64 + lane;
  float acc[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) acc[j] = 0.0f;
  if (i < n) {
    int gq = wid;
    for (; gq + 28 < ngroups; gq += 32) {
#pragma unroll
      for (int j = 0; j < 8; ++j) acc[j] += part[(size_t)(gq + 4 * j) * n + i];
    }
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (gq + 4 * j < ngroups) acc[j] += part[(size_t)(gq + 4 * j) * n + i];
  }
  red[wid][lane] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
  __syncthreads();
  if (wid == 0 && i < n) out[i] = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
}

// a[p][c] = BF16(act(y[p][c])) for c < cin, 0 up to cpad; act = GELU
// (gelu != 0) or identity
static __global__ void k_cnn_act(const float* y, long long npix, int cin, int cpad, int use_gelu,
                          uint16_t* a) {
  const long long n = npix * cpad;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long p = i / cpad;
    const int c = (int)(i % cpad);
    float v = 0.0f;
    if (c < cin) {
      v = y[p * cin + c];
      if (use_gelu) v = gelu(v);
    }
    a[i] = to_bf16(v);
  }
}

// d[p][c] = BF16(g[p][c] * gelu'(y[p][c])) (g: gradient w.r.t. the GELU
// output, channel stride gs; y: pre-activation, c channels)
static __global__ void k_cnn_act_bwd(const float* g, int gs, const float* y, long long npix, int c,
                              uint16_t* d) {
  const long long n = npix * c;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long p = i / c;
    const int k = (int)(i % c);
    d[i] = to_bf16(g[p * gs + k] * gelu_grad(y[i]));
  }
}

// loss partials and gradient of L = mean((out - target)^2):
// d[p][c] = BF16(2 (out - target) / n) for c = 0, 0 up to cpad
static __global__ __launch_bounds__(256) void k_cnn_mse(const float* out, const float* target,
                                                 long long npix, int cpad, uint16_t* d,
                                                 double* part) {
  __shared__ double red[256];
  double s = 0.0;
  const double inv = 1.0 / (double)npix;
  for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < npix;
       p += (long long)gridDim.x * blockDim.x) {
    const float e = out[p] - target[p];
    s += (double)e * e;
    d[p * cpad] = to_bf16((float)(2.0 * inv) * e);
    for (int c = 1; c < cpad; ++c) d[p * cpad + c] = 0;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] * inv;
}

// Adam (FP32 master weights) and the BF16 copy used by the kernels
static __global__ void k_cnn_adam(float* w, const float* gr, float* m1, float* m2, int n, float lr,
                           float b1, float b2, float eps, float c1, float c2, uint16_t* wb) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float gi = gr[i];
    const float a = b1 * m1[i] + (1.0f - b1) * gi;
    const float v = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = v;
    const float wi = w[i] - lr * (a / c1) / (sqrtf(v / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = to_bf16(wi);
  }
}

// Adam over up to kAdamTensors tensors in one launch, the tensors' pointer
// table passed by value: tensor t owns the workgroups [blk[t], blk[t + 1])
// and runs the per-element update of k_cnn_adam over its n[t] elements.
constexpr int kAdamTensors = 16;
struct AdamTable {
  float* w[kAdamTensors];
  const float* g[kAdamTensors];
  float* m1[kAdamTensors];
  float* m2[kAdamTensors];
  uint16_t* wb[kAdamTensors];
  int n[kAdamTensors];
  int blk[kAdamTensors + 1];
  int count;
};

static __global__ __launch_bounds__(256) void k_cnn_adam_multi(AdamTable tb, float lr, float b1,
                                                               float b2, float eps, float c1,
                                                               float c2) {
  int t = 0;
  while (t + 1 < tb.count && (int)blockIdx.x >= tb.blk[t + 1]) ++t;
  const int nb = tb.blk[t + 1] - tb.blk[t];
  float* w = tb.w[t];
  const float* gr = tb.g[t];
  float* m1 = tb.m1[t];
  float* m2 = tb.m2[t];
  uint16_t* wb = tb.wb[t];
  const int n = tb.n[t];
  for (int i = (blockIdx.x - tb.blk[t]) * blockDim.x + threadIdx.x; i < n; i += nb * blockDim.x) {
    const float gi = gr[i];
    const float a = b1 * m1[i] + (1.0f - b1) * gi;
    const float v = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = v;
    const float wi = w[i] - lr * (a / c1) / (sqrtf(v / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = to_bf16(wi);
  }
}

// BF16 copy of FP32 values (weights)
static __global__ void k_cnn_to_bf16(const float* in, int n, uint16_t* out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = to_bf16(in[i]);
}

}  // namespace fsx
