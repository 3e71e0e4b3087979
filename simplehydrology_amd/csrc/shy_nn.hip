// shy_nn.hip -- BF16 MFMA kernels of the discharge/momentum surrogate (gfx950).
//
// Activations are NHWC bf16 (channels contiguous).  A 3x3 (pad 1) or 1x1
// convolution is an implicit GEMM: M = pixels, N = output channels,
// K = taps x input channels (tap-major), weights stored Wt[Cout][K] so both
// MFMA operands load 8 consecutive k as one 16-byte vector.  Every product
// runs on v_mfma_f32_16x16x32_bf16 with FP32 accumulation.  Lane layout of
// that instruction (checked on the device): lane l holds A[l%16][(l/16)*8+i],
// B[(l/16)*8+i][l%16] for i < 8 and C[(l/16)*4+i][l%16] for i < 4.
//
//   conv_tiles_kernel  3x3 forward / data gradient of every layer whose
//                      16x16 pixel tiles give each CU at least one:
//                      persistent blocks (three per CU, 32 output channels
//                      each) over the tiles, patch and weights staged in LDS,
//                      the next chunk's patch loaded into registers during
//                      the current chunk's MFMAs; decoder inputs read as
//                      [up2(x), skip] from two tensors; optionally a 1x1
//                      head on the rounded output, or its 2x2 average pool
//                      (pairs of rows in a lane, pairs of columns across
//                      lanes l and l ^ 1)
//   conv_pf_kernel     the same arithmetic with one tile per block for the
//                      smaller layers (16x8 tiles: more blocks)
//   conv_c8_kernel     the 8-channel first layer, four taps per MFMA k-step
//   conv_fwd_kernel    direct-load convolution of the remaining shapes (the
//                      head's data gradient)
//   epilogue           bias, an optional residual and ReLU, bf16 or f32
//                      store; for the data gradient (the forward kernels on
//                      weights flipped by flip_weights_kernel or written so
//                      by Adam) the residual gradient and the ReLU mask
//   wgrad_lds_kernel   dW[co][k] = sum_p g[p][co] x[p + tap][ci]: K = pixels,
//   conv_wgrad_kernel  staged through LDS (both operands need 8 consecutive
//                      pixels per lane; the second loads directly), split
//                      over pixel ranges into FP32 partials that
//                      wgrad_reduce_kernel sums in a fixed order
//   elementwise        ReLU backward, 2x2 average pool, nearest upsample,
//                      channel concat/split, bias gradient, MSE loss, Adam
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/shy_nn.h"

// Split-K partials of shy_nn_conv_wgrad_noalloc: SHY_NN_NOALLOC_FLOATS
// floats (16 MiB) on every device the library runs on, held until the
// process ends.  On the host this name is a stand-in, not a device address:
// host code uses only the address hipGetSymbolAddress returns for the
// current device (noalloc_scratch below).
static __device__ __align__(256) float scratch[SHY_NN_NOALLOC_FLOATS];

namespace shy_nn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline bf16x8 zero8() {
  bf16x8 z;
#pragma unroll
  for (int i = 0; i < 8; ++i) z[i] = (__bf16)0.0f;
  return z;
}

__device__ inline bf16x8 load8(const __bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }

__device__ inline f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

struct ConvArgs {
  const __bf16* __restrict__ x;    // [P][Cin], P = N*H*W
  const __bf16* __restrict__ w;    // [Cout][K]
  const float* __restrict__ bias;  // [Cout] or null
  const __bf16* __restrict__ res;  // [P][Cout] or null (added before the ReLU)
  __bf16* y_bf;                    // [P][Cout] (one of y_bf / y_f)
  float* y_f;
  int N, H, W, Cin, Cout, K;
  int relu;
  // When x2 is set (conv_pf_kernel only), the input is the concatenation of
  // a nearest x2 upsample of x ([N][H/2][W/2][C1]) and x2 ([N][H][W][Cin-C1]).
  const __bf16* __restrict__ x2;
  int C1;
  // Backward-pass epilogue of the LDS-staged kernels (bf16 output): after
  // rounding, gres ([P][Cout]) is added (in float, rounded again) and the
  // result is zeroed where mask ([P][Cout]) is not positive: the bits of
  // shy_nn_add then shy_nn_relu_bwd on the rounded output.
  const __bf16* __restrict__ gres;
  const __bf16* __restrict__ mask;
  // When head_w is set (conv_tiles_kernel, 32 output channels, BF16 output):
  // the epilogue also applies a 1x1 convolution to the rounded output,
  // head_y [P][head_c] (FP32) = head_b + head_w [head_c][32] . y, with one
  // MFMA per row of pixels; y_bf may then be null (y not stored).
  const __bf16* __restrict__ head_w = nullptr;
  const float* __restrict__ head_b = nullptr;
  float* head_y = nullptr;
  int head_c = 0;
  // When pool_y is set (conv_tiles_kernel, H and W even): also the 2x2
  // average of the rounded output, [N][H/2][W/2][Cout], with the bits of
  // pool2_fwd_v8_kernel on y.
  __bf16* pool_y = nullptr;
};

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

// One lane's 4 consecutive output channels co..co+3 of output element o:
// bias, residual (before rounding), ReLU, then the store, FP32 or BF16 (with
// the backward-pass residual and mask after rounding).  Returns the stored
// BF16 values (zeros for an FP32 output); with HEAD a null y_bf stores
// nothing.
template <bool HEAD>
__device__ inline bf16x4 store_out4(const ConvArgs& a, size_t o, int co, const f32x4& acc) {
  float v[4];
  // the lane's 4 biases in one load: co is a multiple of 4 and the entry
  // points require a 16-byte aligned bias
  const float4 b = a.bias ? *reinterpret_cast<const float4*>(a.bias + co)
                          : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  v[0] = acc[0] + b.x;
  v[1] = acc[1] + b.y;
  v[2] = acc[2] + b.z;
  v[3] = acc[3] + b.w;
  if (a.res) {
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] += (float)a.res[o + i];
  }
  if (a.relu) {
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = v[i] > 0.0f ? v[i] : 0.0f;
  }
  bf16x4 q;
  if (a.y_f) {
    *reinterpret_cast<float4*>(a.y_f + o) = make_float4(v[0], v[1], v[2], v[3]);
#pragma unroll
    for (int i = 0; i < 4; ++i) q[i] = (__bf16)0.0f;
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i) q[i] = (__bf16)v[i];
    if (a.gres) {
#pragma unroll
      for (int i = 0; i < 4; ++i) q[i] = (__bf16)((float)q[i] + (float)a.gres[o + i]);
    }
    if (a.mask) {
#pragma unroll
      for (int i = 0; i < 4; ++i) q[i] = (float)a.mask[o + i] > 0.0f ? q[i] : (__bf16)0.0f;
    }
    if (!HEAD || a.y_bf) *reinterpret_cast<bf16x4*>(a.y_bf + o) = q;
  }
  return q;
}

// The 1x1 head on a tile's rounded 32-channel output q (ConvArgs::head_w),
// one MFMA per row of 16 pixels: the output channels are the M operand and
// the pixels the N operand, and the K operand takes each lane's own 8
// channels in the order it holds them, 4g..4g+3 then 16+4g..16+4g+3 for
// lane group g = lane/16, so the rounded output needs no exchange; the
// head's weights are read in that order.  Lane holds head channels
// 4(lane/16) + i of pixel lane%16.
template <int NT>
__device__ inline void conv_head(const ConvArgs& a, const bf16x4 (&q)[2][NT], int n, int y0,
                                 int x0, int wave, int lane) {
  const int hc = a.head_c, o = lane & 15, g = lane >> 4;
  // two 8-byte loads (head_w is 16-byte aligned, rows of 32)
  bf16x4 w0, w1;
#pragma unroll
  for (int j = 0; j < 4; ++j) w0[j] = w1[j] = (__bf16)0.0f;
  if (o < hc) {
    w0 = *reinterpret_cast<const bf16x4*>(a.head_w + o * 32 + 4 * g);
    w1 = *reinterpret_cast<const bf16x4*>(a.head_w + o * 32 + 16 + 4 * g);
  }
  bf16x8 A;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    A[j] = w0[j];
    A[4 + j] = w1[j];
  }
  const int gx = x0 + (lane & 15), oc = 4 * g;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    bf16x8 B;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      B[j] = q[0][t][j];
      B[4 + j] = q[1][t][j];
    }
    const f32x4 h = mfma(A, B, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
    const int gy = y0 + wave * NT + t;
    if (gy < a.H && gx < a.W && oc < hc) {
      const size_t p = (size_t)(n * a.H + gy) * a.W + gx;
      const float4 b = a.head_b ? *reinterpret_cast<const float4*>(a.head_b + oc)
                                : make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // 16-byte aligned
      *reinterpret_cast<float4*>(a.head_y + p * hc + oc) =
          make_float4(h[0] + b.x, h[1] + b.y, h[2] + b.z, h[3] + b.w);
    }
  }
}

// The epilogue of a block's tile: lane holds channels cb + 16m + 4(lane/16)
// + i of pixel (y0 + NT wave + t, x0 + lane%16).  HEAD: then the 1x1 head
// (its own instantiation: the rounded outputs it keeps cost registers).
template <int MT, int NT, bool HEAD = false, bool POOL = false>
__device__ inline void conv_epilogue(const ConvArgs& a, const f32x4 (&acc)[MT][NT], int n, int y0,
                                     int x0, int cb, int wave, int lane) {
  static_assert(!POOL || NT % 2 == 0, "the pool pairs a wave's rows");
  static_assert(!HEAD || MT == 2, "the head reads a block's 32 output channels");
  const int gx = x0 + (lane & 15);
  bf16x4 q[MT][NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int gy = y0 + wave * NT + t;
    const size_t p = (size_t)(n * a.H + gy) * a.W + gx;
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      const int co = cb + m * 16 + (lane >> 4) * 4;
#pragma unroll
      for (int i = 0; i < 4; ++i) q[m][t][i] = (__bf16)0.0f;
      if (gy < a.H && gx < a.W && co < a.Cout)
        q[m][t] = store_out4<HEAD>(a, p * a.Cout + co, co, acc[m][t]);
    }
  }
  if constexpr (HEAD) conv_head<NT>(a, q, n, y0, x0, wave, lane);
  if constexpr (POOL) {
    // Rows y0 + NT wave + t pair as (t, t + 1) for even t (y0 and NT wave
    // are even); columns pair as lanes (l, l ^ 1).  The even lane of a pair
    // sums (top left, top right, bottom left, bottom right) in the order
    // of pool2_fwd_v8_kernel.  H and W are even, so a pooled pixel lies in
    // bounds exactly when its top-left input pixel does.
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      const int co = cb + m * 16 + (lane >> 4) * 4;
#pragma unroll
      for (int t = 0; t < NT; t += 2) {
        bf16x4 r0, r1;
        {
          int2 u0 = *reinterpret_cast<const int2*>(&q[m][t]);
          int2 u1 = *reinterpret_cast<const int2*>(&q[m][t + 1]);
          u0.x = __shfl_xor(u0.x, 1);
          u0.y = __shfl_xor(u0.y, 1);
          u1.x = __shfl_xor(u1.x, 1);
          u1.y = __shfl_xor(u1.y, 1);
          r0 = *reinterpret_cast<const bf16x4*>(&u0);
          r1 = *reinterpret_cast<const bf16x4*>(&u1);
        }
        const int gy = y0 + wave * NT + t;
        if ((lane & 1) == 0 && gy < a.H && gx < a.W && co < a.Cout) {
          bf16x4 o;
#pragma unroll
          for (int i = 0; i < 4; ++i)
            o[i] = (__bf16)(0.25f * ((float)q[m][t][i] + (float)r0[i] + (float)q[m][t + 1][i] +
                                     (float)r1[i]));
          const size_t pp = ((size_t)(n * (a.H >> 1) + (gy >> 1)) * (a.W >> 1) + (gx >> 1)) * a.Cout;
          *reinterpret_cast<bf16x4*>(a.pool_y + pp + co) = o;
        }
      }
    }
  }
}

// Forward (and dgrad) convolution.  Block = 4 waves; each wave owns 32 pixels
// x 16*NT output channels as 2 x NT MFMA accumulators.
template <int TAPS, int NT>
__global__ void __launch_bounds__(256) conv_fwd_kernel(ConvArgs a) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int P = a.N * a.H * a.W;
  const int pb = (blockIdx.x * 4 + wave) * 32;
  const int cb = blockIdx.y * 16 * NT;
  const int K = a.K, Cin = a.Cin;
  int pn[2], py[2], px[2];
  bool pv[2];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int p = pb + r * 16 + (lane & 15);
    pv[r] = p < P;
    const int pp = pv[r] ? p : 0;
    px[r] = pp % a.W;
    py[r] = (pp / a.W) % a.H;
    pn[r] = pp / (a.W * a.H);
  }
  f32x4 acc[2][NT];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < NT; ++c) acc[r][c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  const int ksteps = (K + 31) / 32;
  for (int ks = 0; ks < ksteps; ++ks) {
    const int k0 = ks * 32 + (lane >> 4) * 8;
    const bool kv = k0 < K;
    const int tap = kv ? k0 / Cin : 0;
    const int ci0 = kv ? k0 - tap * Cin : 0;
    const int dy = TAPS == 9 ? tap / 3 - 1 : 0;
    const int dx = TAPS == 9 ? tap % 3 - 1 : 0;
    bf16x8 A[2], B[NT];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int yy = py[r] + dy, xx = px[r] + dx;
      const bool ok = kv && pv[r] && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W;
      A[r] = ok ? load8(a.x + ((size_t)(pn[r] * a.H + yy) * a.W + xx) * Cin + ci0) : zero8();
    }
#pragma unroll
    for (int c = 0; c < NT; ++c) {
      const int co = cb + c * 16 + (lane & 15);
      B[c] = (kv && co < a.Cout) ? load8(a.w + (size_t)co * K + k0) : zero8();
    }
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int c = 0; c < NT; ++c) acc[r][c] = mfma(A[r], B[c], acc[r][c]);
  }
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < NT; ++c) {
      const int co = cb + c * 16 + (lane & 15);
      if (co >= a.Cout) continue;
      const float bv = a.bias ? a.bias[co] : 0.0f;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int p = pb + r * 16 + (lane >> 4) * 4 + i;
        if (p >= P) continue;
        float v = acc[r][c][i] + bv;
        const size_t o = (size_t)p * a.Cout + co;
        if (a.res) v += (float)a.res[o];
        if (a.relu) v = v > 0.0f ? v : 0.0f;
        if (a.y_f) a.y_f[o] = v;
        else a.y_bf[o] = (__bf16)v;
      }
    }
}

// Tile width of the LDS-staged convolutions, and LDS bf16 per patch pixel
// (32 channels + 8: spreads the banks).
constexpr int kTW = 16;
constexpr int kPS = 40;

// Forward (and dgrad) convolution for Cin and Cout multiples of 32, staged
// through LDS.  A block owns a 16 (x) by 4*NT (y) pixel tile of one image and
// 16*MT output channels.  Per 32-channel chunk of the input it stages the
// tile's input patch with a 1-pixel halo (TAPS = 9) and the chunk's weights
// for every tap; output channels are the MFMA's M operand (A = weights),
// pixels its N operand (B = patch), so a lane ends with 4 consecutive
// channels of one pixel and stores them as one vector.  Wave w owns tile
// rows NT*w .. NT*w+NT-1 (one 16-pixel N tile each).  The next chunk's patch
// and weights are
// loaded into registers while the current chunk's MFMAs run (the global
// loads' latency hides behind them; LDS holds one chunk), and each patch row
// read from LDS once per column shift dx and used for the three row shifts
// dy (a wave's NT output rows need NT + 2 patch rows): per chunk and wave
// 3 (NT + 2) B reads and 9 MT A reads feed 9 MT NT MFMAs.
template <int TAPS, int MT, int NT, bool UPCAT = false, bool PF = true>
__global__ void __launch_bounds__(256, 2) conv_pf_kernel(ConvArgs a) {
  constexpr int kTH = 4 * NT;
  constexpr int HALO = TAPS == 9 ? 1 : 0;
  constexpr int PW = kTW + 2 * HALO, PH = kTH + 2 * HALO;
  constexpr int WS = TAPS * 32 + 8;  // LDS bf16 per output channel of the weight chunk
  constexpr int NPV = PH * PW * 4;             // 16-byte patch vectors per chunk
  constexpr int NWV = MT * 16 * TAPS * 4;      // 16-byte weight vectors per chunk
  constexpr int RP = (NPV + 255) / 256, RW = (NWV + 255) / 256;
  constexpr int NS = 2 * HALO + 1;             // column (and row) shifts
  __shared__ __bf16 patch[PH * PW * kPS];
  __shared__ __bf16 wl[MT * 16 * WS];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tiles_x = (a.W + kTW - 1) / kTW, tiles_y = (a.H + kTH - 1) / kTH;
  int b = blockIdx.x;
  const int tx = b % tiles_x;
  b /= tiles_x;
  const int ty = b % tiles_y;
  const int n = b / tiles_y;
  const int x0 = tx * kTW, y0 = ty * kTH;
  const int cb = blockIdx.y * MT * 16;
  const int Cin = a.Cin, K = a.K;
  bf16x8 rp[RP], rw[RW];
  auto gload = [&](int ci0) {
#pragma unroll
    for (int j = 0; j < RP; ++j) {
      const int v = threadIdx.x + 256 * j;
      bf16x8 val = zero8();
      if (v < NPV) {
        const int pix = v >> 2, q = v & 3;
        const int py = pix / PW, px = pix % PW;
        const int gy = y0 + py - HALO, gx = x0 + px - HALO;
        if (gy >= 0 && gy < a.H && gx >= 0 && gx < a.W) {
          const __bf16* src;
          if (!UPCAT)
            src = a.x + ((size_t)(n * a.H + gy) * a.W + gx) * Cin + ci0;
          else if (ci0 < a.C1)  // upsampled half: pixel (gy/2, gx/2) of x
            src = a.x + ((size_t)(n * (a.H >> 1) + (gy >> 1)) * (a.W >> 1) + (gx >> 1)) * a.C1 + ci0;
          else
            src = a.x2 + ((size_t)(n * a.H + gy) * a.W + gx) * (Cin - a.C1) + (ci0 - a.C1);
          val = load8(src + q * 8);
        }
      }
      rp[j] = val;
    }
#pragma unroll
    for (int j = 0; j < RW; ++j) {
      const int v = threadIdx.x + 256 * j;
      bf16x8 val = zero8();
      if (v < NWV) {
        const int co = v / (TAPS * 4), r = v % (TAPS * 4);
        const int tap = r >> 2, q = r & 3;
        if (cb + co < a.Cout) val = load8(a.w + (size_t)(cb + co) * K + tap * Cin + ci0 + q * 8);
      }
      rw[j] = val;
    }
  };
  auto lstore = [&]() {
#pragma unroll
    for (int j = 0; j < RP; ++j) {
      const int v = threadIdx.x + 256 * j;
      if (v < NPV)
        *reinterpret_cast<bf16x8*>(&patch[(v >> 2) * kPS + (v & 3) * 8]) = rp[j];
    }
#pragma unroll
    for (int j = 0; j < RW; ++j) {
      const int v = threadIdx.x + 256 * j;
      if (v < NWV) {
        const int co = v / (TAPS * 4), r = v % (TAPS * 4);
        *reinterpret_cast<bf16x8*>(&wl[co * WS + (r >> 2) * 32 + (r & 3) * 8]) = rw[j];
      }
    }
  };
  f32x4 acc[MT][NT];
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[m][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  // PF: the next chunk loads into registers during this chunk's MFMAs; else
  // each chunk is loaded and staged before its MFMAs (fewer registers live
  // across them, for single-chunk layers with larger tiles)
  if constexpr (PF) {
    gload(0);
    lstore();
    __syncthreads();
  }
  for (int ci0 = 0; ci0 < Cin; ci0 += 32) {
    if constexpr (!PF) {
      if (ci0 > 0) __syncthreads();
      gload(ci0);
      lstore();
      __syncthreads();
    }
    const bool more = PF && ci0 + 32 < Cin;
    if (more) gload(ci0 + 32);
#pragma unroll
    for (int dx = 0; dx < NS; ++dx) {
      bf16x8 B[NT + 2 * HALO];
#pragma unroll
      for (int r = 0; r < NT + 2 * HALO; ++r) {
        const int row = wave * NT + r, col = (lane & 15) + dx;
        B[r] = *reinterpret_cast<const bf16x8*>(&patch[(row * PW + col) * kPS + (lane >> 4) * 8]);
      }
#pragma unroll
      for (int dy = 0; dy < NS; ++dy) {
        const int tap = dy * NS + dx;
        bf16x8 A[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m)
          A[m] = *reinterpret_cast<const bf16x8*>(&wl[(m * 16 + (lane & 15)) * WS + tap * 32 +
                                                      (lane >> 4) * 8]);
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
          for (int t = 0; t < NT; ++t) acc[m][t] = mfma(A[m], B[t + dy], acc[m][t]);
      }
    }
    if constexpr (PF) {
      __syncthreads();
      if (more) {
        lstore();
        __syncthreads();
      }
    }
  }
  conv_epilogue<MT, NT>(a, acc, n, y0, x0, cb, wave, lane);
}

// The same convolution with persistent blocks: block (b, co) takes the tiles
// b, b + gridDim.x, ... of output-channel block co, and the register
// prefetch of the next input chunk runs across tile boundaries: while the
// last chunk of one tile is multiplied, the first chunk of the block's next
// tile is loaded.  A layer of one or two 32-channel chunks then hides its
// load latency behind MFMAs as a deep layer does, where one tile per block
// leaves every block waiting for its patch.  A single-chunk layer loads its
// weights once per block.  Three blocks per compute unit with 32 output
// channels per block (MT = 2): at most 168 VGPRs and 45 KB of LDS each.
constexpr int kTileBlocksPerCU = 3;

template <int TAPS, int MT, int NT, bool UPCAT = false, bool HEAD = false, bool POOL = false>
__global__ void __launch_bounds__(256, kTileBlocksPerCU) conv_tiles_kernel(ConvArgs a) {
  static_assert(MT == 2, "three blocks per CU fit 32 output channels per block");
  constexpr int kTH = 4 * NT;
  constexpr int HALO = TAPS == 9 ? 1 : 0;
  constexpr int PW = kTW + 2 * HALO, PH = kTH + 2 * HALO;
  constexpr int WS = TAPS * 32 + 8;  // LDS bf16 per output channel of the weight chunk
  constexpr int NPV = PH * PW * 4;             // 16-byte patch vectors per chunk
  constexpr int NWV = MT * 16 * TAPS * 4;      // 16-byte weight vectors per chunk
  constexpr int RP = (NPV + 255) / 256, RW = (NWV + 255) / 256;
  constexpr int NS = 2 * HALO + 1;             // column (and row) shifts
  __shared__ __bf16 patch[PH * PW * kPS];
  __shared__ __bf16 wl[MT * 16 * WS];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tiles_x = (a.W + kTW - 1) / kTW, tiles_y = (a.H + kTH - 1) / kTH;
  const int total = a.N * tiles_x * tiles_y;
  int tile = blockIdx.x;
  if (tile >= total) return;  // uniform over the block
  const int cb = blockIdx.y * MT * 16;
  const int Cin = a.Cin, K = a.K;
  const bool reload_w = Cin > 32;  // one chunk: its weights stay in LDS
  bf16x8 rp[RP], rw[RW];
  auto gload = [&](int t, int ci0, bool wts) {
    const int tx = t % tiles_x, r = t / tiles_x;
    const int x0 = tx * kTW, y0 = (r % tiles_y) * kTH, n = r / tiles_y;
#pragma unroll
    for (int j = 0; j < RP; ++j) {
      const int v = threadIdx.x + 256 * j;
      bf16x8 val = zero8();
      if (v < NPV) {
        const int pix = v >> 2, q = v & 3;
        const int py = pix / PW, px = pix % PW;
        const int gy = y0 + py - HALO, gx = x0 + px - HALO;
        if (gy >= 0 && gy < a.H && gx >= 0 && gx < a.W) {
          const __bf16* src;
          if (!UPCAT)
            src = a.x + ((size_t)(n * a.H + gy) * a.W + gx) * Cin + ci0;
          else if (ci0 < a.C1)  // upsampled half: pixel (gy/2, gx/2) of x
            src = a.x + ((size_t)(n * (a.H >> 1) + (gy >> 1)) * (a.W >> 1) + (gx >> 1)) * a.C1 + ci0;
          else
            src = a.x2 + ((size_t)(n * a.H + gy) * a.W + gx) * (Cin - a.C1) + (ci0 - a.C1);
          val = load8(src + q * 8);
        }
      }
      rp[j] = val;
    }
    if (wts) {
#pragma unroll
      for (int j = 0; j < RW; ++j) {
        const int v = threadIdx.x + 256 * j;
        bf16x8 val = zero8();
        if (v < NWV) {
          const int co = v / (TAPS * 4), r2 = v % (TAPS * 4);
          const int tap = r2 >> 2, q = r2 & 3;
          if (cb + co < a.Cout) val = load8(a.w + (size_t)(cb + co) * K + tap * Cin + ci0 + q * 8);
        }
        rw[j] = val;
      }
    }
  };
  auto lstore = [&](bool wts) {
#pragma unroll
    for (int j = 0; j < RP; ++j) {
      const int v = threadIdx.x + 256 * j;
      if (v < NPV)
        *reinterpret_cast<bf16x8*>(&patch[(v >> 2) * kPS + (v & 3) * 8]) = rp[j];
    }
    if (wts) {
#pragma unroll
      for (int j = 0; j < RW; ++j) {
        const int v = threadIdx.x + 256 * j;
        if (v < NWV) {
          const int co = v / (TAPS * 4), r2 = v % (TAPS * 4);
          *reinterpret_cast<bf16x8*>(&wl[co * WS + (r2 >> 2) * 32 + (r2 & 3) * 8]) = rw[j];
        }
      }
    }
  };
  gload(tile, 0, true);
  lstore(true);
  __syncthreads();
  for (;;) {
    f32x4 acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[m][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int ci0 = 0; ci0 < Cin; ci0 += 32) {
      // the next (tile, chunk) of this block
      int nt = tile, nc = ci0 + 32;
      if (nc >= Cin) {
        nt = tile + (int)gridDim.x;
        nc = 0;
      }
      const bool more = nt < total;
      if (more) gload(nt, nc, reload_w);
#pragma unroll
      for (int dx = 0; dx < NS; ++dx) {
        bf16x8 B[NT + 2 * HALO];
#pragma unroll
        for (int r = 0; r < NT + 2 * HALO; ++r) {
          const int row = wave * NT + r, col = (lane & 15) + dx;
          B[r] = *reinterpret_cast<const bf16x8*>(&patch[(row * PW + col) * kPS + (lane >> 4) * 8]);
        }
#pragma unroll
        for (int dy = 0; dy < NS; ++dy) {
          const int tap = dy * NS + dx;
          bf16x8 A[MT];
#pragma unroll
          for (int m = 0; m < MT; ++m)
            A[m] = *reinterpret_cast<const bf16x8*>(&wl[(m * 16 + (lane & 15)) * WS + tap * 32 +
                                                        (lane >> 4) * 8]);
#pragma unroll
          for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[m][t] = mfma(A[m], B[t + dy], acc[m][t]);
        }
      }
      __syncthreads();
      if (more) {
        lstore(reload_w);
        __syncthreads();
      }
    }
    const int tx = tile % tiles_x, r = tile / tiles_x;
    conv_epilogue<MT, NT, HEAD, POOL>(a, acc, r / tiles_y, (r % tiles_y) * kTH, tx * kTW, cb,
                                      wave, lane);
    tile += (int)gridDim.x;
    if (tile >= total) break;
  }
}

// 3x3 convolution of an 8-channel input (the network's first layer), staged
// through LDS.  K = 9 taps x 8 channels = 72, padded to three MFMA k-steps of
// 32: k-step j takes taps 4j .. 4j+3 (lane group q = lane/16 reads tap 4j+q,
// zero past tap 8), each tap's 8 channels one 16-byte patch vector.  Output
// channels are the M operand and pixels the N operand, as in conv_pf_kernel,
// whose epilogue this shares: a lane stores 4 consecutive channels at once.
template <int MT, int NT>
__global__ void __launch_bounds__(256) conv_c8_kernel(ConvArgs a) {
  constexpr int kTH = 4 * NT, PW = kTW + 2, PH = kTH + 2;
  constexpr int KP = 96 + 8;  // LDS bf16 per output channel of the weights (72 used, padded)
  __shared__ bf16x8 patch[PH * PW];
  __shared__ __bf16 wl[MT * 16 * KP];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tiles_x = (a.W + kTW - 1) / kTW, tiles_y = (a.H + kTH - 1) / kTH;
  int b = blockIdx.x;
  const int tx = b % tiles_x;
  b /= tiles_x;
  const int ty = b % tiles_y;
  const int n = b / tiles_y;
  const int x0 = tx * kTW, y0 = ty * kTH;
  const int cb = blockIdx.y * MT * 16;
  for (int v = threadIdx.x; v < PH * PW; v += 256) {
    const int py = v / PW, px = v % PW;
    const int gy = y0 + py - 1, gx = x0 + px - 1;
    bf16x8 val = zero8();
    if (gy >= 0 && gy < a.H && gx >= 0 && gx < a.W)
      val = load8(a.x + ((size_t)(n * a.H + gy) * a.W + gx) * 8);
    patch[v] = val;
  }
  for (int v = threadIdx.x; v < MT * 16 * 12; v += 256) {  // 12 taps of 8 per channel
    const int co = v / 12, tap = v % 12;
    bf16x8 val = zero8();
    if (tap < 9 && cb + co < a.Cout) val = load8(a.w + (size_t)(cb + co) * 72 + tap * 8);
    *reinterpret_cast<bf16x8*>(&wl[co * KP + tap * 8]) = val;
  }
  __syncthreads();
  f32x4 acc[MT][NT];
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[m][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  const int q = lane >> 4;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int tap = 4 * j + q;
    const int dy = tap < 9 ? tap / 3 : 0, dx = tap < 9 ? tap % 3 : 0;
    bf16x8 A[MT], B[NT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
      A[m] = *reinterpret_cast<const bf16x8*>(&wl[(m * 16 + (lane & 15)) * KP + tap * 8]);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const bf16x8 v = patch[(wave * NT + t + dy) * PW + (lane & 15) + dx];
      B[t] = tap < 9 ? v : zero8();
    }
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[m][t] = mfma(A[m], B[t], acc[m][t]);
  }
  conv_epilogue<MT, NT>(a, acc, n, y0, x0, cb, wave, lane);
}

// dgrad weights: Wf[ci][tap'*Cout + co] = W[co][(TAPS-1-tap')*Cin + ci]
__global__ void flip_weights_kernel(const __bf16* __restrict__ w, __bf16* wf, int Cin, int Cout,
                                    int taps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = Cin * Cout * taps;
  if (i >= n) return;
  const int ci = i / (taps * Cout);
  const int rem = i % (taps * Cout);
  const int tp = rem / Cout, co = rem % Cout;
  wf[i] = w[(size_t)co * taps * Cin + (taps - 1 - tp) * Cin + ci];
}

struct WgradArgs {
  const __bf16* __restrict__ x;  // [P][Cin] input of the convolution
  const __bf16* __restrict__ g;  // [P][Cout] gradient of its output
  float* part;                   // [splits][Cout][K]
  int N, H, W, Cin, Cout, K, taps;
  int chunk;                     // pixels per split (multiple of 32)
};

// One wave per block: a 32 (co) x 32 (k) tile of dW over one pixel range.
__global__ void __launch_bounds__(64) conv_wgrad_kernel(WgradArgs a) {
  __shared__ __bf16 Gs[32][32 + 8];  // [pixel][co]
  __shared__ __bf16 Xs[32][32 + 8];  // [pixel][k]
  const int lane = threadIdx.x;
  const int P = a.N * a.H * a.W;
  const int cb = blockIdx.x * 32, kb = blockIdx.y * 32;
  const int p_begin = blockIdx.z * a.chunk;
  const int p_end = min(P, p_begin + a.chunk);
  f32x4 acc[2][2];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c) acc[r][c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int p0 = p_begin; p0 < p_end; p0 += 32) {
    // stage: 32 pixels x 4 groups of 8 channels, for G and for the tap-shifted X
    for (int t = lane; t < 128; t += 64) {
      const int pi = t >> 2, grp = t & 3;
      const int p = p0 + pi;
      const int co = cb + grp * 8;
      bf16x8 gv = zero8();
      if (p < p_end && co < a.Cout) gv = load8(a.g + (size_t)p * a.Cout + co);
      const int k0 = kb + grp * 8;
      bf16x8 xv = zero8();
      if (p < p_end && k0 < a.K) {
        const int tap = k0 / a.Cin, ci0 = k0 - tap * a.Cin;
        const int dy = a.taps == 9 ? tap / 3 - 1 : 0, dx = a.taps == 9 ? tap % 3 - 1 : 0;
        const int xx = p % a.W, yy = (p / a.W) % a.H, nn = p / (a.W * a.H);
        const int y2 = yy + dy, x2 = xx + dx;
        if (y2 >= 0 && y2 < a.H && x2 >= 0 && x2 < a.W)
          xv = load8(a.x + ((size_t)(nn * a.H + y2) * a.W + x2) * a.Cin + ci0);
      }
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        Gs[pi][grp * 8 + i] = gv[i];
        Xs[pi][grp * 8 + i] = xv[i];
      }
    }
    __syncthreads();
    // A[m = co][kk = pixel], B[kk = pixel][n = k]
    bf16x8 A[2], B[2];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int i = 0; i < 8; ++i) A[r][i] = Gs[(lane >> 4) * 8 + i][r * 16 + (lane & 15)];
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int i = 0; i < 8; ++i) B[c][i] = Xs[(lane >> 4) * 8 + i][c * 16 + (lane & 15)];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int c = 0; c < 2; ++c) acc[r][c] = mfma(A[r], B[c], acc[r][c]);
    __syncthreads();
  }
  float* out = a.part + (size_t)blockIdx.z * a.Cout * a.K;
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int co = cb + r * 16 + (lane >> 4) * 4 + i;
        const int k = kb + c * 16 + (lane & 15);
        if (co < a.Cout && k < a.K) out[(size_t)co * a.K + k] = acc[r][c][i];
      }
}

// Weight gradient for Cin and Cout multiples of 32, TAPS = 9, staged through
// LDS: dW[co][tap*Cin + ci] = sum_p g[p][co] x[p + tap][ci].  The reduction
// runs over pixels, so both MFMA operands need 8 consecutive pixels of one
// channel per lane: each 16x8 pixel tile is staged transposed (channel-major)
// -- G^T [16*MT][128 pixels] and X^T [32][10 rows][24] (the tile with its
// 1-pixel halo) -- four pixels per 8-byte LDS write.  A block owns 16*MT
// output channels x one 32-channel input chunk x all 9 taps (18 N tiles of
// 16), wave w the N tiles w, w+4, ...; it sums its range of pixel tiles into
// one FP32 partial.  The +1 / +2 pixel shifts of the taps read two aligned
// 16-byte rows of X^T and realign in registers.
constexpr int kWgTW = 16, kWgTH = 8, kWgPix = kWgTW * kWgTH;
constexpr int kXR = 24;  // X^T row: 18 halo columns padded to 24

__device__ inline bf16x8 shift_pixels(bf16x8 lo, bf16x8 hi, int dx) {
  typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
  if (dx == 0) return lo;
  const u32x4 a = __builtin_bit_cast(u32x4, lo), b = __builtin_bit_cast(u32x4, hi);
  u32x4 r;
  if (dx == 2) {
    r = u32x4{a[1], a[2], a[3], b[0]};
  } else {  // one bf16 to the left: v_alignbyte of neighbouring dwords
    r[0] = __builtin_amdgcn_alignbyte(a[1], a[0], 2);
    r[1] = __builtin_amdgcn_alignbyte(a[2], a[1], 2);
    r[2] = __builtin_amdgcn_alignbyte(a[3], a[2], 2);
    r[3] = __builtin_amdgcn_alignbyte(b[0], a[3], 2);
  }
  return __builtin_bit_cast(bf16x8, r);
}

struct WgArgs {
  const __bf16* __restrict__ x;
  const __bf16* __restrict__ g;
  float* part;  // [splits][Cout][K]
  int N, H, W, Cin, Cout, K;
  int tiles_x, tiles_y, tiles;
  int tiles_per_split;
};

template <int MT>
__global__ void __launch_bounds__(256) wgrad_lds_kernel(WgArgs a) {
  constexpr int NCO = 16 * MT;
  constexpr int GS = kWgPix + 8;  // G^T row (pixels) with padding
  __shared__ __bf16 gt[NCO * GS];
  __shared__ __bf16 xt[32 * (kWgTH + 2) * kXR];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int cb = blockIdx.x * NCO, ci0 = blockIdx.y * 32;
  const int t_begin = blockIdx.z * a.tiles_per_split;
  const int t_end = min(a.tiles, t_begin + a.tiles_per_split);
  const int Cin = a.Cin, Cout = a.Cout;
  typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
  f32x4 acc[MT][5];
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int j = 0; j < 5; ++j) acc[m][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int t = t_begin; t < t_end; ++t) {
    const int tx = t % a.tiles_x, ty = (t / a.tiles_x) % a.tiles_y;
    const int n = t / (a.tiles_x * a.tiles_y);
    const int x0 = tx * kWgTW, y0 = ty * kWgTH;
    // G^T: task = (4 pixels of a row, 8 channels); pixels off the image are 0
    for (int task = threadIdx.x; task < (kWgPix / 4) * (NCO / 8); task += 256) {
      const int pq = task % (kWgPix / 4), cq = task / (kWgPix / 4);
      const int pix = pq * 4, row = pix / kWgTW, col = pix % kWgTW;
      bf16x8 v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int gy = y0 + row, gx = x0 + col + u;
        v[u] = zero8();
        if (gy < a.H && gx < a.W)
          v[u] = load8(a.g + ((size_t)(n * a.H + gy) * a.W + gx) * Cout + cb + cq * 8);
      }
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        bf16x4 w4 = {v[0][c], v[1][c], v[2][c], v[3][c]};
        *reinterpret_cast<bf16x4*>(&gt[(cq * 8 + c) * GS + pix]) = w4;
      }
    }
    // X^T: task = (row of the halo tile, 4 halo columns, 8 channels)
    for (int task = threadIdx.x; task < (kWgTH + 2) * 5 * 4; task += 256) {
      const int r = task % (kWgTH + 2), rest = task / (kWgTH + 2);
      const int cg = rest % 5, cq = rest / 5;
      bf16x8 v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int hc = cg * 4 + u;  // halo column 0..19 (18, 19 are padding)
        const int gy = y0 + r - 1, gx = x0 + hc - 1;
        v[u] = zero8();
        if (hc < kWgTW + 2 && gy >= 0 && gy < a.H && gx >= 0 && gx < a.W)
          v[u] = load8(a.x + ((size_t)(n * a.H + gy) * a.W + gx) * Cin + ci0 + cq * 8);
      }
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        bf16x4 w4 = {v[0][c], v[1][c], v[2][c], v[3][c]};
        *reinterpret_cast<bf16x4*>(&xt[((cq * 8 + c) * (kWgTH + 2) + r) * kXR + cg * 4]) = w4;
      }
    }
    __syncthreads();
    // 4 k-steps of 32 pixels (tile rows 2s, 2s+1)
#pragma unroll
    for (int s = 0; s < kWgPix / 32; ++s) {
      const int grp = lane >> 4;  // 8-pixel group of this lane
      const int R = 2 * s + (grp >> 1), c0 = (grp & 1) * 8;
      bf16x8 A[MT];
#pragma unroll
      for (int m = 0; m < MT; ++m)
        A[m] = *reinterpret_cast<const bf16x8*>(&gt[(m * 16 + (lane & 15)) * GS + R * kWgTW + c0]);
#pragma unroll
      for (int j = 0; j < 5; ++j) {
        const int nt = wave + 4 * j;  // N tile: tap nt/2, channels 16*(nt%2) + 0..15
        if (nt >= 18) continue;
        const int tap = nt >> 1, dy = tap / 3, dx = tap % 3;
        const int ci = (nt & 1) * 16 + (lane & 15);
        const __bf16* row = &xt[(ci * (kWgTH + 2) + R + dy) * kXR + c0];
        const bf16x8 lo = *reinterpret_cast<const bf16x8*>(row);
        const bf16x8 hi = *reinterpret_cast<const bf16x8*>(row + 8);
        const bf16x8 B = shift_pixels(lo, hi, dx);
#pragma unroll
        for (int m = 0; m < MT; ++m) acc[m][j] = mfma(A[m], B, acc[m][j]);
      }
    }
    __syncthreads();
  }
  float* out = a.part + (size_t)blockIdx.z * Cout * a.K;
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    const int nt = wave + 4 * j;
    if (nt >= 18) continue;
    const int k = (nt >> 1) * Cin + ci0 + (nt & 1) * 16 + (lane & 15);
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int co = cb + m * 16 + (lane >> 4) * 4 + i;
        out[(size_t)co * a.K + k] = acc[m][j][i];
      }
  }
}

// dW[i] = sum over splits of part[k][i] in a fixed order (deterministic):
// 8 interleaved slices of k, each summed in ascending k, then the slice sums
// in slice order.  32 outputs x 8 slices per block keep the loads coalesced
// and give every output 8 independent chains.
constexpr int kRedSlices = 8;

__global__ void __launch_bounds__(256) wgrad_reduce_kernel(const float* __restrict__ part,
                                                           float* dw, int n, int splits) {
  __shared__ float acc[kRedSlices][32];
  const int e = threadIdx.x & 31, sl = threadIdx.x >> 5;
  const int i = blockIdx.x * 32 + e;
  float s = 0.0f;
  if (i < n) {
    // eight loads in flight per iteration, added in ascending k as before
    int k = sl;
    for (; k + 7 * kRedSlices < splits; k += 8 * kRedSlices) {
      float v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = part[(size_t)(k + u * kRedSlices) * n + i];
#pragma unroll
      for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; k < splits; k += kRedSlices) s += part[(size_t)k * n + i];
  }
  acc[sl][e] = s;
  __syncthreads();
  if (sl == 0 && i < n) {
    float t = 0.0f;
#pragma unroll
    for (int q = 0; q < kRedSlices; ++q) t += acc[q][e];
    dw[i] = t;
  }
}

// part[b][c] = sum of g[p][c] over block b's rows, in a fixed order: for
// C <= 256 thread t covers channel t % C and every (256/C)-th row from row
// t / C (coalesced: neighbouring threads read neighbouring channels), then
// the row-group sums are added in group order.
__global__ void __launch_bounds__(256) bias_grad_kernel(const __bf16* __restrict__ g, float* part,
                                                        int P, int C, int rows_per_block) {
  __shared__ float acc[256];
  const int t = threadIdx.x;
  const int p0 = blockIdx.x * rows_per_block;
  const int p1 = min(P, p0 + rows_per_block);
  if (C <= 256) {
    const int groups = 256 / C;
    const int c = t % C, r = t / C;
    float s = 0.0f;
    if (r < groups)
      for (int p = p0 + r; p < p1; p += groups) s += (float)g[(size_t)p * C + c];
    acc[t] = s;
    __syncthreads();
    if (t < C) {
      float u = 0.0f;
      for (int q = 0; q < groups; ++q) u += acc[q * C + t];
      part[(size_t)blockIdx.x * C + t] = u;
    }
  } else {
    for (int c = t; c < C; c += 256) {
      float s = 0.0f;
      for (int p = p0; p < p1; ++p) s += (float)g[(size_t)p * C + c];
      part[(size_t)blockIdx.x * C + c] = s;
    }
  }
}

// The same sums 8 channels per thread (C % 8 == 0, C <= 1024, g 16-byte
// aligned), 1024 threads per block: thread t covers channels 8 (t % C8) ..
// + 7 of every G-th row from row t / C8 (G = 1024 / C8 row groups), each
// load 16 bytes, eight rows in flight; then the G group sums are folded in
// LDS by halves (group r + h onto group r), a fixed order.
__device__ inline void fold_groups(float* acc, int t, int C8, int groups) {
  const int c8 = t % C8, r = t / C8;
  for (int n = groups; n > 1;) {
    const int h = (n + 1) / 2;  // groups [h, n) onto [0, n - h)
    if (r < n - h) {
#pragma unroll
      for (int k = 0; k < 8; ++k) acc[(r * C8 + c8) * 8 + k] += acc[((r + h) * C8 + c8) * 8 + k];
    }
    __syncthreads();
    n = h;
  }
}

__global__ void __launch_bounds__(1024) bias_grad_v8_kernel(const __bf16* __restrict__ g,
                                                            float* part, int P, int C,
                                                            int rows_per_block) {
  __shared__ float acc[1024 * 8];
  const int t = threadIdx.x, C8 = C / 8, groups = 1024 / C8;
  const int c8 = t % C8, r = t / C8;
  const int p0 = blockIdx.x * rows_per_block;
  const int p1 = min(P, p0 + rows_per_block);
  float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (r < groups) {
    int p = p0 + r;
    for (; p + 7 * groups < p1; p += 8 * groups) {
      bf16x8 v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = load8(g + (size_t)(p + u * groups) * C + 8 * c8);
#pragma unroll
      for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int k = 0; k < 8; ++k) s[k] += (float)v[u][k];
    }
    for (; p < p1; p += groups) {
      const bf16x8 v = load8(g + (size_t)p * C + 8 * c8);
#pragma unroll
      for (int k = 0; k < 8; ++k) s[k] += (float)v[k];
    }
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) acc[t * 8 + k] = s[k];
  __syncthreads();
  fold_groups(acc, t, C8, groups);
  for (int c = t; c < C; c += 1024) part[(size_t)blockIdx.x * C + c] = acc[c];
}

// db[c] = sum over the nb block partials part[b][c] (C % 8 == 0, C <= 1024):
// one block of 1024 threads, thread t adds the partials b = t / C8, + G, ...
// of 8 channels (four in flight), then the groups fold as above.
__global__ void __launch_bounds__(1024) bias_reduce_kernel(const float* __restrict__ part,
                                                           float* db, int C, int nb) {
  __shared__ float acc[1024 * 8];
  const int t = threadIdx.x, C8 = C / 8, groups = 1024 / C8;
  const int c8 = t % C8, r = t / C8;
  float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (r < groups) {
    int b = r;
    for (; b + 3 * groups < nb; b += 4 * groups) {
      float4 v[4][2];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float4* q = reinterpret_cast<const float4*>(part + (size_t)(b + u * groups) * C + 8 * c8);
        v[u][0] = q[0];
        v[u][1] = q[1];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        s[0] += v[u][0].x; s[1] += v[u][0].y; s[2] += v[u][0].z; s[3] += v[u][0].w;
        s[4] += v[u][1].x; s[5] += v[u][1].y; s[6] += v[u][1].z; s[7] += v[u][1].w;
      }
    }
    for (; b < nb; b += groups) {
      const float4* q = reinterpret_cast<const float4*>(part + (size_t)b * C + 8 * c8);
      const float4 v0 = q[0], v1 = q[1];
      s[0] += v0.x; s[1] += v0.y; s[2] += v0.z; s[3] += v0.w;
      s[4] += v1.x; s[5] += v1.y; s[6] += v1.z; s[7] += v1.w;
    }
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) acc[t * 8 + k] = s[k];
  __syncthreads();
  fold_groups(acc, t, C8, groups);
  for (int c = t; c < C; c += 1024) db[c] = acc[c];
}

__global__ void relu_bwd_kernel(const __bf16* __restrict__ y, const __bf16* __restrict__ g,
                                __bf16* out, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = (float)y[i] > 0.0f ? g[i] : (__bf16)0.0f;
}

// 2x2 average pool, NHWC, H and W even
__global__ void pool2_fwd_kernel(const __bf16* __restrict__ x, __bf16* y, int N, int H, int W,
                                 int C) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int Ho = H / 2, Wo = W / 2;
  const size_t n = (size_t)N * Ho * Wo * C;
  if (i >= n) return;
  const int c = (int)(i % C);
  const size_t q = i / C;
  const int xo = (int)(q % Wo), yo = (int)((q / Wo) % Ho), nn = (int)(q / ((size_t)Wo * Ho));
  const size_t b = ((size_t)(nn * H + 2 * yo) * W + 2 * xo) * C + c;
  const float s = (float)x[b] + (float)x[b + C] + (float)x[b + (size_t)W * C] +
                  (float)x[b + (size_t)W * C + C];
  y[i] = (__bf16)(0.25f * s);
}

__global__ void pool2_bwd_kernel(const __bf16* __restrict__ gy, __bf16* gx, int N, int H, int W,
                                 int C) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n = (size_t)N * H * W * C;
  if (i >= n) return;
  const int c = (int)(i % C);
  const size_t q = i / C;
  const int x = (int)(q % W), y = (int)((q / W) % H), nn = (int)(q / ((size_t)W * H));
  const int Ho = H / 2, Wo = W / 2;
  gx[i] = (__bf16)(0.25f * (float)gy[((size_t)(nn * Ho + y / 2) * Wo + x / 2) * C + c]);
}

// nearest x2 upsample of x [N][H][W][C] into channels [c_off, c_off+C) of
// y [N][2H][2W][Cy] (the first half of a concat)
__global__ void up2_fwd_kernel(const __bf16* __restrict__ x, __bf16* y, int N, int H, int W, int C,
                               int Cy, int c_off) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int Ho = 2 * H, Wo = 2 * W;
  const size_t n = (size_t)N * Ho * Wo * C;
  if (i >= n) return;
  const int c = (int)(i % C);
  const size_t q = i / C;
  const int xo = (int)(q % Wo), yo = (int)((q / Wo) % Ho), nn = (int)(q / ((size_t)Wo * Ho));
  y[q * Cy + c_off + c] = x[((size_t)(nn * H + yo / 2) * W + xo / 2) * C + c];
}

__global__ void up2_bwd_kernel(const __bf16* __restrict__ gy, __bf16* gx, int N, int H, int W,
                               int C, int Cy, int c_off) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n = (size_t)N * H * W * C;
  if (i >= n) return;
  const int c = (int)(i % C);
  const size_t q = i / C;
  const int x = (int)(q % W), y = (int)((q / W) % H), nn = (int)(q / ((size_t)W * H));
  const int Wo = 2 * W;
  const size_t b = ((size_t)(nn * 2 * H + 2 * y) * Wo + 2 * x) * Cy + c_off + c;
  const float s = (float)gy[b] + (float)gy[b + Cy] + (float)gy[b + (size_t)Wo * Cy] +
                  (float)gy[b + (size_t)Wo * Cy + Cy];
  gx[i] = (__bf16)s;
}

// copy x [P][C] into channels [c_off, c_off+C) of y [P][Cy] (dir 0), or back (dir 1)
__global__ void channels_copy_kernel(const __bf16* __restrict__ src, __bf16* dst, size_t P, int C,
                                     int Cy, int c_off, int dir) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P * C) return;
  const size_t p = i / C;
  const int c = (int)(i % C);
  if (dir == 0) dst[p * Cy + c_off + c] = src[i];
  else dst[i] = src[p * Cy + c_off + c];
}

__global__ void add_kernel(const __bf16* __restrict__ a, const __bf16* __restrict__ b, __bf16* y,
                           size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  y[i] = (__bf16)((float)a[i] + (float)b[i]);
}

// The same element-wise kernels on 8 channels (16 bytes) per thread, used
// when every channel count and offset is a multiple of 8.
__device__ inline void store8(__bf16* p, bf16x8 v) { *reinterpret_cast<bf16x8*>(p) = v; }

__global__ void relu_bwd_v8_kernel(const __bf16* __restrict__ y, const __bf16* __restrict__ g,
                                   __bf16* out, size_t n8) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n8) return;
  const bf16x8 a = load8(y + 8 * i), b = load8(g + 8 * i);
  bf16x8 r;
#pragma unroll
  for (int k = 0; k < 8; ++k) r[k] = (float)a[k] > 0.0f ? b[k] : (__bf16)0.0f;
  store8(out + 8 * i, r);
}

__global__ void add_v8_kernel(const __bf16* __restrict__ a, const __bf16* __restrict__ b,
                              __bf16* y, size_t n8) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n8) return;
  const bf16x8 u = load8(a + 8 * i), v = load8(b + 8 * i);
  bf16x8 r;
#pragma unroll
  for (int k = 0; k < 8; ++k) r[k] = (__bf16)((float)u[k] + (float)v[k]);
  store8(y + 8 * i, r);
}

__global__ void pool2_fwd_v8_kernel(const __bf16* __restrict__ x, __bf16* y, int N, int H, int W,
                                    int C) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int Ho = H / 2, Wo = W / 2, C8 = C / 8;
  if (i >= (size_t)N * Ho * Wo * C8) return;
  const int c = (int)(i % C8) * 8;
  const size_t q = i / C8;
  const int xo = (int)(q % Wo), yo = (int)((q / Wo) % Ho), nn = (int)(q / ((size_t)Wo * Ho));
  const size_t b = ((size_t)(nn * H + 2 * yo) * W + 2 * xo) * C + c;
  const bf16x8 v0 = load8(x + b), v1 = load8(x + b + C), v2 = load8(x + b + (size_t)W * C),
               v3 = load8(x + b + (size_t)W * C + C);
  bf16x8 r;
#pragma unroll
  for (int k = 0; k < 8; ++k)
    r[k] = (__bf16)(0.25f * ((float)v0[k] + (float)v1[k] + (float)v2[k] + (float)v3[k]));
  store8(y + q * C + c, r);
}

__global__ void pool2_bwd_v8_kernel(const __bf16* __restrict__ gy, __bf16* gx, int N, int H,
                                    int W, int C) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int C8 = C / 8;
  if (i >= (size_t)N * H * W * C8) return;
  const int c = (int)(i % C8) * 8;
  const size_t q = i / C8;
  const int x = (int)(q % W), y = (int)((q / W) % H), nn = (int)(q / ((size_t)W * H));
  const int Ho = H / 2, Wo = W / 2;
  const bf16x8 v = load8(gy + ((size_t)(nn * Ho + y / 2) * Wo + x / 2) * C + c);
  bf16x8 r;
#pragma unroll
  for (int k = 0; k < 8; ++k) r[k] = (__bf16)(0.25f * (float)v[k]);
  store8(gx + q * C + c, r);
}

__global__ void up2_fwd_v8_kernel(const __bf16* __restrict__ x, __bf16* y, int N, int H, int W,
                                  int C, int Cy, int c_off) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int Ho = 2 * H, Wo = 2 * W, C8 = C / 8;
  if (i >= (size_t)N * Ho * Wo * C8) return;
  const int c = (int)(i % C8) * 8;
  const size_t q = i / C8;
  const int xo = (int)(q % Wo), yo = (int)((q / Wo) % Ho), nn = (int)(q / ((size_t)Wo * Ho));
  store8(y + q * Cy + c_off + c, load8(x + ((size_t)(nn * H + yo / 2) * W + xo / 2) * C + c));
}

__global__ void up2_bwd_v8_kernel(const __bf16* __restrict__ gy, __bf16* gx, int N, int H, int W,
                                  int C, int Cy, int c_off) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int C8 = C / 8;
  if (i >= (size_t)N * H * W * C8) return;
  const int c = (int)(i % C8) * 8;
  const size_t q = i / C8;
  const int x = (int)(q % W), y = (int)((q / W) % H), nn = (int)(q / ((size_t)W * H));
  const int Wo = 2 * W;
  const size_t b = ((size_t)(nn * 2 * H + 2 * y) * Wo + 2 * x) * Cy + c_off + c;
  const bf16x8 v0 = load8(gy + b), v1 = load8(gy + b + Cy), v2 = load8(gy + b + (size_t)Wo * Cy),
               v3 = load8(gy + b + (size_t)Wo * Cy + Cy);
  bf16x8 r;
#pragma unroll
  for (int k = 0; k < 8; ++k)
    r[k] = (__bf16)((float)v0[k] + (float)v1[k] + (float)v2[k] + (float)v3[k]);
  store8(gx + q * C + c, r);
}

__global__ void channels_copy_v8_kernel(const __bf16* __restrict__ src, __bf16* dst, size_t P,
                                        int C, int Cy, int c_off, int dir) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int C8 = C / 8;
  if (i >= P * C8) return;
  const size_t p = i / C8;
  const int c = (int)(i % C8) * 8;
  if (dir == 0) store8(dst + p * Cy + c_off + c, load8(src + p * C + c));
  else store8(dst + p * C + c, load8(src + p * Cy + c_off + c));
}

// masked MSE: loss partials and d(loss)/dy = 2 (y - t) m / count
__global__ void mse_kernel(const float* __restrict__ y, const float* __restrict__ t,
                           const float* __restrict__ m, __bf16* g, float* part, size_t n,
                           float inv_count) {
  __shared__ float red[256];
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (i < n) {
    const float mk = m ? m[i] : 1.0f;
    const float d = (y[i] - t[i]) * mk;
    l = d * d;
    g[i] = (__bf16)(2.0f * d * inv_count);
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Adam (Kingma & Ba 2015) on FP32 master weights; optional bf16 copy
// With wf, the rounded weight also goes to its place in the data gradient's
// flipped weights (flip_weights_kernel's layout), so the backward pass needs
// no flip of its own: w[co][t*Cin + ci] -> wf[ci][(taps-1-t)*Cout + co].
__global__ void adam_kernel(float* w, const float* __restrict__ g, float* m, float* v,
                            __bf16* w_bf, size_t n, float lr, float b1, float b2, float eps,
                            float c1, float c2, float wd, __bf16* wf = nullptr, int Cin = 0,
                            int Cout = 0, int taps = 0) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i] + wd * w[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (w_bf) w_bf[i] = (__bf16)wi;
  if (wf) {
    const int k = taps * Cin;
    const int co = (int)(i / k), rem = (int)(i % k);
    const int t = rem / Cin, ci = rem % Cin;
    wf[(size_t)ci * taps * Cout + (size_t)(taps - 1 - t) * Cout + co] = (__bf16)wi;
  }
}

__global__ void f32_to_bf16_kernel(const float* __restrict__ x, __bf16* y, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (__bf16)x[i];
}

__global__ void bf16_to_f32_kernel(const __bf16* __restrict__ x, float* y, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (float)x[i];
}

}  // namespace shy_nn

using namespace shy_nn;

namespace {
thread_local std::string g_nn_err;
int nn_fail(const std::string& m) {
  g_nn_err = m;
  return SHY_NN_ERR;
}
int nn_check(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return nn_fail(std::string(what) + ": " + hipGetErrorString(e));
  return 0;
}
inline unsigned blocks(size_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

// Compute units of the current device (looked up once per device).
int device_cus() {
  static std::atomic<int> cus[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  int c = cus[dev].load(std::memory_order_relaxed);
  if (c == 0) {
    if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c < 1)
      c = 256;
    cus[dev].store(c, std::memory_order_relaxed);
  }
  return c;
}

// Persistent tile loop (conv_tiles_kernel): three blocks per compute unit
// over the layer's tiles of 16 x 16 pixels, 32 output channels per block.  A
// 64-channel layer takes two such blocks per tile: at 16 x 128^2, 32->64
// 31.1 -> 27.8 us and 64->64 41.7 -> 40.7 us against one 64-channel block
// at two blocks per CU (the 64-channel block needs 255 VGPRs).
int launch_tiles(const ConvArgs& a, hipStream_t s, const char* what) {
  constexpr int MT = 2, NT = 4;
  const long long tiles = (long long)a.N * ((a.H + 4 * NT - 1) / (4 * NT)) * ((a.W + kTW - 1) / kTW);
  if (tiles > INT32_MAX) return nn_fail(std::string(what) + ": too many tiles");
  const int gy = a.Cout / (16 * MT);
  const long long cap = std::max(1LL, (long long)kTileBlocksPerCU * device_cus() / gy);
  const dim3 g((unsigned)std::min(tiles, cap), (unsigned)gy);
  if (a.x2) conv_tiles_kernel<9, MT, NT, true, false><<<g, 256, 0, s>>>(a);
  else if (a.head_w) conv_tiles_kernel<9, MT, NT, false, true><<<g, 256, 0, s>>>(a);
  else if (a.pool_y) conv_tiles_kernel<9, MT, NT, false, false, true><<<g, 256, 0, s>>>(a);
  else conv_tiles_kernel<9, MT, NT, false, false><<<g, 256, 0, s>>>(a);
  return nn_check(what);
}

// conv_pf_kernel for Cin and Cout multiples of 32, 32 output channels per
// block (MT = 2; 64-channel blocks at 32^2 and 128^2 ran 4-16% slower: half
// the blocks, and 210-256 VGPRs), 16x16 tiles (4 rows per wave) when they
// still give >= 1024 blocks, else 16x8.
int launch_pf(const ConvArgs& a, int taps, hipStream_t s, const char* what) {
  const long long tiles16 = (long long)a.N * ((a.H + 15) / 16) * ((a.W + kTW - 1) / kTW);
  // 3x3 layers whose 16x16 tiles give every CU at least one: the persistent
  // tile loop, decoder inputs included (at batch 16, against the one-tile
  // blocks below: 32->32 at 256^2 50.7 -> 43.4 us, [up2, skip] 96->32 at
  // 256^2 105 -> 83 us, 96->96 at 64^2 23.0 -> 21.5 us; at 32^2, 64 tiles,
  // the one-tile blocks' 16x8 tiles are faster)
  if (taps == 9 && tiles16 >= device_cus()) return launch_tiles(a, s, what);
  const unsigned gy = (unsigned)(a.Cout / 32);
  const long long tiles8 = (long long)a.N * ((a.H + 7) / 8) * ((a.W + kTW - 1) / kTW);
  if (tiles8 > INT32_MAX) return nn_fail(std::string(what) + ": too many tiles");
  const bool big = tiles16 * gy >= 1024;
  const dim3 g((unsigned)(big ? tiles16 : tiles8), gy);
  const long long tiles32 = (long long)a.N * ((a.H + 31) / 32) * ((a.W + kTW - 1) / kTW);
  if (!a.x2 && taps == 9 && a.Cin == 32 && tiles32 * gy >= 1024) {
    // one input chunk: 16x32 tiles staged without prefetch, so each block
    // reloads the layer's weights for twice the pixels
    conv_pf_kernel<9, 2, 8, false, false><<<dim3((unsigned)tiles32, gy), 256, 0, s>>>(a);
    return nn_check(what);
  }
  if (a.x2) {  // the input is [up2(x), x2] (taps 9)
    if (big) conv_pf_kernel<9, 2, 4, true><<<g, 256, 0, s>>>(a);
    else conv_pf_kernel<9, 2, 2, true><<<g, 256, 0, s>>>(a);
  } else if (big) {
    if (taps == 9) conv_pf_kernel<9, 2, 4><<<g, 256, 0, s>>>(a);
    else conv_pf_kernel<1, 2, 4><<<g, 256, 0, s>>>(a);
  } else {
    if (taps == 9) conv_pf_kernel<9, 2, 2><<<g, 256, 0, s>>>(a);
    else conv_pf_kernel<1, 2, 2><<<g, 256, 0, s>>>(a);
  }
  return nn_check(what);
}

// Device address of `scratch` on the current device: looked up once per
// device (on that device's first allocation-free call) and kept.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_dev[kMaxDevices] = {};

int noalloc_scratch(float** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return nn_fail(std::string("noalloc: hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return nn_fail("noalloc: device index beyond 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return nn_fail(std::string("noalloc: hipGetSymbolAddress(scratch): ") +
                     hipGetErrorString(e));
    g_scratch_dev[dev] = static_cast<float*>(p);
  }
  *out = g_scratch_dev[dev];
  return 0;
}
}  // namespace

extern "C" {

const char* shy_nn_last_error(void) { return g_nn_err.c_str(); }

int shy_nn_conv_fwd(const void* x, const void* w, const float* bias, const void* res, void* y_bf16,
                    float* y_f32, int N, int H, int W, int Cin, int Cout, int taps, int relu,
                    void* stream) {
  if (!x || !w || (!y_bf16 && !y_f32) || (y_bf16 && y_f32)) return nn_fail("conv_fwd: bad pointers");
  if ((uintptr_t)bias & 15) return nn_fail("conv_fwd: bias must be 16-byte aligned");
  if (N < 1 || H < 1 || W < 1 || Cin < 8 || Cin % 8 || Cout < 1 || (taps != 9 && taps != 1))
    return nn_fail("conv_fwd: need Cin a multiple of 8 and taps 9 or 1");
  ConvArgs a;
  a.x = (const __bf16*)x;
  a.w = (const __bf16*)w;
  a.bias = bias;
  a.res = (const __bf16*)res;
  a.y_bf = (__bf16*)y_bf16;
  a.y_f = y_f32;
  a.N = N; a.H = H; a.W = W; a.Cin = Cin; a.Cout = Cout; a.K = taps * Cin;
  a.relu = relu;
  a.x2 = nullptr;
  a.C1 = 0;
  a.gres = nullptr;
  a.mask = nullptr;
  const long long P = (long long)N * H * W;
  if (P > INT32_MAX / 2) return nn_fail("conv_fwd: too many pixels");
  hipStream_t s = (hipStream_t)stream;
  if (Cin % 32 == 0 && Cout % 32 == 0) return launch_pf(a, taps, s, "conv_fwd");
  if (Cin == 8 && taps == 9 && Cout % 32 == 0) {  // the first layer
    const int mt = Cout % 64 == 0 ? 4 : 2;
    const long long tiles = (long long)N * ((H + 15) / 16) * ((W + kTW - 1) / kTW);
    if (tiles > INT32_MAX) return nn_fail("conv_fwd: too many tiles");
    const dim3 g((unsigned)tiles, (unsigned)(Cout / (16 * mt)));
    if (mt == 4) conv_c8_kernel<4, 4><<<g, 256, 0, s>>>(a);
    else conv_c8_kernel<2, 4><<<g, 256, 0, s>>>(a);
    return nn_check("conv_fwd");
  }
  const unsigned gx = (unsigned)((P + 127) / 128);
  if (Cout > 16) {
    dim3 g(gx, (unsigned)((Cout + 31) / 32));
    if (taps == 9) conv_fwd_kernel<9, 2><<<g, 256, 0, s>>>(a);
    else conv_fwd_kernel<1, 2><<<g, 256, 0, s>>>(a);
  } else {
    dim3 g(gx, 1);
    if (taps == 9) conv_fwd_kernel<9, 1><<<g, 256, 0, s>>>(a);
    else conv_fwd_kernel<1, 1><<<g, 256, 0, s>>>(a);
  }
  return nn_check("conv_fwd");
}

int shy_nn_conv_fwd_head(const void* x, const void* w, const float* bias, const void* res,
                         void* y_bf16, const void* head_w, const float* head_b, float* head_y,
                         int N, int H, int W, int Cin, int head_c, int relu, void* stream) {
  if (!x || !w || !head_w || !head_y) return nn_fail("conv_fwd_head: bad pointers");
  if (((uintptr_t)bias | (uintptr_t)head_b | (uintptr_t)head_w) & 15)
    return nn_fail("conv_fwd_head: bias, head_w and head_b must be 16-byte aligned");
  if (N < 1 || H < 1 || W < 1 || (Cin != 32 && Cin != 64) || head_c < 4 || head_c > 16 ||
      head_c % 4)
    return nn_fail("conv_fwd_head: need Cin 32 or 64 and head_c in {4, 8, 12, 16}");
  if ((long long)N * H * W > INT32_MAX / 2) return nn_fail("conv_fwd_head: too many pixels");
  ConvArgs a;
  a.x = (const __bf16*)x;
  a.w = (const __bf16*)w;
  a.bias = bias;
  a.res = (const __bf16*)res;
  a.y_bf = (__bf16*)y_bf16;
  a.y_f = nullptr;
  a.N = N; a.H = H; a.W = W; a.Cin = Cin; a.Cout = 32; a.K = 9 * Cin;
  a.relu = relu;
  a.x2 = nullptr;
  a.C1 = 0;
  a.gres = nullptr;
  a.mask = nullptr;
  a.head_w = (const __bf16*)head_w;
  a.head_b = head_b;
  a.head_y = head_y;
  a.head_c = head_c;
  return launch_tiles(a, (hipStream_t)stream, "conv_fwd_head");
}

int shy_nn_conv_fwd_upcat(const void* x_half, const void* skip, const void* w, const float* bias,
                          const void* res, void* y_bf16, float* y_f32, int N, int H, int W,
                          int C1, int C2, int Cout, int relu, void* stream) {
  if (!x_half || !skip || !w || (!y_bf16 && !y_f32) || (y_bf16 && y_f32))
    return nn_fail("conv_fwd_upcat: bad pointers");
  if ((uintptr_t)bias & 15) return nn_fail("conv_fwd_upcat: bias must be 16-byte aligned");
  if (N < 1 || H < 2 || W < 2 || H % 2 || W % 2 || C1 < 32 || C1 % 32 || C2 < 32 || C2 % 32 ||
      Cout < 32 || Cout % 32)
    return nn_fail("conv_fwd_upcat: need even H, W and C1, C2, Cout multiples of 32");
  if ((long long)N * H * W > INT32_MAX / 2) return nn_fail("conv_fwd_upcat: too many pixels");
  ConvArgs a;
  a.x = (const __bf16*)x_half;
  a.x2 = (const __bf16*)skip;
  a.C1 = C1;
  a.w = (const __bf16*)w;
  a.bias = bias;
  a.res = (const __bf16*)res;
  a.y_bf = (__bf16*)y_bf16;
  a.y_f = y_f32;
  a.N = N; a.H = H; a.W = W; a.Cin = C1 + C2; a.Cout = Cout; a.K = 9 * (C1 + C2);
  a.relu = relu;
  a.gres = nullptr;
  a.mask = nullptr;
  return launch_pf(a, 9, (hipStream_t)stream, "conv_fwd_upcat");
}


int shy_nn_conv_fwd_pool(const void* x, const void* w, const float* bias, const void* res,
                         void* y_bf16, void* pool_y, int N, int H, int W, int Cin, int Cout,
                         int relu, void* stream) {
  if (!x || !w || !y_bf16 || !pool_y) return nn_fail("conv_fwd_pool: bad pointers");
  if (((uintptr_t)bias & 15) || ((uintptr_t)pool_y & 7))
    return nn_fail("conv_fwd_pool: bias must be 16-byte and pool_y 8-byte aligned");
  if (N < 1 || H < 2 || W < 2 || H % 2 || W % 2 || Cin < 32 || Cin % 32 || Cout < 32 || Cout % 32)
    return nn_fail("conv_fwd_pool: need H, W even and Cin, Cout multiples of 32");
  if ((long long)N * H * W > INT32_MAX / 2) return nn_fail("conv_fwd_pool: too many pixels");
  const long long tiles16 = (long long)N * ((H + 15) / 16) * ((W + kTW - 1) / kTW);
  if (tiles16 < device_cus()) {
    // layers the tile loop does not take: the convolution, then the pool
    if (int r = shy_nn_conv_fwd(x, w, bias, res, y_bf16, nullptr, N, H, W, Cin, Cout, 9, relu,
                                stream))
      return r;
    return shy_nn_pool2(y_bf16, pool_y, N, H, W, Cout, 0, stream);
  }
  ConvArgs a;
  a.x = (const __bf16*)x;
  a.w = (const __bf16*)w;
  a.bias = bias;
  a.res = (const __bf16*)res;
  a.y_bf = (__bf16*)y_bf16;
  a.y_f = nullptr;
  a.N = N; a.H = H; a.W = W; a.Cin = Cin; a.Cout = Cout; a.K = 9 * Cin;
  a.relu = relu;
  a.x2 = nullptr;
  a.C1 = 0;
  a.gres = nullptr;
  a.mask = nullptr;
  a.pool_y = (__bf16*)pool_y;
  return launch_tiles(a, (hipStream_t)stream, "conv_fwd_pool");
}

int shy_nn_conv_dgrad(const void* g, const void* wf, const void* gres, const void* mask,
                      void* dx_bf16, int N, int H, int W, int Cin, int Cout, int taps,
                      void* stream) {
  if (!g || !wf || !dx_bf16) return nn_fail("conv_dgrad: bad pointers");
  if (N < 1 || H < 1 || W < 1 || Cin < 32 || Cin % 32 || Cout < 32 || Cout % 32 ||
      (taps != 9 && taps != 1))
    return nn_fail("conv_dgrad: need Cin and Cout multiples of 32 and taps 9 or 1");
  if ((long long)N * H * W > INT32_MAX / 2) return nn_fail("conv_dgrad: too many pixels");
  ConvArgs a;
  a.x = (const __bf16*)g;
  a.w = (const __bf16*)wf;
  a.bias = nullptr;
  a.res = nullptr;
  a.y_bf = (__bf16*)dx_bf16;
  a.y_f = nullptr;
  // the data gradient is the convolution of g (Cout channels) with the
  // flipped weights into Cin channels
  a.N = N; a.H = H; a.W = W; a.Cin = Cout; a.Cout = Cin; a.K = taps * Cout;
  a.relu = 0;
  a.x2 = nullptr;
  a.C1 = 0;
  a.gres = (const __bf16*)gres;
  a.mask = (const __bf16*)mask;
  return launch_pf(a, taps, (hipStream_t)stream, "conv_dgrad");
}

int shy_nn_flip_weights(const void* w, void* wf, int Cin, int Cout, int taps, void* stream) {
  if (!w || !wf || (taps != 9 && taps != 1)) return nn_fail("flip_weights: bad argument");
  const int n = Cin * Cout * taps;
  flip_weights_kernel<<<blocks(n), 256, 0, (hipStream_t)stream>>>((const __bf16*)w, (__bf16*)wf,
                                                                  Cin, Cout, taps);
  return nn_check("flip_weights");
}

int shy_nn_conv_wgrad(const void* x, const void* g, float* dw, float* part, int part_splits,
                      int N, int H, int W, int Cin, int Cout, int taps, void* stream) {
  if (!x || !g || !dw || !part || part_splits < 1) return nn_fail("conv_wgrad: bad pointers");
  if (Cin % 8 || Cout % 8 || (taps != 9 && taps != 1)) return nn_fail("conv_wgrad: channels must be multiples of 8");
  const long long P = (long long)N * H * W;
  if (taps == 9 && Cin % 32 == 0 && Cout % 32 == 0) {
    WgArgs b;
    b.x = (const __bf16*)x;
    b.g = (const __bf16*)g;
    b.part = part;
    b.N = N; b.H = H; b.W = W; b.Cin = Cin; b.Cout = Cout; b.K = 9 * Cin;
    b.tiles_x = (W + kWgTW - 1) / kWgTW;
    b.tiles_y = (H + kWgTH - 1) / kWgTH;
    const long long tiles = (long long)N * b.tiles_x * b.tiles_y;
    if (tiles > INT32_MAX) return nn_fail("conv_wgrad: too many tiles");
    b.tiles = (int)tiles;
    // 32 output channels per block, as in the forward (64-channel blocks
    // ran at two waves per SIMD: training step 4.538 -> 4.513 ms)
    const int mt = 2;
    const int bxy = (Cout / (16 * mt)) * (Cin / 32);
    // about 1024 blocks (4 waves per CU), at most part_splits partials, at
    // least one tile each: fewer partials to sum afterwards
    long long splits = std::max(1LL, std::min<long long>((1024 + bxy - 1) / bxy, tiles));
    splits = std::min<long long>(splits, part_splits);
    b.tiles_per_split = (int)((tiles + splits - 1) / splits);
    splits = (tiles + b.tiles_per_split - 1) / b.tiles_per_split;
    hipStream_t s = (hipStream_t)stream;
    dim3 grid((unsigned)(Cout / (16 * mt)), (unsigned)(Cin / 32), (unsigned)splits);
    wgrad_lds_kernel<2><<<grid, 256, 0, s>>>(b);
    if (int r = nn_check("conv_wgrad")) return r;
    const int n = Cout * b.K;
    wgrad_reduce_kernel<<<(n + 31) / 32, 256, 0, s>>>(part, dw, n, (int)splits);
    return nn_check("wgrad_reduce");
  }
  int chunk = (int)((P + part_splits - 1) / part_splits);

  chunk = (chunk + 31) / 32 * 32;
  const int splits = (int)((P + chunk - 1) / chunk);
  WgradArgs a;
  a.x = (const __bf16*)x;
  a.g = (const __bf16*)g;
  a.part = part;
  a.N = N; a.H = H; a.W = W; a.Cin = Cin; a.Cout = Cout; a.K = taps * Cin; a.taps = taps;
  a.chunk = chunk;
  hipStream_t s = (hipStream_t)stream;
  dim3 grid((unsigned)((Cout + 31) / 32), (unsigned)((a.K + 31) / 32), (unsigned)splits);
  conv_wgrad_kernel<<<grid, 64, 0, s>>>(a);
  if (int r = nn_check("conv_wgrad")) return r;
  const int n = Cout * a.K;
  wgrad_reduce_kernel<<<(n + 31) / 32, 256, 0, s>>>(part, dw, n, splits);
  return nn_check("wgrad_reduce");
}

int shy_nn_conv_wgrad_noalloc(const void* x, const void* g, float* dw, int N, int H, int W,
                              int Cin, int Cout, int taps, void* stream) {
  if (Cin < 1 || Cout < 1 || (taps != 9 && taps != 1))
    return nn_fail("conv_wgrad_noalloc: bad shape");
  const size_t per = (size_t)Cout * taps * Cin;
  if (per > (size_t)SHY_NN_NOALLOC_FLOATS)
    return nn_fail("conv_wgrad_noalloc: Cout*taps*Cin = " + std::to_string(per) +
                   " floats exceeds the scratch of " + std::to_string(SHY_NN_NOALLOC_FLOATS));
  float* part = nullptr;
  if (int r = noalloc_scratch(&part)) return r;
  const int splits = (int)std::min<size_t>((size_t)SHY_NN_NOALLOC_FLOATS / per,
                                           (size_t)SHY_NN_NOALLOC_MAX_SPLITS);
  return shy_nn_conv_wgrad(x, g, dw, part, splits, N, H, W, Cin, Cout, taps, stream);
}

int shy_nn_bias_grad(const void* g, float* db, float* part, int part_rows, int P, int C,
                     void* stream) {
  if (!g || !db || !part || C < 1 || C > 1024 || P < 1 || part_rows < 1)
    return nn_fail("bias_grad: bad argument");
  hipStream_t s = (hipStream_t)stream;
  if (C % 8 == 0 && ((uintptr_t)g & 15) == 0 && ((uintptr_t)part & 15) == 0) {
    // at least 2048 pixels per block (two 1024-thread blocks per CU on the
    // largest layers, eight rows in flight per thread)
    const int rows = std::max((P + part_rows - 1) / part_rows, 2048);
    const int nb = (P + rows - 1) / rows;
    bias_grad_v8_kernel<<<nb, 1024, 0, s>>>((const __bf16*)g, part, P, C, rows);
    if (int r = nn_check("bias_grad")) return r;
    bias_reduce_kernel<<<1, 1024, 0, s>>>(part, db, C, nb);
    return nn_check("bias_grad_reduce");
  }
  const int rows = (P + part_rows - 1) / part_rows;
  const int nb = (P + rows - 1) / rows;
  bias_grad_kernel<<<nb, 256, 0, s>>>((const __bf16*)g, part, P, C, rows);
  if (int r = nn_check("bias_grad")) return r;
  wgrad_reduce_kernel<<<(C + 31) / 32, 256, 0, s>>>(part, db, C, nb);
  return nn_check("bias_grad_reduce");
}

inline bool a16(const void* p) { return ((uintptr_t)p & 15) == 0; }

int shy_nn_relu_bwd(const void* y, const void* g, void* out, size_t n, void* stream) {
  if (n % 8 == 0 && a16(y) && a16(g) && a16(out))
    relu_bwd_v8_kernel<<<blocks(n / 8), 256, 0, (hipStream_t)stream>>>(
        (const __bf16*)y, (const __bf16*)g, (__bf16*)out, n / 8);
  else
    relu_bwd_kernel<<<blocks(n), 256, 0, (hipStream_t)stream>>>((const __bf16*)y,
                                                                 (const __bf16*)g, (__bf16*)out, n);
  return nn_check("relu_bwd");
}

int shy_nn_pool2(const void* x, void* y, int N, int H, int W, int C, int backward, void* stream) {
  if (H % 2 || W % 2) return nn_fail("pool2: H and W must be even");
  hipStream_t s = (hipStream_t)stream;
  const bool v8 = C % 8 == 0 && a16(x) && a16(y);
  if (v8 && !backward)
    pool2_fwd_v8_kernel<<<blocks((size_t)N * H * W * C / 32), 256, 0, s>>>(
        (const __bf16*)x, (__bf16*)y, N, H, W, C);
  else if (v8)
    pool2_bwd_v8_kernel<<<blocks((size_t)N * H * W * C / 8), 256, 0, s>>>(
        (const __bf16*)x, (__bf16*)y, N, H, W, C);
  else if (!backward)
    pool2_fwd_kernel<<<blocks((size_t)N * H * W * C / 4), 256, 0, s>>>((const __bf16*)x, (__bf16*)y,
                                                                       N, H, W, C);
  else
    pool2_bwd_kernel<<<blocks((size_t)N * H * W * C), 256, 0, s>>>((const __bf16*)x, (__bf16*)y, N,
                                                                   H, W, C);
  return nn_check("pool2");
}

int shy_nn_up2(const void* x, void* y, int N, int H, int W, int C, int Cy, int c_off, int backward,
               void* stream) {
  if (c_off < 0 || c_off + C > Cy) return nn_fail("up2: channel range outside the target");
  hipStream_t s = (hipStream_t)stream;
  const bool v8 = C % 8 == 0 && Cy % 8 == 0 && c_off % 8 == 0 && a16(x) && a16(y);
  if (v8 && !backward)
    up2_fwd_v8_kernel<<<blocks((size_t)N * 4 * H * W * C / 8), 256, 0, s>>>(
        (const __bf16*)x, (__bf16*)y, N, H, W, C, Cy, c_off);
  else if (v8)
    up2_bwd_v8_kernel<<<blocks((size_t)N * H * W * C / 8), 256, 0, s>>>(
        (const __bf16*)x, (__bf16*)y, N, H, W, C, Cy, c_off);
  else if (!backward)
    up2_fwd_kernel<<<blocks((size_t)N * 4 * H * W * C), 256, 0, s>>>(
        (const __bf16*)x, (__bf16*)y, N, H, W, C, Cy, c_off);
  else
    up2_bwd_kernel<<<blocks((size_t)N * H * W * C), 256, 0, s>>>((const __bf16*)x, (__bf16*)y, N, H,
                                                                 W, C, Cy, c_off);
  return nn_check("up2");
}

int shy_nn_channels(const void* src, void* dst, size_t P, int C, int Cy, int c_off, int dir,
                    void* stream) {
  if (c_off < 0 || c_off + C > Cy) return nn_fail("channels: range outside the target");
  if (C % 8 == 0 && Cy % 8 == 0 && c_off % 8 == 0 && a16(src) && a16(dst))
    channels_copy_v8_kernel<<<blocks(P * C / 8), 256, 0, (hipStream_t)stream>>>(
        (const __bf16*)src, (__bf16*)dst, P, C, Cy, c_off, dir);
  else
    channels_copy_kernel<<<blocks(P * C), 256, 0, (hipStream_t)stream>>>(
        (const __bf16*)src, (__bf16*)dst, P, C, Cy, c_off, dir);
  return nn_check("channels");
}

int shy_nn_add(const void* a, const void* b, void* y, size_t n, void* stream) {
  if (n % 8 == 0 && a16(a) && a16(b) && a16(y))
    add_v8_kernel<<<blocks(n / 8), 256, 0, (hipStream_t)stream>>>(
        (const __bf16*)a, (const __bf16*)b, (__bf16*)y, n / 8);
  else
    add_kernel<<<blocks(n), 256, 0, (hipStream_t)stream>>>((const __bf16*)a, (const __bf16*)b,
                                                            (__bf16*)y, n);
  return nn_check("add");
}

int shy_nn_mse(const float* y, const float* t, const float* mask, void* g_bf16, float* part,
               size_t n, float inv_count, void* stream) {
  mse_kernel<<<blocks(n), 256, 0, (hipStream_t)stream>>>(y, t, mask, (__bf16*)g_bf16, part, n,
                                                          inv_count);
  return nn_check("mse");
}

int shy_nn_adam(float* w, const float* g, float* m, float* v, void* w_bf16, size_t n, float lr,
                float b1, float b2, float eps, int step, float weight_decay, void* stream) {
  if (step < 1) return nn_fail("adam: step must be >= 1");
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  adam_kernel<<<blocks(n), 256, 0, (hipStream_t)stream>>>(w, g, m, v, (__bf16*)w_bf16, n, lr, b1,
                                                           b2, eps, c1, c2, weight_decay);
  return nn_check("adam");
}

int shy_nn_adam_conv(float* w, const float* g, float* m, float* v, void* w_bf16, void* wf_bf16,
                     int Cin, int Cout, int taps, float lr, float b1, float b2, float eps, int step,
                     float weight_decay, void* stream) {
  if (step < 1) return nn_fail("adam_conv: step must be >= 1");
  if (!w || !g || !m || !v || !w_bf16 || !wf_bf16 || Cin < 1 || Cout < 1 ||
      (taps != 9 && taps != 1))
    return nn_fail("adam_conv: bad argument");
  const size_t n = (size_t)Cout * taps * Cin;
  if (n > (size_t)INT32_MAX) return nn_fail("adam_conv: too many weights");
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  adam_kernel<<<blocks(n), 256, 0, (hipStream_t)stream>>>(w, g, m, v, (__bf16*)w_bf16, n, lr, b1,
                                                           b2, eps, c1, c2, weight_decay,
                                                           (__bf16*)wf_bf16, Cin, Cout, taps);
  return nn_check("adam_conv");
}

int shy_nn_cast(const void* src, void* dst, size_t n, int to_bf16, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (to_bf16)
    f32_to_bf16_kernel<<<blocks(n), 256, 0, s>>>((const float*)src, (__bf16*)dst, n);
  else
    bf16_to_f32_kernel<<<blocks(n), 256, 0, s>>>((const __bf16*)src, (float*)dst, n);
  return nn_check("cast");
}

}  // extern "C"
