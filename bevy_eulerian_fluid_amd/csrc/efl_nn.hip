// efl_nn.hip - the learned preconditioner: U-Net forward on MFMA, its
// backward (input and weight gradients), the loss and Adam, the dataset
// sample builder, and the CG hooks (features from the residual, z and the
// flexible-CG partials from the output).
//
// Tensors are NHWC BF16 (channels innermost, a multiple of 16), one plane per
// resolution level; weights are BF16 in [tap][cout][cin] order for the
// forward and [tap'][cin][cout] (taps flipped) for the input gradient, both
// packed from FP32 master weights [cout][cin][ky][kx].
//
// A 3x3 convolution is an implicit GEMM: M = pixels, N = output channels,
// K = 9 taps x input channels.  One wavefront owns 64 consecutive pixels of a
// row (four 16-pixel M blocks) and every output channel; per 32-wide K step
// (32 channels of one tap, or 16 channels of two taps when the input has 16)
// it loads one 8-channel fragment per lane per M block (16 bytes, zero
// outside the image) and one weight fragment per N block, and issues gfx950's
// v_mfma_f32_16x16x32_bf16 (FP32 accumulation).  Concatenated inputs
// (decoder: upsampled coarse tensor + skip tensor) are read in place: K
// chunks below CA come from A at (y/2, x/2), the rest from B.  The weight
// gradient (K = pixels) runs v_mfma_f32_16x16x16_bf16.
#include "efl_nn.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "efl_launch.h"
#include "efl_reduce.h"

namespace efl {
namespace nn {

typedef short bf16x4 __attribute__((__vector_size__(4 * sizeof(short))));
typedef float f32x4 __attribute__((__vector_size__(4 * sizeof(float))));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// gfx950's v_mfma_f32_16x16x32_bf16: D (16 x 16) += A (16 x 32) B (32 x 16);
// lane l supplies A[l % 16][8 (l / 16) .. + 7] and B[8 (l / 16) .. + 7][l % 16]
// (8 BF16 each, one 16-byte value) and holds D[4 (l / 16) + i][l % 16].
__device__ __forceinline__ f32x4 mfma32(uint4 a, uint4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0, 0,
                                                 0);
}

// ---------------------------------------------------------------- network

const LayerDef kNet[kLayers] = {
    // ca, cb, cout, level, upa, relu, in_a, in_b
    {16, 0, 16, 0, 0, 1, kTX, -1},      // 0  e0a
    {16, 0, 16, 0, 0, 1, 0, -1},        // 1  e0b -> skip 0
    {16, 0, 32, 1, 0, 1, kTP0, -1},     // 2  e1a (pool of 1)
    {32, 0, 32, 1, 0, 1, 2, -1},        // 3  e1b -> skip 1
    {32, 0, 64, 2, 0, 1, kTP0 + 1, -1}, // 4  e2a (pool of 3)
    {64, 0, 64, 2, 0, 1, 4, -1},        // 5  e2b -> skip 2
    {64, 0, 64, 3, 0, 1, kTP0 + 2, -1}, // 6  b3a (pool of 5)
    {64, 0, 64, 3, 0, 1, 6, -1},        // 7  b3b
    {64, 64, 64, 2, 1, 1, 7, 5},        // 8  d2a: up(7) | 5
    {64, 0, 32, 2, 0, 1, 8, -1},        // 9  d2b
    {32, 32, 32, 1, 1, 1, 9, 3},        // 10 d1a: up(9) | 3
    {32, 0, 16, 1, 0, 1, 10, -1},       // 11 d1b
    {16, 16, 16, 0, 1, 1, 11, 1},       // 12 d0a: up(11) | 1
    {16, 0, 16, 0, 0, 1, 12, -1},       // 13 d0b
    {16, 0, 16, 0, 0, 0, 13, -1},       // 14 out (channel 0)
};
// pooled tensors: P_l = avgpool2(Y_src[l])
static const int kPoolSrc[3] = {1, 3, 5};

size_t layer_weight_offset(int k) {
  size_t off = 0;
  for (int j = 0; j < k; ++j) off += (size_t)kNet[j].cout * (kNet[j].ca + kNet[j].cb) * 9 + kNet[j].cout;
  return off;
}
size_t layer_bias_offset(int k) { return layer_weight_offset(k) + (size_t)kNet[k].cout * (kNet[k].ca + kNet[k].cb) * 9; }
size_t param_count() { return layer_weight_offset(kLayers); }
double flops_per_pixel() {
  double f = 0;
  for (int k = 0; k < kLayers; ++k)
    f += 2.0 * 9.0 * (kNet[k].ca + kNet[k].cb) * kNet[k].cout / double(1 << (2 * kNet[k].level));
  return f;
}

int Shapes::level(int t) const {
  if (t < kLayers) return kNet[t].level;
  if (t == kTX) return 0;
  return t - kTP0 + 1;
}
int Shapes::channels(int t) const {
  if (t < kLayers) return kNet[t].cout;
  if (t == kTX) return kInCh;
  return kNet[kPoolSrc[t - kTP0]].cout;
}
size_t Shapes::elems(int t) const {
  const int l = level(t);
  return (size_t)n * (w >> l) * (h >> l) * channels(t);
}

// ---------------------------------------------------------------- helpers

__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ bool stopped(const double* stop) { return stop && *stop != 0.0; }

// ---------------------------------------------------------------- convolution

struct ConvArgs {
  const uint16_t* a;  // CA channels, (W/2 x H/2) when UPA
  const uint16_t* b;  // CB channels, W x H
  const uint16_t* w;  // [9][COUT][CA + CB]
  const float* bias;  // [COUT] or null
  uint16_t* y;        // COUT channels, W x H
  int W, H;
  long sa, sb, sy;    // per-sample strides (elements)
  const double* stop; // CG done flag: skip the launch when set
  int strip;          // host side: pixels per wavefront strip (64 or 16), 0 = by size
};

// One wavefront: MBW x 16 consecutive pixels of row y and every output
// channel (each input fragment feeds COUT / 16 MFMAs); MBW = 1 on the small
// levels so that they still launch enough workgroups.
template <int CA, int CB, int COUT, bool UPA, bool RELU, int MBW>
__global__ __launch_bounds__(256) void k_conv3x3(ConvArgs g) {
  constexpr int CIN = CA + CB, NB = COUT / 16;
  if (stopped(g.stop)) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lq = lane >> 4;
  const int n = blockIdx.z;
  const int y = blockIdx.y * 4 + wave;
  const int x0 = blockIdx.x * (16 * MBW);
  if (y >= g.H) return;  // no barriers below: a whole wavefront leaves
  const uint16_t* A = g.a + n * g.sa;
  const uint16_t* B = CB ? g.b + n * g.sb : nullptr;
  const int Wa = UPA ? g.W >> 1 : g.W;
  f32x4 acc[MBW][NB];
#pragma unroll
  for (int mb = 0; mb < MBW; ++mb)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  // K = 32 per MFMA over (tap, channel): with a multiple of 32 channels a
  // lane supplies channels c + 8 lq of one tap; with 16, channels 8 (lq & 1)
  // of tap 2 p + (lq >> 1), the ninth tap beside zeros.  k_conv3x3_lds runs
  // the same sequence.
  constexpr int KSTEPS = CIN % 32 == 0 ? 9 * (CIN / 32) : 5;
  static_assert(CIN % 32 == 0 || CIN == 16, "16 channels or a multiple of 32");
#pragma unroll 2
  for (int ks = 0; ks < KSTEPS; ++ks) {
    const int tap = CIN % 32 == 0 ? ks / (CIN / 32) : 2 * ks + (lq >> 1);
    const int c0 = CIN % 32 == 0 ? (ks % (CIN / 32)) * 32 + 8 * lq : 8 * (lq & 1);
    const bool has = tap < 9;
    const int tt = has ? tap : 8;
    const int dy = tt / 3 - 1, dx = tt % 3 - 1;
    const int yy = y + dy;
    const bool rowok = has && yy >= 0 && yy < g.H;
    const int ya = UPA ? (yy >> 1) : yy;
    uint4 bw[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
      bw[nb] = has ? *(const uint4*)(g.w + ((size_t)(tt * COUT + nb * 16 + lr) * CIN + c0)) : uint4{0u, 0u, 0u, 0u};
#pragma unroll
    for (int mb = 0; mb < MBW; ++mb) {
      const int xx = x0 + mb * 16 + lr + dx;
      uint4 av = {0u, 0u, 0u, 0u};
      if (rowok && xx >= 0 && xx < g.W) {
        if (c0 < CA) {
          const int xa = UPA ? (xx >> 1) : xx;
          av = *(const uint4*)(A + ((size_t)ya * Wa + xa) * CA + c0);
        } else {
          av = *(const uint4*)(B + ((size_t)yy * g.W + xx) * CB + (c0 - CA));
        }
      }
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = mfma32(av, bw[nb], acc[mb][nb]);
    }
  }
  // lane holds D[pixel 4*lq + i][channel lr] of every (M, N) block
  uint16_t* Y = g.y + n * g.sy + (size_t)y * g.W * COUT;
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    const int co = nb * 16 + lr;
    const float bv = g.bias ? g.bias[co] : 0.0f;
#pragma unroll
    for (int mb = 0; mb < MBW; ++mb)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int x = x0 + mb * 16 + 4 * lq + i;
        if (x >= g.W) continue;
        float v = acc[mb][nb][i] + bv;
        if (RELU) v = fmaxf(v, 0.0f);
        Y[(size_t)x * COUT + co] = f2bf(v);
      }
  }
}

// The same convolution with the block's input staged in LDS once: rows
// y0 - 1 .. y0 + 4 and pixels x0 - 1 .. x0 + 16 MBW of all CIN channels (A
// upsampled while staging, zero outside the image), so each input element
// comes from L2 about once instead of once per tap.  Every wavefront then
// runs k_conv3x3's MFMA sequence with the same operands in the same order
// (fragments from LDS), so the results are the same bits.  A pixel takes
// CIN + 8 elements of LDS: 16-byte stores and fewer bank conflicts on the
// 8-byte fragment reads.
template <int CIN, int MBW>
constexpr size_t conv_lds_bytes() { return (size_t)6 * (16 * MBW + 2) * (CIN + 8) * 2; }

template <int CA, int CB, int COUT, bool UPA, bool RELU, int MBW>
__global__ __launch_bounds__(256) void k_conv3x3_lds(ConvArgs g) {
  constexpr int CIN = CA + CB, NB = COUT / 16, TW = 16 * MBW, RW = TW + 2, PS = CIN + 8;
  constexpr int CH = CIN / 8, NCH = 6 * RW * CH, U = (NCH + 255) / 256 < 8 ? (NCH + 255) / 256 : 8;
  static_assert(CA % 8 == 0 && CB % 8 == 0, "16-byte channel chunks");
  extern __shared__ uint4 conv_smem[];
  uint16_t* sm = (uint16_t*)conv_smem;
  if (stopped(g.stop)) return;  // uniform for the block
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lq = lane >> 4;
  const int n = blockIdx.z;
  const int y0 = blockIdx.y * 4, x0 = blockIdx.x * TW;
  const uint16_t* A = g.a + n * g.sa;
  const uint16_t* B = CB ? g.b + n * g.sb : nullptr;
  const int Wa = UPA ? g.W >> 1 : g.W;
  // staging, U chunks of 16 bytes per thread in flight at a time (a chunk
  // outside the image reads A's first chunk and stores zeros)
  for (int u0 = threadIdx.x; u0 < NCH; u0 += U * 256) {
    uint4 v[U];
    int at[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const int u = u0 + k * 256;
      const int c8 = u % CH, p = u / CH, px = p % RW, py = p / RW;
      const int xx = x0 - 1 + px, yy = y0 - 1 + py, c = c8 * 8;
      const bool in = u < NCH && xx >= 0 && xx < g.W && yy >= 0 && yy < g.H;
      const uint16_t* src = A;
      if (in) {
        if (c < CA) {
          const int xa = UPA ? xx >> 1 : xx, ya = UPA ? yy >> 1 : yy;
          src = A + ((size_t)ya * Wa + xa) * CA + c;
        } else {
          src = B + ((size_t)yy * g.W + xx) * CB + (c - CA);
        }
      }
      v[k] = *(const uint4*)src;
      if (!in) v[k] = uint4{0u, 0u, 0u, 0u};
      at[k] = u < NCH ? p * PS + c : -1;
    }
#pragma unroll
    for (int k = 0; k < U; ++k)
      if (at[k] >= 0) *(uint4*)(sm + at[k]) = v[k];
  }
  __syncthreads();
  const int y = y0 + wave;
  if (y >= g.H) return;  // no barriers below: a whole wavefront leaves
  f32x4 acc[MBW][NB];
#pragma unroll
  for (int mb = 0; mb < MBW; ++mb)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  // the K = 32 sequence of k_conv3x3 (the same operands in the same order),
  // fragments from LDS; unrolled, so the weights of the next steps load
  // while these multiply
  constexpr int KSTEPS = CIN % 32 == 0 ? 9 * (CIN / 32) : 5;
  static_assert(CIN % 32 == 0 || CIN == 16, "16 channels or a multiple of 32");
#pragma unroll
  for (int ks = 0; ks < KSTEPS; ++ks) {
    const int tap = CIN % 32 == 0 ? ks / (CIN / 32) : 2 * ks + (lq >> 1);
    const int c0 = CIN % 32 == 0 ? (ks % (CIN / 32)) * 32 + 8 * lq : 8 * (lq & 1);
    const bool has = tap < 9;
    const int tt = has ? tap : 8;
    const int dy = tt / 3 - 1, dx = tt % 3 - 1;
    const uint16_t* row = sm + (size_t)(wave + 1 + dy) * RW * PS;
    uint4 bw[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
      bw[nb] = has ? *(const uint4*)(g.w + ((size_t)(tt * COUT + nb * 16 + lr) * CIN + c0)) : uint4{0u, 0u, 0u, 0u};
#pragma unroll
    for (int mb = 0; mb < MBW; ++mb) {
      uint4 av = *(const uint4*)(row + (mb * 16 + lr + dx + 1) * PS + c0);
      if (!has) av = uint4{0u, 0u, 0u, 0u};
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = mfma32(av, bw[nb], acc[mb][nb]);
    }
  }
  // lane holds D[pixel 4*lq + i][channel lr] of every (M, N) block
  uint16_t* Y = g.y + n * g.sy + (size_t)y * g.W * COUT;
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    const int co = nb * 16 + lr;
    const float bv = g.bias ? g.bias[co] : 0.0f;
#pragma unroll
    for (int mb = 0; mb < MBW; ++mb)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int x = x0 + mb * 16 + 4 * lq + i;
        if (x >= g.W) continue;
        float v = acc[mb][nb][i] + bv;
        if (RELU) v = fmaxf(v, 0.0f);
        Y[(size_t)x * COUT + co] = f2bf(v);
      }
  }
}

// ---------------------------------------------------------------- pooling

// out (W x H, C) = mean of the 2 x 2 children of in (2W x 2H, C)
template <int C>
__global__ __launch_bounds__(256) void k_pool2(const uint16_t* in, uint16_t* out, int W, int H, long sin, long sout,
                                               const double* stop) {
  if (stopped(stop)) return;
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  const long total = (long)W * H * (C / 4);
  if (i >= total) return;
  const int cq = (int)(i % (C / 4));
  const long p = i / (C / 4);
  const int x = (int)(p % W), y = (int)(p / W);
  const uint16_t* src = in + blockIdx.y * sin;
  float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int d = 0; d < 4; ++d) {
    const bf16x4 v = *(const bf16x4*)(src + ((size_t)(2 * y + (d >> 1)) * (2 * W) + 2 * x + (d & 1)) * C + cq * 4);
#pragma unroll
    for (int j = 0; j < 4; ++j) s[j] += bf2f((uint16_t)v[j]);
  }
  bf16x4 o;
#pragma unroll
  for (int j = 0; j < 4; ++j) o[j] = (short)f2bf(0.25f * s[j]);
  *(bf16x4*)(out + blockIdx.y * sout + p * C + cq * 4) = o;
}

// g_in (2W x 2H) += g_out / 4 at each child
template <int C>
__global__ __launch_bounds__(256) void k_pool2_bwd(const uint16_t* gout, uint16_t* gin, int W, int H, long sout, long sin) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  const long total = (long)W * H * (C / 4);
  if (i >= total) return;
  const int cq = (int)(i % (C / 4));
  const long p = i / (C / 4);
  const int x = (int)(p % W), y = (int)(p / W);
  const bf16x4 go = *(const bf16x4*)(gout + blockIdx.y * sout + p * C + cq * 4);
  uint16_t* dst = gin + blockIdx.y * sin;
#pragma unroll
  for (int d = 0; d < 4; ++d) {
    bf16x4* q = (bf16x4*)(dst + ((size_t)(2 * y + (d >> 1)) * (2 * W) + 2 * x + (d & 1)) * C + cq * 4);
    bf16x4 v = *q;
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = (short)f2bf(bf2f((uint16_t)v[j]) + 0.25f * bf2f((uint16_t)go[j]));
    *q = v;
  }
}

// Gradient of cat(up(A), B) (2W x 2H, CA + CB channels) split into A
// (W x H: sum over the 4 children) and B (2W x 2H: copied).
template <int CA, int CB>
__global__ __launch_bounds__(256) void k_cat_bwd(const uint16_t* gcat, uint16_t* ga, uint16_t* gb, int W, int H,
                                                 long scat, long sa, long sb) {
  constexpr int G = (CA + CB) / 4;
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  const long total = (long)W * H * G;
  if (i >= total) return;
  const int cq = (int)(i % G);
  const long p = i / G;
  const int x = (int)(p % W), y = (int)(p / W);
  const uint16_t* src = gcat + blockIdx.y * scat;
  if (cq < CA / 4) {
    float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      const bf16x4 v =
          *(const bf16x4*)(src + ((size_t)(2 * y + (d >> 1)) * (2 * W) + 2 * x + (d & 1)) * (CA + CB) + cq * 4);
#pragma unroll
      for (int j = 0; j < 4; ++j) s[j] += bf2f((uint16_t)v[j]);
    }
    bf16x4 o;
#pragma unroll
    for (int j = 0; j < 4; ++j) o[j] = (short)f2bf(s[j]);
    *(bf16x4*)(ga + blockIdx.y * sa + p * CA + cq * 4) = o;
  } else {
    const int cb = cq - CA / 4;
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      const size_t q = (size_t)(2 * y + (d >> 1)) * (2 * W) + 2 * x + (d & 1);
      *(bf16x4*)(gb + blockIdx.y * sb + q * CB + cb * 4) = *(const bf16x4*)(src + q * (CA + CB) + CA + cb * 4);
    }
  }
}

// g = y > 0 ? g : 0 (ReLU), 4 elements per thread
__global__ __launch_bounds__(256) void k_relu_bwd(uint16_t* g, const uint16_t* y, long n4) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n4) return;
  bf16x4 gv = ((bf16x4*)g)[i];
  const bf16x4 yv = ((const bf16x4*)y)[i];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint16_t h = (uint16_t)yv[j];
    if ((h & 0x8000u) || (h & 0x7fffu) == 0) gv[j] = 0;
  }
  ((bf16x4*)g)[i] = gv;
}

// ---------------------------------------------------------------- weight gradient

struct WgradArgs {
  const uint16_t* a;
  const uint16_t* b;
  const uint16_t* dy;  // COUT channels, W x H
  float* part;         // [(n * R + region) * 9 + tap][COUT][CIN]
  float* bpart;        // [n * R + region][COUT]
  int W, H, rx;        // regions of 64 x 64 pixels, rx per row
  long sa, sb, sdy;
};

// dW[tap][co][ci] = sum_p dY[p][co] * In[p + off(tap)][ci]: M = co (the block's
// 16), N = ci, K = pixels.  Grid: x = regions, y = 9 taps x (COUT / 16), z = n.
// Each wavefront takes 16 rows of the region; the 4 wavefront sums are added
// in LDS and written as this block's partial (summed later in a fixed order).
template <int CA, int CB, int COUT, bool UPA>
__global__ __launch_bounds__(256) void k_wgrad(WgradArgs g) {
  constexpr int CIN = CA + CB, NB = CIN / 16, MB = COUT / 16;
  __shared__ float red[4 * 16 * CIN];
  __shared__ float bred[4 * 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lq = lane >> 4;
  const int region = blockIdx.x, tap = blockIdx.y / MB, mb = blockIdx.y % MB, n = blockIdx.z;
  const int dy = tap / 3 - 1, dx = tap % 3 - 1;
  const int x0 = (region % g.rx) * 64, y0 = (region / g.rx) * 64 + wave * 16;
  const uint16_t* A = g.a + n * g.sa;
  const uint16_t* B = CB ? g.b + n * g.sb : nullptr;
  const uint16_t* DY = g.dy + n * g.sdy;
  const int Wa = UPA ? g.W >> 1 : g.W;
  f32x4 acc[NB];
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) acc[nb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  float bsum = 0.0f;
  for (int r = 0; r < 16; ++r) {
    const int y = y0 + r;
    if (y >= g.H) break;
    const int yy = y + dy;
    const bool rowok = yy >= 0 && yy < g.H;
    for (int kc = 0; kc < 4; ++kc) {
      const int xb = x0 + kc * 16 + 4 * lq;
      if (x0 + kc * 16 >= g.W) break;
      bf16x4 av;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int x = xb + i;
        av[i] = x < g.W ? (short)DY[((size_t)y * g.W + x) * COUT + mb * 16 + lr] : (short)0;
      }
      if (tap == 4) {
#pragma unroll
        for (int i = 0; i < 4; ++i) bsum += bf2f((uint16_t)av[i]);
      }
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
        const int ci = nb * 16 + lr;
        bf16x4 bv;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int xx = xb + i + dx;
          short v = 0;
          if (rowok && xx >= 0 && xx < g.W && xb + i < g.W) {
            if (nb * 16 < CA) {
              const int xa = UPA ? (xx >> 1) : xx, ya = UPA ? (yy >> 1) : yy;
              v = (short)A[((size_t)ya * Wa + xa) * CA + ci];
            } else {
              v = (short)B[((size_t)yy * g.W + xx) * CB + ci - CA];
            }
          }
          bv[i] = v;
        }
        acc[nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[nb], 0, 0, 0);
      }
    }
  }
  // lane holds D[co 4*lq + i][ci lr] of each N block
#pragma unroll
  for (int nb = 0; nb < NB; ++nb)
#pragma unroll
    for (int i = 0; i < 4; ++i) red[wave * 16 * CIN + (4 * lq + i) * CIN + nb * 16 + lr] = acc[nb][i];
  bred[wave * 64 + lane] = bsum;
  __syncthreads();
  float* out = g.part + (((size_t)n * gridDim.x + region) * 9 + tap) * COUT * CIN + (size_t)mb * 16 * CIN;
  for (int e = threadIdx.x; e < 16 * CIN; e += 256)
    out[e] = red[e] + red[16 * CIN + e] + red[2 * 16 * CIN + e] + red[3 * 16 * CIN + e];
  if (tap == 4 && threadIdx.x < 16) {
    float s = 0.0f;
    for (int w = 0; w < 4; ++w)
      for (int q = 0; q < 4; ++q) s += bred[w * 64 + q * 16 + threadIdx.x];
    g.bpart[((size_t)n * gridDim.x + region) * COUT + mb * 16 + threadIdx.x] = s;
  }
}

// grad[co][ci][tap] = sum over s < S of part[s][tap][co][ci]; bias likewise
__global__ __launch_bounds__(256) void k_wreduce(const float* part, const float* bpart, int S, int cout, int cin,
                                                 float* gw, float* gb) {
  const long j = (long)blockIdx.x * 256 + threadIdx.x;
  const long M = 9L * cout * cin;
  if (j < M) {
    float s = 0.0f;
    for (int k = 0; k < S; ++k) s += part[(size_t)k * M + j];
    const int tap = (int)(j / ((long)cout * cin));
    const int co = (int)((j / cin) % cout), ci = (int)(j % cin);
    gw[((size_t)co * cin + ci) * 9 + tap] = s;
  } else if (j < M + cout) {
    const int co = (int)(j - M);
    float s = 0.0f;
    for (int k = 0; k < S; ++k) s += bpart[(size_t)k * cout + co];
    gb[co] = s;
  }
}

// ---------------------------------------------------------------- parameters

// master [cout][cin][9] -> wf [9][cout][cin], wb [8 - tap][cin][cout]
__global__ __launch_bounds__(256) void k_pack(const float* w, uint16_t* wf, uint16_t* wb, int cout, int cin) {
  const long j = (long)blockIdx.x * 256 + threadIdx.x;
  if (j >= 9L * cout * cin) return;
  const int tap = (int)(j % 9), ci = (int)((j / 9) % cin), co = (int)(j / (9L * cin));
  const uint16_t h = f2bf(w[j]);
  wf[((size_t)tap * cout + co) * cin + ci] = h;
  wb[((size_t)(8 - tap) * cin + ci) * cout + co] = h;
}

__global__ __launch_bounds__(256) void k_adam(float* p, const float* g, float* m, float* v, long n, float lr,
                                              float c1, float c2) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---------------------------------------------------------------- data

// A dataset sample (kSamplePlanes planes): r / sigma, solid fraction,
// clamped level set, mask (pressure unknown), target / rms(target), then the
// stencil of the pressure operator at that cell.  Flexible CG fixes the
// length of every step itself, so only the shape of A^-1 r matters and each
// target is scaled to unit rms.  The stencil is residual.wgsl's operator
// without its positive factor dt / (rho dx^2): (A y)_c = diag_c y_c -
// sum_k off_k y_k over the neighbours -x, +x, -y, +y inside the grid, with
// off_k = f_k for a fluid neighbour and f_k (1 - phi_k / phi_c) added to the
// diagonal for an air one (f_k the area fraction of the shared face); rows
// of cells that are not unknowns are zero.
__global__ __launch_bounds__(256) void k_make_sample(float* dst, const DevInst* I, int W, int H, const float* r,
                                                     const float* t, float inv_sigma, float inv_trms) {
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  const long n = (long)W * H;
  if (p >= n) return;
  const float4 a = I->afs[p];
  // the pressure unknowns of the current fields (k_mg_setup's MG_LSNEG and
  // not MG_CLOSED)
  const float phi = I->ls0[p];
  const bool act = phi < 0.0f && !(a.x == 0.0f && a.y == 0.0f && a.z == 0.0f && a.w == 0.0f);
  dst[p] = r[p] * inv_sigma;
  dst[n + p] = 1.0f - 0.25f * (a.x + a.y + a.z + a.w);
  dst[2 * n + p] = fminf(fmaxf(phi * kLsScale, -1.0f), 1.0f);
  dst[3 * n + p] = act ? 1.0f : 0.0f;
  dst[4 * n + p] = t[p] * inv_trms;
  const int x = (int)(p % W), y = (int)(p / W);
  const float f[4] = {a.x, a.y, a.z, a.w};
  const int ox[4] = {-1, 1, 0, 0}, oy[4] = {0, 0, -1, 1};
  float diag = 0.0f, off[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  if (act) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int jx = x + ox[k], jy = y + oy[k];
      if (jx < 0 || jy < 0 || jx >= W || jy >= H) continue;
      const float pn = I->ls0[(long)jy * W + jx];
      if (pn < 0.0f) {
        off[k] = f[k];
        diag += f[k];
      } else {
        diag += f[k] * (1.0f - pn / phi);
      }
    }
  }
  dst[5 * n + p] = diag;
#pragma unroll
  for (int k = 0; k < 4; ++k) dst[(6 + k) * n + p] = off[k];
}

// batch input tensor, target and mask from dataset samples idx[n]
__global__ __launch_bounds__(256) void k_gather(const float* data, const int* idx, uint16_t* x, float* target,
                                                float* mask, int W, int H) {
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  const long npx = (long)W * H;
  if (p >= npx) return;
  const int n = blockIdx.y;
  const float* s = data + (size_t)idx[n] * kSamplePlanes * npx;
  uint16_t* o = x + ((size_t)n * npx + p) * kInCh;
  bf16x4 v0 = {(short)f2bf(s[p]), (short)f2bf(s[npx + p]), (short)f2bf(s[2 * npx + p]), 0};
  *(bf16x4*)o = v0;
  const bf16x4 zero = {0, 0, 0, 0};
#pragma unroll
  for (int q = 1; q < kInCh / 4; ++q) *(bf16x4*)(o + 4 * q) = zero;
  target[(size_t)n * npx + p] = s[4 * npx + p];
  mask[(size_t)n * npx + p] = s[3 * npx + p];
}

// loss = sum mask (o - t)^2 / count; d o = 2 mask (o - t) / count (channel 0)
__global__ __launch_bounds__(256) void k_loss(const uint16_t* o, const float* target, const float* mask, uint16_t* g,
                                              long npx, float inv_count, double* loss) {
  __shared__ double sh[4];
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  double l = 0.0;
  if (p < npx) {
    const float d = mask[p] != 0.0f ? bf2f(o[p * kInCh]) - target[p] : 0.0f;
    l = (double)d * d;
    if (g) {
      const bf16x4 zero = {0, 0, 0, 0};
      bf16x4 v0 = {(short)f2bf(2.0f * d * inv_count), 0, 0, 0};
      *(bf16x4*)(g + p * kInCh) = v0;
#pragma unroll
      for (int q = 1; q < kInCh / 4; ++q) *(bf16x4*)(g + p * kInCh + 4 * q) = zero;
    }
  }
  for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(loss, (sh[0] + sh[1] + sh[2] + sh[3]) * (double)inv_count);
}

// Residual-norm loss, first pass: A y of each batch sample (y = output
// channel 0; the stencil's zero rows and columns leave cells that are not
// unknowns out) into ay, and the sums r.r, r.Ay, Ay.Ay (r = plane 0 on the
// unknowns) into sums[3 n ..].  Grid (pixel blocks, samples).
__global__ __launch_bounds__(256) void k_rloss_fwd(const float* data, const int* idx, const uint16_t* o, float* ay,
                                                   double* sums, int W, int H) {
  __shared__ double sh[3][4];
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  const long npx = (long)W * H;
  const int n = blockIdx.y;
  const float* s = data + (size_t)idx[n] * kSamplePlanes * npx;
  const uint16_t* y = o + (size_t)n * npx * kInCh;
  double rr = 0.0, ra = 0.0, aa = 0.0;
  if (p < npx) {
    const int x = (int)(p % W), yy = (int)(p / W);
    float v = s[5 * npx + p] * bf2f(y[p * kInCh]);
    if (x > 0) v -= s[6 * npx + p] * bf2f(y[(p - 1) * kInCh]);
    if (x + 1 < W) v -= s[7 * npx + p] * bf2f(y[(p + 1) * kInCh]);
    if (yy > 0) v -= s[8 * npx + p] * bf2f(y[(p - W) * kInCh]);
    if (yy + 1 < H) v -= s[9 * npx + p] * bf2f(y[(p + W) * kInCh]);
    const float r = s[3 * npx + p] != 0.0f ? s[p] : 0.0f;
    ay[(size_t)n * npx + p] = v;
    rr = (double)r * r;
    ra = (double)r * v;
    aa = (double)v * v;
  }
  for (int off = 32; off > 0; off >>= 1) {
    rr += __shfl_xor(rr, off, 64);
    ra += __shfl_xor(ra, off, 64);
    aa += __shfl_xor(aa, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = rr;
    sh[1][threadIdx.x >> 6] = ra;
    sh[2][threadIdx.x >> 6] = aa;
  }
  __syncthreads();
  if (threadIdx.x < 3) atomicAdd(sums + 3 * n + threadIdx.x, sh[threadIdx.x][0] + sh[threadIdx.x][1] + sh[threadIdx.x][2] + sh[threadIdx.x][3]);
}

// Second pass: loss += (1 - ra^2 / (rr aa)) / N per sample and, with g,
// d loss / d y = A v (A is symmetric on the unknowns) with v = d loss / d Ay
// = 2 ra / (rr aa) (alpha Ay - r) / N, alpha = ra / aa, in output channel 0
// of g (other channels zero).  A sample whose A y is zero adds 1 and no
// gradient.
__global__ __launch_bounds__(256) void k_rloss_bwd(const float* data, const int* idx, const float* ay,
                                                   const double* sums, uint16_t* g, double* loss, int W, int H,
                                                   float inv_n) {
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  const long npx = (long)W * H;
  const int n = blockIdx.y;
  const double rr = sums[3 * n], ra = sums[3 * n + 1], aa = sums[3 * n + 2];
  const bool ok = rr > 0.0 && aa > 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(loss, (ok ? 1.0 - ra * ra / (rr * aa) : 1.0) * (double)inv_n);
  if (!g || p >= npx) return;
  const float* s = data + (size_t)idx[n] * kSamplePlanes * npx;
  const float* a = ay + (size_t)n * npx;
  const float alpha = ok ? (float)(ra / aa) : 0.0f;
  const float coef = ok ? (float)(2.0 * ra / (rr * aa)) * inv_n : 0.0f;
  auto v = [&](long q) { return s[3 * npx + q] != 0.0f ? coef * (alpha * a[q] - s[q]) : 0.0f; };
  const int x = (int)(p % W), yy = (int)(p / W);
  float d = 0.0f;
  if (s[3 * npx + p] != 0.0f) {
    d = s[5 * npx + p] * v(p);
    if (x > 0) d -= s[6 * npx + p] * v(p - 1);
    if (x + 1 < W) d -= s[7 * npx + p] * v(p + 1);
    if (yy > 0) d -= s[8 * npx + p] * v(p - W);
    if (yy + 1 < H) d -= s[9 * npx + p] * v(p + W);
  }
  uint16_t* o = g + ((size_t)n * npx + p) * kInCh;
  const bf16x4 zero = {0, 0, 0, 0};
  bf16x4 v0 = {(short)f2bf(d), 0, 0, 0};
  *(bf16x4*)o = v0;
#pragma unroll
  for (int q = 1; q < kInCh / 4; ++q) *(bf16x4*)(o + 4 * q) = zero;
}

// network output channel 0 -> fp32 plane
__global__ __launch_bounds__(256) void k_read_out(const uint16_t* o, float* out, long npx) {
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if (p < npx) out[p] = bf2f(o[p * kInCh]);
}

// ---------------------------------------------------------------- CG hooks

__device__ __forceinline__ bool active_cell(const DevInst* I, size_t id) {
  const int fl = I->lev[0].fl[id];
  return (fl & MG_LSNEG) && !(fl & MG_CLOSED);
}

// input tensor (Wp x Hp, zero outside W x H) from the CG residual:
// sigma = sqrt(||r||^2 / (W H)) from the stopping test of this iteration
__global__ __launch_bounds__(256) void k_features(const DevInst* I, uint16_t* x, int Wp, int Hp, const double* stop) {
  if (stopped(stop)) return;
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if (p >= (long)Wp * Hp) return;
  const int xp = (int)(p % Wp), yp = (int)(p / Wp);
  bf16x4 v0 = {0, 0, 0, 0};
  if (xp < I->W && yp < I->H) {
    const size_t id = (size_t)yp * I->W + xp;
    const double rr = I->scal[SC_RR];
    const float inv_sigma = rr > 0.0 ? (float)(1.0 / sqrt(rr / ((double)I->W * I->H))) : 1.0f;
    const float4 a = I->afs[id];
    v0[0] = (short)f2bf(I->pcg_r[id] * inv_sigma);
    v0[1] = (short)f2bf(1.0f - 0.25f * (a.x + a.y + a.z + a.w));
    v0[2] = (short)f2bf(fminf(fmaxf(I->ls0[id] * kLsScale, -1.0f), 1.0f));
  }
  uint16_t* o = x + p * kInCh;
  *(bf16x4*)o = v0;
  const bf16x4 zero = {0, 0, 0, 0};
#pragma unroll
  for (int q = 1; q < kInCh / 4; ++q) *(bf16x4*)(o + 4 * q) = zero;
}

// z = sigma / factor * out (unknowns only), partials of r . z (RED_RZ) and
// z . q_prev (RED_ZQ, for the flexible beta) in blocks of 1024 cells
__global__ __launch_bounds__(256) void k_output(const DevInst* I, const uint16_t* o, int Wp, const double* stop) {
  __shared__ double sh[2][4];
  if (stopped(stop)) return;
  const double rr = I->scal[SC_RR];
  const float sigma = rr > 0.0 ? (float)sqrt(rr / ((double)I->W * I->H)) : 1.0f;
  const float factor = I->dt / I->rho / I->dx / I->dx;
  const float scale = sigma / factor;
  const long n = (long)I->W * I->H;
  double rz = 0.0, zq = 0.0;
  for (int k = 0; k < 4; ++k) {
    const long id = (long)blockIdx.x * 1024 + k * 256 + threadIdx.x;
    if (id >= n) break;
    const int x = (int)(id % I->W), y = (int)(id / I->W);
    float z = 0.0f;
    if (active_cell(I, id)) z = scale * bf2f(o[((size_t)y * Wp + x) * kInCh]);
    I->pcg_z[id] = z;
    rz += (double)I->pcg_r[id] * z;
    zq += (double)z * I->pcg_q[id];
  }
  for (int off = 32; off > 0; off >>= 1) {
    rz += __shfl_xor(rz, off, 64);
    zq += __shfl_xor(zq, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = rz;
    sh[1][threadIdx.x >> 6] = zq;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    I->red[(size_t)RED_RZ * I->red_stride + blockIdx.x] = sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3];
    I->red[(size_t)RED_ZQ * I->red_stride + blockIdx.x] = sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3];
    if (blockIdx.x == 0) I->scal[SC_NRZ] = (double)gridDim.x;
  }
}

// ---------------------------------------------------------------- host side

namespace {

inline unsigned blocks(long n) { return (unsigned)((n + 255) / 256); }

// 64-pixel strips while that gives >= 256 workgroups (one per CU), else
// 16-pixel strips (measured at 1024^2: 16-pixel strips cost 1.7x on the
// 256^2 level and save 2x on the 128^2 level)
template <int CA, int CB, int COUT, bool UPA, bool RELU, int MBW>
void conv_lds_launch(hipStream_t st, const ConvArgs& g, int n, unsigned rows) {
  constexpr size_t bytes = conv_lds_bytes<CA + CB, MBW>();
  static bool granted = bytes <= 65536 ||
                        hipFuncSetAttribute(reinterpret_cast<const void*>(k_conv3x3_lds<CA, CB, COUT, UPA, RELU, MBW>),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) == hipSuccess;
  if (granted) {
    hipLaunchKernelGGL((k_conv3x3_lds<CA, CB, COUT, UPA, RELU, MBW>), dim3((g.W + 16 * MBW - 1) / (16 * MBW), rows, n),
                       dim3(256), bytes, st, g);
  } else {  // the device refused the LDS: the unstaged kernel (same results)
    (void)hipGetLastError();
    hipLaunchKernelGGL((k_conv3x3<CA, CB, COUT, UPA, RELU, MBW>), dim3((g.W + 16 * MBW - 1) / (16 * MBW), rows, n),
                       dim3(256), 0, st, g);
  }
}

template <int CA, int CB, int COUT, bool UPA, bool RELU>
void conv_t(hipStream_t st, const ConvArgs& g, int n) {
  const unsigned rows = (unsigned)(g.H + 3) / 4;
  const bool wide = g.strip ? g.strip == 64 : (unsigned)((g.W + 63) / 64) * rows * (unsigned)n >= 256;
  if (wide) conv_lds_launch<CA, CB, COUT, UPA, RELU, 4>(st, g, n, rows);
  else conv_lds_launch<CA, CB, COUT, UPA, RELU, 1>(st, g, n, rows);
}

// every (ca, cb, cout, upa, relu) the network and its input gradient use
void conv(hipStream_t st, int ca, int cb, int cout, bool upa, bool relu, const ConvArgs& g, int n) {
#define EFL_NN_CONV(A, B, O, U, R) \
  if (ca == A && cb == B && cout == O && upa == U && relu == R) return conv_t<A, B, O, U, R>(st, g, n);
  // forward
  EFL_NN_CONV(16, 0, 16, false, true)
  EFL_NN_CONV(16, 0, 32, false, true)
  EFL_NN_CONV(32, 0, 32, false, true)
  EFL_NN_CONV(32, 0, 64, false, true)
  EFL_NN_CONV(64, 0, 64, false, true)
  EFL_NN_CONV(64, 64, 64, true, true)
  EFL_NN_CONV(64, 0, 32, false, true)
  EFL_NN_CONV(32, 32, 32, true, true)
  EFL_NN_CONV(32, 0, 16, false, true)
  EFL_NN_CONV(16, 16, 16, true, true)
  EFL_NN_CONV(16, 0, 16, false, false)
  // input gradients (A = dY with cout channels, output = cin channels)
  EFL_NN_CONV(32, 0, 16, false, false)
  EFL_NN_CONV(16, 0, 32, false, false)
  EFL_NN_CONV(32, 0, 32, false, false)
  EFL_NN_CONV(64, 0, 32, false, false)
  EFL_NN_CONV(64, 0, 64, false, false)
  EFL_NN_CONV(64, 0, 128, false, false)
  EFL_NN_CONV(32, 0, 64, false, false)
#undef EFL_NN_CONV
  set_launch_error("nn conv: layer shape without an instantiation");  // a layer table mismatch
}

template <int CA, int CB, int COUT, bool UPA>
void wgrad_t(hipStream_t st, const WgradArgs& g, int regions, int n) {
  dim3 grid(regions, 9 * (COUT / 16), n);
  hipLaunchKernelGGL((k_wgrad<CA, CB, COUT, UPA>), grid, dim3(256), 0, st, g);
}

void wgrad(hipStream_t st, const LayerDef& L, const WgradArgs& g, int regions, int n) {
#define EFL_NN_WG(A, B, O, U) \
  if (L.ca == A && L.cb == B && L.cout == O && (bool)L.upa == U) return wgrad_t<A, B, O, U>(st, g, regions, n);
  EFL_NN_WG(16, 0, 16, false)
  EFL_NN_WG(16, 0, 32, false)
  EFL_NN_WG(32, 0, 32, false)
  EFL_NN_WG(32, 0, 64, false)
  EFL_NN_WG(64, 0, 64, false)
  EFL_NN_WG(64, 64, 64, true)
  EFL_NN_WG(64, 0, 32, false)
  EFL_NN_WG(32, 32, 32, true)
  EFL_NN_WG(32, 0, 16, false)
  EFL_NN_WG(16, 16, 16, true)
#undef EFL_NN_WG
  set_launch_error("nn wgrad: layer shape without an instantiation");
}

void pool(hipStream_t st, int c, const uint16_t* in, uint16_t* out, int W, int H, long sin, long sout, int n,
          const double* stop) {
  dim3 grid(blocks((long)W * H * (c / 4)), n);
  if (c == 16) hipLaunchKernelGGL(k_pool2<16>, grid, dim3(256), 0, st, in, out, W, H, sin, sout, stop);
  else if (c == 32) hipLaunchKernelGGL(k_pool2<32>, grid, dim3(256), 0, st, in, out, W, H, sin, sout, stop);
  else hipLaunchKernelGGL(k_pool2<64>, grid, dim3(256), 0, st, in, out, W, H, sin, sout, stop);
}

void pool_bwd(hipStream_t st, int c, const uint16_t* gout, uint16_t* gin, int W, int H, long sout, long sin, int n) {
  dim3 grid(blocks((long)W * H * (c / 4)), n);
  if (c == 16) hipLaunchKernelGGL(k_pool2_bwd<16>, grid, dim3(256), 0, st, gout, gin, W, H, sout, sin);
  else if (c == 32) hipLaunchKernelGGL(k_pool2_bwd<32>, grid, dim3(256), 0, st, gout, gin, W, H, sout, sin);
  else hipLaunchKernelGGL(k_pool2_bwd<64>, grid, dim3(256), 0, st, gout, gin, W, H, sout, sin);
}

void cat_bwd(hipStream_t st, const LayerDef& L, const uint16_t* gcat, uint16_t* ga, uint16_t* gb, int W, int H,
             long scat, long sa, long sb, int n) {
  dim3 grid(blocks((long)W * H * ((L.ca + L.cb) / 4)), n);
  if (L.ca == 64) hipLaunchKernelGGL((k_cat_bwd<64, 64>), grid, dim3(256), 0, st, gcat, ga, gb, W, H, scat, sa, sb);
  else if (L.ca == 32) hipLaunchKernelGGL((k_cat_bwd<32, 32>), grid, dim3(256), 0, st, gcat, ga, gb, W, H, scat, sa, sb);
  else hipLaunchKernelGGL((k_cat_bwd<16, 16>), grid, dim3(256), 0, st, gcat, ga, gb, W, H, scat, sa, sb);
}

// per-sample element stride of tensor t
long stride(const Shapes& s, int t) {
  const int l = s.level(t);
  return (long)(s.w >> l) * (s.h >> l) * s.channels(t);
}

void forward_impl(hipStream_t st, const Packed& p, Activations& a, const double* stop) {
  const Shapes& s = a.s;
  for (int k = 0; k < kLayers; ++k) {
    const LayerDef& L = kNet[k];
    if (k == 2 || k == 4 || k == 6) {  // pool the skip tensor below this level
      const int src = kPoolSrc[(k - 2) / 2], dst = kTP0 + (k - 2) / 2;
      const int l = L.level;
      pool(st, s.channels(src), a.t[src], a.t[dst], s.w >> l, s.h >> l, stride(s, src), stride(s, dst), s.n, stop);
    }
    ConvArgs g{};
    g.a = a.t[L.in_a];
    g.b = L.in_b >= 0 ? a.t[L.in_b] : nullptr;
    g.w = p.wf[k];
    g.bias = p.bias[k];
    g.y = a.t[k];
    g.W = s.w >> L.level;
    g.H = s.h >> L.level;
    g.sa = stride(s, L.in_a);
    g.sb = L.in_b >= 0 ? stride(s, L.in_b) : 0;
    g.sy = stride(s, k);
    g.stop = stop;
    g.strip = a.strip[L.level];
    conv(st, L.ca, L.cb, L.cout, L.upa, L.relu, g, s.n);
  }
}

}  // namespace

void forward(hipStream_t st, const Packed& p, Activations& a) { forward_impl(st, p, a, nullptr); }

void precondition(hipStream_t st, const Packed& p, Activations& a, const DevInst* I, int W, int H,
                  const double* done_flag) {
  hipLaunchKernelGGL(k_features, dim3(blocks((long)a.s.w * a.s.h)), dim3(256), 0, st, I, a.t[kTX], a.s.w, a.s.h,
                     done_flag);
  forward_impl(st, p, a, done_flag);
  hipLaunchKernelGGL(k_output, dim3((unsigned)(((long)W * H + 1023) / 1024)), dim3(256), 0, st, I, a.t[14], a.s.w,
                     done_flag);
}

void pack(hipStream_t st, const float* params, Packed& p) {
  for (int k = 0; k < kLayers; ++k) {
    const LayerDef& L = kNet[k];
    const int cin = L.ca + L.cb;
    hipLaunchKernelGGL(k_pack, dim3(blocks(9L * L.cout * cin)), dim3(256), 0, st, params + layer_weight_offset(k),
                       p.wf[k], p.wb[k], L.cout, cin);
    p.bias[k] = const_cast<float*>(params + layer_bias_offset(k));
  }
}

int trainer_reserve_batch(Trainer& t, int n) {
  if (t.act_n >= n) return 0;
  Activations& a = t.act;
  for (int k = 0; k < kTensors; ++k) {
    if (a.t[k]) hipFree(a.t[k]);
    if (a.g[k]) hipFree(a.g[k]);
    a.t[k] = a.g[k] = nullptr;
  }
  if (a.scratch) hipFree(a.scratch);
  if (t.partial) hipFree(t.partial);
  if (t.idx) hipFree(t.idx);
  if (t.target) hipFree(t.target);
  if (t.mask) hipFree(t.mask);
  if (t.rsums) hipFree(t.rsums);
  a.scratch = nullptr;
  t.partial = nullptr;
  t.idx = nullptr;
  t.target = t.mask = nullptr;
  t.rsums = nullptr;
  t.act_n = 0;
  a.s = Shapes{n, t.w, t.h};
  size_t scratch = 0, part = 0;
  for (int k = 0; k < kTensors; ++k) {
    if (hipMalloc(&a.t[k], a.s.elems(k) * 2) != hipSuccess) return -1;
    if (hipMalloc(&a.g[k], a.s.elems(k) * 2) != hipSuccess) return -1;
  }
  for (int k = 0; k < kLayers; ++k) {
    const LayerDef& L = kNet[k];
    const int w = t.w >> L.level, h = t.h >> L.level;
    scratch = std::max(scratch, (size_t)n * w * h * (L.ca + L.cb));
    const size_t regions = (size_t)((w + 63) / 64) * ((h + 63) / 64);
    part = std::max(part, regions * n * 9 * L.cout * (L.ca + L.cb) + regions * n * L.cout);
  }
  if (hipMalloc(&a.scratch, scratch * 2) != hipSuccess) return -1;
  if (hipMalloc(&t.partial, part * 4) != hipSuccess) return -1;
  t.partial_floats = part;
  if (hipMalloc(&t.idx, n * sizeof(int)) != hipSuccess) return -1;
  if (hipMalloc(&t.target, (size_t)n * t.w * t.h * 4) != hipSuccess) return -1;
  if (hipMalloc(&t.mask, (size_t)n * t.w * t.h * 4) != hipSuccess) return -1;
  if (hipMalloc(&t.rsums, (size_t)n * 3 * sizeof(double)) != hipSuccess) return -1;
  t.act_n = n;
  return 0;
}

void eval_batch(hipStream_t st, Trainer& t, const Packed& p, const int* host_idx, int n, bool grads) {
  Activations& a = t.act;
  a.s.n = n;
  const Shapes& s = a.s;
  const long npx = (long)t.w * t.h;
  hipMemcpyAsync(t.idx, host_idx, n * sizeof(int), hipMemcpyHostToDevice, st);
  hipLaunchKernelGGL(k_gather, dim3(blocks(npx), n), dim3(256), 0, st, t.data, t.idx, a.t[kTX], t.target, t.mask,
                     t.w, t.h);
  forward_impl(st, p, a, nullptr);
  hipMemsetAsync(t.loss, 0, sizeof(double), st);
  if (t.loss_mode == kLossResidual) {
    hipMemsetAsync(t.rsums, 0, sizeof(double) * 3 * n, st);
    hipLaunchKernelGGL(k_rloss_fwd, dim3(blocks(npx), n), dim3(256), 0, st, t.data, t.idx, a.t[14], t.target,
                       t.rsums, t.w, t.h);
    hipLaunchKernelGGL(k_rloss_bwd, dim3(blocks(npx), n), dim3(256), 0, st, t.data, t.idx, t.target, t.rsums,
                       grads ? a.g[14] : nullptr, t.loss, t.w, t.h, 1.0f / (float)n);
  } else {
    const float inv_count = 1.0f / (float)(npx * n);
    hipLaunchKernelGGL(k_loss, dim3(blocks(npx * n)), dim3(256), 0, st, a.t[14], t.target, t.mask,
                       grads ? a.g[14] : nullptr, npx * n, inv_count, t.loss);
  }
  if (!grads) return;
  for (int k = kLayers - 1; k >= 0; --k) {
    const LayerDef& L = kNet[k];
    const int w = t.w >> L.level, h = t.h >> L.level;
    const long sy = stride(s, k);
    if (L.relu) hipLaunchKernelGGL(k_relu_bwd, dim3(blocks(sy * n / 4)), dim3(256), 0, st, a.g[k], a.t[k], sy * n / 4);
    // weight and bias gradients
    WgradArgs wg{};
    wg.a = a.t[L.in_a];
    wg.b = L.in_b >= 0 ? a.t[L.in_b] : nullptr;
    wg.dy = a.g[k];
    const int rx = (w + 63) / 64, ry = (h + 63) / 64;
    const size_t M = 9L * L.cout * (L.ca + L.cb);
    wg.part = t.partial;
    wg.bpart = t.partial + (size_t)rx * ry * n * M;
    wg.W = w;
    wg.H = h;
    wg.rx = rx;
    wg.sa = stride(s, L.in_a);
    wg.sb = L.in_b >= 0 ? stride(s, L.in_b) : 0;
    wg.sdy = sy;
    wgrad(st, L, wg, rx * ry, n);
    hipLaunchKernelGGL(k_wreduce, dim3(blocks((long)M + L.cout)), dim3(256), 0, st, wg.part, wg.bpart, rx * ry * n,
                       L.cout, L.ca + L.cb, t.grads + layer_weight_offset(k), t.grads + layer_bias_offset(k));
    if (k == 0) break;
    // input gradient: conv of dY with the flipped, transposed weights
    const bool cat = L.in_b >= 0;
    ConvArgs g{};
    g.a = a.g[k];
    g.w = p.wb[k];
    g.y = cat ? a.scratch : a.g[L.in_a];
    g.W = w;
    g.H = h;
    g.sa = sy;
    g.sy = (long)w * h * (L.ca + L.cb);
    conv(st, L.cout, 0, L.ca + L.cb, false, false, g, n);
    if (cat)
      cat_bwd(st, L, a.scratch, a.g[L.in_a], a.g[L.in_b], w / 2, h / 2, g.sy, stride(s, L.in_a), stride(s, L.in_b), n);
    if (k == 2 || k == 4 || k == 6) {  // through the pooling into the skip tensor
      const int src = kPoolSrc[(k - 2) / 2], dst = kTP0 + (k - 2) / 2;
      pool_bwd(st, s.channels(src), a.g[dst], a.g[src], w, h, stride(s, dst), stride(s, src), n);
    }
  }
}

void adam_step(hipStream_t st, Trainer& t, float lr) {
  t.adam_t += 1;
  const float c1 = 1.0f - std::pow(0.9f, (float)t.adam_t), c2 = 1.0f - std::pow(0.999f, (float)t.adam_t);
  const long n = (long)param_count();
  hipLaunchKernelGGL(k_adam, dim3(blocks(n)), dim3(256), 0, st, t.params, t.grads, t.m, t.v, n, lr, c1, c2);
}

void make_sample(hipStream_t st, Trainer& t, uint32_t slot, const DevInst* I, int W, int H, const float* dev_r,
                 const float* dev_t, float sigma, float trms) {
  float* dst = t.data + (size_t)slot * kSamplePlanes * W * H;
  hipLaunchKernelGGL(k_make_sample, dim3(blocks((long)W * H)), dim3(256), 0, st, dst, I, W, H, dev_r, dev_t,
                     1.0f / sigma, 1.0f / trms);
}

void read_output(hipStream_t st, const Activations& a, int n, float* dev_out) {
  const long npx = (long)a.s.w * a.s.h * n;
  hipLaunchKernelGGL(k_read_out, dim3(blocks(npx)), dim3(256), 0, st, a.t[14], dev_out, npx);
}

// He-normal weights (std sqrt(2 / fan_in)), zero biases; the output layer
// starts small and only its channel 0 is used.  SplitMix64 + Box-Muller.
void init_params(float* host, uint64_t seed) {
  uint64_t s = seed;
  auto next = [&s]() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  auto uni = [&]() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); };
  for (int k = 0; k < kLayers; ++k) {
    const LayerDef& L = kNet[k];
    const int cin = L.ca + L.cb;
    const int fan_in = 9 * (k == 0 ? kFeatures : cin);
    const double sd = k == kLayers - 1 ? 1e-2 / std::sqrt((double)fan_in) : std::sqrt(2.0 / fan_in);
    float* w = host + layer_weight_offset(k);
    for (int co = 0; co < L.cout; ++co)
      for (int ci = 0; ci < cin; ++ci)
        for (int tap = 0; tap < 9; ++tap) {
          const double g = std::sqrt(-2.0 * std::log(uni())) * std::cos(6.283185307179586 * uni());
          bool used = !(k == 0 && ci >= kFeatures) && !(k == kLayers - 1 && co != 0);
          w[((size_t)co * cin + ci) * 9 + tap] = used ? (float)(sd * g) : 0.0f;
        }
    float* b = host + layer_bias_offset(k);
    for (int co = 0; co < L.cout; ++co) b[co] = 0.0f;
  }
}

}  // namespace nn
}  // namespace efl
