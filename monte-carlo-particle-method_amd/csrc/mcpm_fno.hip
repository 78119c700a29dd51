// mcpm_fno.hip — spectral neural-operator surrogate of the periodic plane's
// screened-Poisson solve: forward, backward and Adam kernels, the snapshot
// store, and the mcpm_fno_* entry points of the C ABI.
//
// The network (width 64, four Fourier layers, m retained modes per axis)
// maps the nodal density of a periodic plane to the Q1 FEM solution of
// (K + lambda^2 M) phi = M n that compute_SP_2D (FEM.cpp:249-610) solves
// before its core subtraction (FEM.cpp:497-554).  The reference has no such
// network; the exact solve stays the arbiter: a field whose FEM residual
// exceeds the caller's bound is finished by conjugate gradients started from
// the network's field.
//
// Layouts (p = j N0 + i over the unique nodes of the periodic plane):
//   activations  H[p][c] BF16, Z[p][c] FP32, channel fastest
//   x-DFT table  Tf[k'][x] BF16 (k' < m: cos, k' >= m: -sin), Tinv[x][k']
//   y twiddles   twy[a][y] = (cos, sin)(2 pi ky(a) y / N1), ky(a) = a (a < m)
//                or N1 - 2m + a
//   x-DFT output A[y][k'][c] FP32 (k' < m real parts, k' >= m imaginary)
//   modes        X[a][kx][c] complex FP32; spectral weights R[a][kx][i][o]
//   y-inverse    Cp[y][o][k'] BF16 (B operand of the combine GEMM)
// GEMM-shaped stages run on v_mfma_f32_16x16x16_bf16 (BF16 in, FP32
// accumulate): the x-DFT (2m x 64 per row, K = N0), the combine
// z = W h + Tinv Cp + b (64 nodes x 64 channels per workgroup,
// K = 64 + 2m) and the weight gradients (64 x 64, K = nodes).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "mcpm_internal.h"
#include "mcpm_kernels.h"
#include "mcpm_plane.h"

namespace mcpm {
namespace fno {

constexpr int C = 64;
constexpr int LAYERS = 4;
constexpr int CHUNK = 1024;  // nodes per workgroup of the gradient reductions

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ short bits(uint16_t v) { return (short)v; }
__device__ __forceinline__ short bits(float v) { return (short)f2bf(v); }

// four consecutive elements as BF16 bits
__device__ __forceinline__ bf16x4 load4(const uint16_t *p) {
  const ushort4 v = *(const ushort4 *)p;
  bf16x4 r;
  r[0] = (short)v.x;
  r[1] = (short)v.y;
  r[2] = (short)v.z;
  r[3] = (short)v.w;
  return r;
}
__device__ __forceinline__ bf16x4 load4(const float *p) {
  const float4 v = *(const float4 *)p;
  bf16x4 r;
  r[0] = bits(v.x);
  r[1] = bits(v.y);
  r[2] = bits(v.z);
  r[3] = bits(v.w);
  return r;
}

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// GELU in the tanh form, written as z sigma(2u), u = sqrt(2/pi)(z + 0.044715 z^3)
// (0.5 (1 + tanh u) = sigma(2u)), with the native exponential and division.
constexpr float GELU_K = 0.7978845608028654f;  // sqrt(2/pi)
__device__ __forceinline__ float gelu_sig(float z) {
  const float u = GELU_K * (z + 0.044715f * z * z * z);
  return __fdividef(1.0f, 1.0f + __expf(-2.0f * u));
}
__device__ __forceinline__ float gelu(float z) { return z * gelu_sig(z); }
__device__ __forceinline__ float gelu_d(float z) {
  const float sg = gelu_sig(z);
  return sg + 2.0f * z * sg * (1.0f - sg) * GELU_K * (1.0f + 3.0f * 0.044715f * z * z);
}

// x-DFT of one row per workgroup: A[y] (2m x 64) = Tf (2m x N0) . act[y] (N0 x 64).
// Wave w takes a quarter of the row (all 64 channels, all 2m/16 row tiles);
// the four partial sums meet in LDS.
template <typename T>
__global__ __launch_bounds__(256) void k_dftx(const T *__restrict__ act,
                                              const uint16_t *__restrict__ tf,
                                              float *__restrict__ A, int n0, int m2) {
  __shared__ float red[4][64 * 64];
  const int y = blockIdx.x, w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int r16 = l & 15, q = l >> 4;
  const int nrt = m2 >> 4;
  f32x4 acc[4][4];  // [row tile][channel tile]
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[t][c] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int xq = n0 / 4, xb = w * xq;  // n0 is a multiple of 64
  const T *row = act + (size_t)y * n0 * C + r16;
  for (int x0 = xb; x0 < xb + xq; x0 += 16) {
    bf16x4 b[4];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int j = 0; j < 4; ++j) b[c][j] = bits(row[(size_t)(x0 + 4 * q + j) * C + 16 * c]);
#pragma unroll
    for (int t = 0; t < 4; ++t)
      if (t < nrt) {
        const bf16x4 a = load4(tf + (size_t)(16 * t + r16) * n0 + x0 + 4 * q);
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[t][c] = mfma(a, b[c], acc[t][c]);
      }
  }
  // D[row = k' = 16 t + 4 q + j][col = channel 16 c + r16]
#pragma unroll
  for (int t = 0; t < 4; ++t)
    if (t < nrt)
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int j = 0; j < 4; ++j) red[w][(16 * t + 4 * q + j) * 64 + 16 * c + r16] = acc[t][c][j];
  __syncthreads();
  for (int e = threadIdx.x; e < m2 * 64; e += 256)
    A[(size_t)y * m2 * C + e] = (red[0][e] + red[1][e]) + (red[2][e] + red[3][e]);
}

__device__ __forceinline__ f32x4 mfma4(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// y-DFT of the retained kx on FP32 MFMA:
// X[a][kx][c] = s(kx) sum_y A[y][kx][c] e^{-i phi(a, y)} (A complex from the
// packed re | im rows).  Workgroup: one kx (blockIdx.x), a 16-row tile of a
// (blockIdx.y), a 16-channel tile (blockIdx.z); its four waves split the rows
// y and meet in LDS.  D[a][c] = sum_y tw(a, y) A[y][c].
__global__ __launch_bounds__(256) void k_dfty(const float *__restrict__ A,
                                              const float2 *__restrict__ twy,
                                              float2 *__restrict__ X, int n1, int m, float s0,
                                              float s1) {
  __shared__ float red[4][64][8];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int r16 = l & 15, q = l >> 4;
  const int kx = blockIdx.x, a0 = blockIdx.y * 16, c0 = blockIdx.z * 16, m2 = 2 * m;
  f32x4 xr = {0.f, 0.f, 0.f, 0.f}, xi = {0.f, 0.f, 0.f, 0.f};
  const int yq = n1 / 4, yb = w * yq;  // n1 is a multiple of 16
  const float2 *tw = twy + (size_t)(a0 + r16) * n1 + q;
  const float *ar = A + (size_t)kx * C + c0 + r16 + (size_t)q * m2 * C;
  const float *ai = ar + (size_t)m * C;
  for (int y0 = yb; y0 < yb + yq; y0 += 8) {  // yq is a multiple of 4; two steps in flight
    const float2 t = tw[y0];
    const float br = ar[(size_t)y0 * m2 * C], bi = ai[(size_t)y0 * m2 * C];
    const bool two = y0 + 4 < yb + yq;
    const float2 t2 = two ? tw[y0 + 4] : make_float2(0.f, 0.f);
    const float br2 = two ? ar[(size_t)(y0 + 4) * m2 * C] : 0.f;
    const float bi2 = two ? ai[(size_t)(y0 + 4) * m2 * C] : 0.f;
    xr = mfma4(t.x, br, xr);
    xr = mfma4(t.y, bi, xr);
    xi = mfma4(t.x, bi, xi);
    xi = mfma4(-t.y, br, xi);
    xr = mfma4(t2.x, br2, xr);
    xr = mfma4(t2.y, bi2, xr);
    xi = mfma4(t2.x, bi2, xi);
    xi = mfma4(-t2.y, br2, xi);
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    red[w][l][j] = xr[j];
    red[w][l][4 + j] = xi[j];
  }
  __syncthreads();
  if (w == 0) {
    const float sc = kx ? s1 : s0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float vr = (red[0][l][j] + red[1][l][j]) + (red[2][l][j] + red[3][l][j]);
      const float vi = (red[0][l][4 + j] + red[1][l][4 + j]) + (red[2][l][4 + j] + red[3][l][4 + j]);
      X[((size_t)(a0 + 4 * q + j) * m + kx) * C + c0 + r16] = make_float2(vr * sc, vi * sc);
    }
  }
}

// Spectral mixing Y[mode][o] = sum_i X[mode][i] R[mode][i][o]; wave -> mode, lane -> o.
__global__ __launch_bounds__(256) void k_mix(const float2 *__restrict__ X,
                                             const float2 *__restrict__ R,
                                             float2 *__restrict__ Y) {
  const int mode = blockIdx.x * 4 + (threadIdx.x >> 6), o = threadIdx.x & 63;
  const float2 *xm = X + (size_t)mode * C;
  const float2 *rm = R + (size_t)mode * C * C + o;
  float yr = 0.f, yi = 0.f;
  for (int i = 0; i < C; ++i) {
    const float2 x = xm[i], r = rm[(size_t)i * C];
    yr += x.x * r.x - x.y * r.y;
    yi += x.x * r.y + x.y * r.x;
  }
  Y[(size_t)mode * C + o] = make_float2(yr, yi);
}

// Backward of the mixing: dX[mode][i] = sum_o dY[mode][o] conj(R[mode][i][o]).
__global__ __launch_bounds__(256) void k_mix_dx(const float2 *__restrict__ dY,
                                                const float2 *__restrict__ R,
                                                float2 *__restrict__ dX) {
  const int mode = blockIdx.x * 4 + (threadIdx.x >> 6), i = threadIdx.x & 63;
  const float2 *dym = dY + (size_t)mode * C;
  const float2 *rm = R + ((size_t)mode * C + i) * C;
  float xr = 0.f, xi = 0.f;
  for (int o = 0; o < C; ++o) {
    const float2 d = dym[o], r = rm[o];
    xr += d.x * r.x + d.y * r.y;
    xi += d.y * r.x - d.x * r.y;
  }
  dX[(size_t)mode * C + i] = make_float2(xr, xi);
}

// dR[mode][i][o] = conj(X[mode][i]) dY[mode][o]; wave -> mode, lane -> o.
__global__ __launch_bounds__(256) void k_mix_dr(const float2 *__restrict__ X,
                                                const float2 *__restrict__ dY,
                                                float2 *__restrict__ dR, int accumulate) {
  const int mode = blockIdx.x * 4 + (threadIdx.x >> 6), o = threadIdx.x & 63;
  const float2 d = dY[(size_t)mode * C + o];
  const float2 *xm = X + (size_t)mode * C;
  float2 *gm = dR + (size_t)mode * C * C + o;
  for (int i = 0; i < C; ++i) {
    const float2 x = xm[i];
    float2 g = make_float2(x.x * d.x + x.y * d.y, x.x * d.y - x.y * d.x);
    if (accumulate) {
      const float2 g0 = gm[(size_t)i * C];
      g.x += g0.x;
      g.y += g0.y;
    }
    gm[(size_t)i * C] = g;
  }
}

__device__ __forceinline__ uint32_t pack2(float a, float b) {
  return (uint32_t)f2bf(a) | ((uint32_t)f2bf(b) << 16);
}

// y-inverse: Cp[y][o][kx | m + kx] = bf16(s(kx) sum_a Y[a][kx][o] e^{+i phi(a, y)}),
// on FP32 MFMA: D[o][y] = sum_a Y[a][kx][o] (cos, sin)(a, y) for a 16 x 16
// tile (wave w: channels 16w.., workgroup: rows 16 blockIdx.x..) and a chunk
// of 8 kx (blockIdx.y), so that a lane stores 8 consecutive k' of one (y, o).
__global__ __launch_bounds__(256) void k_idfty(const float2 *__restrict__ Y,
                                               const float2 *__restrict__ twy,
                                               uint16_t *__restrict__ Cp, int n1, int m, float s0,
                                               float s1) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int r16 = l & 15, q = l >> 4;
  const int y0 = blockIdx.x * 16, kc = blockIdx.y * 8, m2 = 2 * m;
  f32x4 ar[8], ai[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) ar[k] = ai[k] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int a0 = 0; a0 < m2; a0 += 4) {
    const int a = a0 + q;
    const float2 tw = twy[(size_t)a * n1 + y0 + r16];
    const float2 *ya = Y + ((size_t)a * m + kc) * C + 16 * w + r16;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const float2 v = ya[(size_t)k * C];
      ar[k] = mfma4(v.x, tw.x, ar[k]);
      ar[k] = mfma4(-v.y, tw.y, ar[k]);
      ai[k] = mfma4(v.y, tw.x, ai[k]);
      ai[k] = mfma4(v.x, tw.y, ai[k]);
    }
  }
  float sc[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) sc[k] = (kc + k) ? s1 : s0;
  const int y = y0 + r16;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int o = 16 * w + 4 * q + j;
    uint16_t *dst = Cp + ((size_t)y * C + o) * m2;
    *(uint4 *)(dst + kc) =
        make_uint4(pack2(ar[0][j] * sc[0], ar[1][j] * sc[1]), pack2(ar[2][j] * sc[2], ar[3][j] * sc[3]),
                   pack2(ar[4][j] * sc[4], ar[5][j] * sc[5]), pack2(ar[6][j] * sc[6], ar[7][j] * sc[7]));
    *(uint4 *)(dst + m + kc) =
        make_uint4(pack2(ai[0][j] * sc[0], ai[1][j] * sc[1]), pack2(ai[2][j] * sc[2], ai[3][j] * sc[3]),
                   pack2(ai[4][j] * sc[4], ai[5][j] * sc[5]), pack2(ai[6][j] * sc[6], ai[7][j] * sc[7]));
  }
}

enum : int { CMB_FWD = 0, CMB_LAST = 1, CMB_BWD = 2 };

// out(64 nodes x 64 channels) = in (64 x 64) . Wt^T + Tinv (64 x 2m) . Cp[y] (2m x 64),
// Wt[n][k] (n = output channel).  FWD: z = out + b, Z = z (if given),
// H = bf16(gelu(z)).  LAST: z = out + b, Z = z (if given), yout = wq . z + bq.
// BWD: dh = out; outZ = dh gelu'(Zprev) (or dh when Zprev is null).
// The MFMA computes the transpose, D[channel][node], so that a lane holds
// four consecutive channels of one node: 8- and 16-byte stores.
template <int MODE, typename TIn, int TILES>
__global__ __launch_bounds__(256) void k_combine(
    const TIn *__restrict__ in, const uint16_t *__restrict__ Wt, const uint16_t *__restrict__ tinv,
    const uint16_t *__restrict__ Cp, const float *__restrict__ bias, const float *__restrict__ wq,
    const float *__restrict__ Zprev, float *__restrict__ Z, uint16_t *__restrict__ H,
    float *__restrict__ yout, int n0, int m2) {
  constexpr int tiles = TILES;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int r16 = l & 15, q = l >> 4;
  // `tiles` consecutive 64-node tiles of one row; the weights and the row's
  // spectral coefficients (A operands) stay in registers across them
  const size_t pbase = (size_t)blockIdx.x * 64 * tiles;
  const int y = (int)(pbase / n0), xbase = (int)(pbase - (size_t)y * n0);
  bf16x4 wa[4][4], ca[4][4];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      wa[t][ks] = load4(Wt + (16 * t + r16) * C + 16 * ks + 4 * q);
      if (16 * ks < m2)
        ca[t][ks] = load4(Cp + ((size_t)y * C + 16 * t + r16) * m2 + 16 * ks + 4 * q);
      else
        ca[t][ks] = bf16x4{0, 0, 0, 0};
    }
  float bb[4][4], wqv[4][4];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      bb[t][j] = MODE == CMB_BWD ? 0.f : bias[16 * t + 4 * q + j];
      wqv[t][j] = MODE == CMB_LAST ? wq[16 * t + 4 * q + j] : 0.f;
    }
#pragma unroll
  for (int tt = 0; tt < tiles; ++tt) {
    f32x4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const size_t node = pbase + (size_t)tt * 64 + 16 * w + r16;
    const TIn *brow = in + node * C + 4 * q;
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const bf16x4 b = load4(brow + 16 * ks);
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[t] = mfma(wa[t][ks], b, acc[t]);
    }
    const uint16_t *trow = tinv + (size_t)(xbase + tt * 64 + 16 * w + r16) * m2 + 4 * q;
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      if (16 * ks < m2) {
        const bf16x4 b = load4(trow + 16 * ks);
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[t] = mfma(ca[t][ks], b, acc[t]);
      }
    }
    // lane: node `node`, channels 16 t + 4 q + j
    float pj = 0.f;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const size_t e = node * C + 16 * t + 4 * q;
      if (MODE == CMB_BWD) {
        float4 r = make_float4(acc[t][0], acc[t][1], acc[t][2], acc[t][3]);
        if (Zprev) {
          const float4 zp = *(const float4 *)(Zprev + e);
          r.x *= gelu_d(zp.x);
          r.y *= gelu_d(zp.y);
          r.z *= gelu_d(zp.z);
          r.w *= gelu_d(zp.w);
        }
        *(float4 *)(Z + e) = r;
      } else {
        const float4 z = make_float4(acc[t][0] + bb[t][0], acc[t][1] + bb[t][1],
                                     acc[t][2] + bb[t][2], acc[t][3] + bb[t][3]);
        if (Z) *(float4 *)(Z + e) = z;
        if (MODE == CMB_FWD) {
          ushort4 hv;
          hv.x = f2bf(gelu(z.x));
          hv.y = f2bf(gelu(z.y));
          hv.z = f2bf(gelu(z.z));
          hv.w = f2bf(gelu(z.w));
          *(ushort4 *)(H + e) = hv;
        } else {
          pj += wqv[t][0] * z.x + wqv[t][1] * z.y + wqv[t][2] * z.z + wqv[t][3] * z.w;
        }
      }
    }
    if (MODE == CMB_LAST) {
      pj += __shfl_xor(pj, 16);
      pj += __shfl_xor(pj, 32);
      if (q == 0) yout[node] = pj + wq[C];
    }
  }
}

// Weight-gradient partials of one chunk of nodes:
// part[chunk][o][i] = sum_p bf16(dZ[p][o]) H[p][i], part[chunk][C*C + o] = sum_p dZ[p][o].
// Wave w owns output rows 16w..16w+15.
__global__ __launch_bounds__(256) void k_wgrad(const float *__restrict__ dZ,
                                               const uint16_t *__restrict__ H,
                                               float *__restrict__ part, int64_t P) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int r16 = l & 15, q = l >> 4;
  const int64_t pb = (int64_t)blockIdx.x * CHUNK, pe = min(P, pb + (int64_t)CHUNK);
  f32x4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  float bs = 0.f;
  for (int64_t p0 = pb; p0 < pe; p0 += 16) {
    bf16x4 a;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float d = dZ[(size_t)(p0 + 4 * q + j) * C + 16 * w + r16];
      bs += d;
      a[j] = bits(d);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      bf16x4 b;
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = bits(H[(size_t)(p0 + 4 * q + j) * C + 16 * t + r16]);
      acc[t] = mfma(a, b, acc[t]);
    }
  }
  float *dst = part + (size_t)blockIdx.x * (C * C + C);
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int j = 0; j < 4; ++j) dst[(16 * w + 4 * q + j) * C + 16 * t + r16] = acc[t][j];
  bs += __shfl_xor(bs, 16);
  bs += __shfl_xor(bs, 32);
  if (q == 0) dst[C * C + 16 * w + r16] = bs;
}

// Column sums of 256 nodes per workgroup:
// part[blk] = {sum_p M[p][c] v[p] (64), sum_p M[p][c] (64), sum_p v[p]}.
constexpr int COLSUM_NODES = 256;
__global__ __launch_bounds__(256) void k_colsum(const float *__restrict__ M,
                                                const float *__restrict__ v,
                                                float *__restrict__ part, int64_t P) {
  __shared__ float sh[4][2 * C + 1];
  const int w = threadIdx.x >> 6, c = threadIdx.x & 63;
  const int64_t pb = (int64_t)blockIdx.x * COLSUM_NODES + w * (COLSUM_NODES / 4);
  const int64_t pe = min(P, pb + (int64_t)(COLSUM_NODES / 4));
  float smv = 0.f, sm = 0.f, sv = 0.f;
  for (int64_t p = pb; p < pe; ++p) {
    const float mv = M[(size_t)p * C + c], vv = v[p];
    smv += mv * vv;
    sm += mv;
    sv += vv;
  }
  sh[w][c] = smv;
  sh[w][C + c] = sm;
  if (c == 0) sh[w][2 * C] = sv;
  __syncthreads();
  if (w == 0) {
    float *dst = part + (size_t)blockIdx.x * (2 * C + 1);
    dst[c] = sh[0][c] + sh[1][c] + sh[2][c] + sh[3][c];
    dst[C + c] = sh[0][C + c] + sh[1][C + c] + sh[2][C + c] + sh[3][C + c];
    if (c == 0) dst[2 * C] = sh[0][2 * C] + sh[1][2 * C] + sh[2][2 * C] + sh[3][2 * C];
  }
}

// dst[k] = sum_chunk part[chunk * stride + off + k], k < len: 64 entries per
// workgroup, the 16 waves split the chunks.
__global__ __launch_bounds__(1024) void k_reduce(const float *__restrict__ part, int nchunk,
                                                 int stride, int off, int len,
                                                 float *__restrict__ dst) {
  __shared__ float sh[16][64];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int k = blockIdx.x * 64 + l;
  float s = 0.f;
  if (k < len)
    for (int b = w; b < nchunk; b += 16) s += part[(size_t)b * stride + off + k];
  sh[w][l] = s;
  __syncthreads();
  if (w == 0 && k < len) {
    float t = 0.f;
    for (int i = 0; i < 16; ++i) t += sh[i][l];
    dst[k] = t;
  }
}

// h0[p][c] = bf16(wl[c] u[p] + bl[c]); a thread writes 8 channels.
__global__ __launch_bounds__(256) void k_lift(const float *__restrict__ u,
                                              const float *__restrict__ wl,
                                              uint16_t *__restrict__ H, int64_t P) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= P * (C / 8)) return;
  const int c0 = (int)(id & (C / 8 - 1)) * 8;
  const float uv = u[id >> 3];
  uint32_t wv[4];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    wv[k] = (uint32_t)f2bf(wl[c0 + 2 * k] * uv + wl[C + c0 + 2 * k]) |
            ((uint32_t)f2bf(wl[c0 + 2 * k + 1] * uv + wl[C + c0 + 2 * k + 1]) << 16);
  *(uint4 *)(H + id * 8) = make_uint4(wv[0], wv[1], wv[2], wv[3]);
}

// dZ[p][c] = dy[p] wq[c]
__global__ __launch_bounds__(256) void k_proj_bwd(const float *__restrict__ dy,
                                                  const float *__restrict__ wq,
                                                  float *__restrict__ dZ, int64_t P) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= P * C) return;
  dZ[id] = dy[id >> 6] * wq[id & (C - 1)];
}

__device__ double block_sum(double v, double *sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sh[i];
  return s;
}

// Loss partials and dL/dy (normalised units, e = y - t, periodic forward
// differences): part[b] = sum (|grad e|^2 + w0 e^2); dy = 2 (-lap e + w0 e) / D.
__global__ __launch_bounds__(256) void k_loss(const float *__restrict__ yv,
                                              const float *__restrict__ t,
                                              const double *__restrict__ D, float *__restrict__ dy,
                                              double *__restrict__ part, int n0, int n1, float w0) {
  __shared__ double sh[4];
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  if (p < (int64_t)n0 * n1) {
    const int j = (int)(p / n0), i = (int)(p - (int64_t)j * n0);
    const int ir = i + 1 == n0 ? 0 : i + 1, il = i ? i - 1 : n0 - 1;
    const int ju = j + 1 == n1 ? 0 : j + 1, jd = j ? j - 1 : n1 - 1;
    auto E = [&](int ii, int jj) {
      const int64_t q = (int64_t)jj * n0 + ii;
      return yv[q] - t[q];
    };
    const float e = E(i, j), er = E(ir, j), el = E(il, j), eu = E(i, ju), ed = E(i, jd);
    const float gx = er - e, gy = eu - e;
    acc = (double)gx * gx + (double)gy * gy + (double)w0 * e * e;
    const double inv = 2.0 / D[0];
    dy[p] = (float)(inv * ((double)(4.0f * e - er - el - eu - ed) + (double)w0 * e));
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Partial denominator sums (|grad t|^2 + w0 t^2) of a snapshot, 256 nodes per workgroup.
__global__ __launch_bounds__(256) void k_tden(const float *__restrict__ t,
                                              double *__restrict__ part, int n0, int n1,
                                              float w0) {
  __shared__ double sh[4];
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  if (p < (int64_t)n0 * n1) {
    const int j = (int)(p / n0), i = (int)(p - (int64_t)j * n0);
    const int ir = i + 1 == n0 ? 0 : i + 1, ju = j + 1 == n1 ? 0 : j + 1;
    const double v = t[p], gx = t[(int64_t)j * n0 + ir] - v, gy = t[(int64_t)ju * n0 + i] - v;
    acc = gx * gx + gy * gy + (double)w0 * v * v;
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// D[0] = sum of the partials.
__global__ __launch_bounds__(1024) void k_sum_final(const double *__restrict__ part, int nb,
                                                    double *__restrict__ D) {
  __shared__ double sh[16];
  double acc = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) acc += part[b];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) D[0] = s;
}

__global__ __launch_bounds__(1024) void k_loss_final(const double *__restrict__ part, int nb,
                                                     const double *__restrict__ D,
                                                     float *__restrict__ out) {
  __shared__ double sh[16];
  double acc = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) acc += part[b];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) out[0] = (float)(s / D[0]);
}

__global__ __launch_bounds__(256) void k_adam(float *__restrict__ p, const float *__restrict__ g,
                                              float *__restrict__ m, float *__restrict__ v,
                                              int64_t n, float lr, float b1, float b2, float eps,
                                              float bc1, float bc2) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float gk = g[k];
  const float mk = b1 * m[k] + (1.0f - b1) * gk;
  const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
  m[k] = mk;
  v[k] = vk;
  p[k] -= lr * (mk / bc1) / (sqrtf(vk / bc2) + eps);
}

// BF16 copies of the four channel-mixing matrices: Wf[l][o][i], Wb[l][i][o].
__global__ __launch_bounds__(256) void k_cast(const float *__restrict__ params, int64_t base,
                                              int64_t lsz, uint16_t *__restrict__ Wf,
                                              uint16_t *__restrict__ Wb) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= LAYERS * C * C) return;
  const int l = id / (C * C), k = id - l * C * C, o = k / C, i = k - o * C;
  const uint16_t b = f2bf(params[base + l * lsz + k]);
  Wf[id] = b;
  Wb[l * C * C + i * C + o] = b;
}

// Plane nodes (N0+1)(N1+1) <-> unique periodic nodes N0 N1.
__global__ __launch_bounds__(256) void k_from_nodes(const double *__restrict__ nodes,
                                                    float *__restrict__ out, int n0, int n1,
                                                    double shift, double scale) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (int64_t)n0 * n1) return;
  const int j = (int)(p / n0), i = (int)(p - (int64_t)j * n0);
  out[p] = (float)((nodes[(int64_t)j * (n0 + 1) + i] - shift) * scale);
}

__global__ __launch_bounds__(256) void k_to_nodes(const float *__restrict__ yv,
                                                  double *__restrict__ nodes, int n0, int n1,
                                                  double shift, double inv_scale) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int w = n0 + 1;
  if (id >= (int64_t)w * (n1 + 1)) return;
  const int j = (int)(id / w), i = (int)(id - (int64_t)j * w);
  const int i0 = i == n0 ? 0 : i, j0 = j == n1 ? 0 : j;
  nodes[id] = (double)yv[(int64_t)j0 * n0 + i0] * inv_scale + shift;
}

}  // namespace fno
}  // namespace mcpm

using namespace mcpm;
using namespace mcpm::fno;

namespace {

#define FHIP(expr)                                                                 \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      fprintf(stderr, "mcpm_fno: %s failed: %s (%s:%d)\n", #expr,                 \
              hipGetErrorString(e_), __FILE__, __LINE__);                          \
      return MCPM_ERR_HIP;                                                         \
    }                                                                              \
  } while (0)

#define FLAUNCH()                                                                  \
  do {                                                                             \
    hipError_t e_ = hipGetLastError();                                             \
    if (e_ != hipSuccess) {                                                        \
      fprintf(stderr, "mcpm_fno: launch failed: %s (%s:%d)\n",                    \
              hipGetErrorString(e_), __FILE__, __LINE__);                          \
      return MCPM_ERR_HIP;                                                         \
    }                                                                              \
  } while (0)

struct Layout {
  int64_t wl = 0, base = 2 * C, lsz = 0, wq = 0, n = 0;
  int64_t W(int l) const { return base + l * lsz; }
  int64_t b(int l) const { return W(l) + C * C; }
  int64_t R(int l) const { return W(l) + C * C + C; }
};

Layout make_layout(int m) {
  Layout L;
  L.lsz = (int64_t)C * C + C + (int64_t)2 * m * m * C * C * 2;
  L.wq = L.base + LAYERS * L.lsz;
  L.n = (L.wq + C + 1 + 3) / 4 * 4;
  return L;
}

}  // namespace

struct mcpm_fno {
  int n0 = 0, n1 = 0, m = 0;
  int64_t P = 0;
  Layout L;
  hipStream_t s = nullptr;
  hipStream_t pending = nullptr;  // a context stream with work on these buffers
  double in_shift = 0.0, in_scale = 1.0, out_shift = 0.0, out_scale = 1.0;
  int64_t adam_t = 0;
  float w0 = -1.f;  // loss weight of the potential term (fixed by the first snapshot)
  // device
  float *params = nullptr, *grad = nullptr, *am = nullptr, *av = nullptr;
  uint16_t *Wf = nullptr, *Wb = nullptr, *tf = nullptr, *tinv = nullptr;
  float2 *twy = nullptr;
  float *u = nullptr, *t = nullptr, *y = nullptr, *dy = nullptr;
  uint16_t *H[LAYERS] = {};
  float *Z[LAYERS] = {};
  float *dZ[2] = {};
  float *A = nullptr;
  float2 *X[LAYERS] = {};
  float2 *Y = nullptr, *dX = nullptr;
  uint16_t *Cp = nullptr;
  float *part = nullptr;
  double *lpart = nullptr, *dscr = nullptr;
  float *lhist = nullptr;
  int64_t lhist_cap = 0;
  // snapshots
  int32_t snap_cap = 0, snap_n = 0;
  float *su = nullptr, *st = nullptr;
  double *sD = nullptr;
  std::vector<void *> allocs;
};

namespace {

int falloc(mcpm_fno *f, void **p, size_t bytes) {
  FHIP(hipMalloc(p, bytes ? bytes : 256));
  f->allocs.push_back(*p);
  return MCPM_OK;
}

int nchunks(const mcpm_fno *f) { return (int)((f->P + CHUNK - 1) / CHUNK); }
int ncolsum(const mcpm_fno *f) { return (int)((f->P + COLSUM_NODES - 1) / COLSUM_NODES); }

// X = s(kx) DFT_y(A).
void dfty(mcpm_fno *f, hipStream_t s, float2 *X, float s0, float s1) {
  const int m = f->m;
  hipLaunchKernelGGL(k_dfty, dim3(m, 2 * m / 16, C / 16), dim3(256), 0, s, f->A, f->twy, X, f->n1,
                     m, s0, s1);
}

void idfty(mcpm_fno *f, hipStream_t s, const float2 *Y, float s0, float s1) {
  hipLaunchKernelGGL(k_idfty, dim3(f->n1 / 16, f->m / 8), dim3(256), 0, s, Y, f->twy, f->Cp, f->n1,
                     f->m, s0, s1);
}

// 64-node tiles per combine workgroup: up to 4, all in one row.
int combine_tiles(const mcpm_fno *f) {
  const int r = f->n0 / 64;
  return r % 4 == 0 ? 4 : (r % 2 == 0 ? 2 : 1);
}

template <int MODE, typename TIn>
void combine(const mcpm_fno *f, hipStream_t s, const TIn *in, const uint16_t *Wt, const float *bias,
             const float *wq, const float *Zprev, float *Z, uint16_t *H, float *yout) {
  const int ct = combine_tiles(f);
  const dim3 g((unsigned)(f->P / (64 * ct))), b(256);
  const int m2 = 2 * f->m;
  if (ct == 4)
    hipLaunchKernelGGL((k_combine<MODE, TIn, 4>), g, b, 0, s, in, Wt, f->tinv, f->Cp, bias, wq,
                       Zprev, Z, H, yout, f->n0, m2);
  else if (ct == 2)
    hipLaunchKernelGGL((k_combine<MODE, TIn, 2>), g, b, 0, s, in, Wt, f->tinv, f->Cp, bias, wq,
                       Zprev, Z, H, yout, f->n0, m2);
  else
    hipLaunchKernelGGL((k_combine<MODE, TIn, 1>), g, b, 0, s, in, Wt, f->tinv, f->Cp, bias, wq,
                       Zprev, Z, H, yout, f->n0, m2);
}

// D[0] = sum over the snapshot t of (|grad t|^2 + w0 t^2).
void tden(mcpm_fno *f, hipStream_t s, const float *t, double *D, float w0) {
  const unsigned nb = (unsigned)((f->P + 255) / 256);
  hipLaunchKernelGGL(k_tden, dim3(nb), dim3(256), 0, s, t, f->lpart, f->n0, f->n1, w0);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(1024), 0, s, f->lpart, (int)nb, D);
}

void settle(mcpm_fno *f) {
  if (f->pending) {
    (void)hipStreamSynchronize(f->pending);
    f->pending = nullptr;
  }
}

// One forward pass from u (normalised density).  train: keep H, Z and X of
// every layer for the backward pass (inference alternates H[0] and H[1]).
// The result is f->y.
int forward(mcpm_fno *f, hipStream_t s, bool train, const float *u) {
  const int m = f->m, m2 = 2 * m, n0 = f->n0, n1 = f->n1;
  const float sinv0 = (float)(1.0 / ((double)n0 * n1)), sinv1 = (float)(2.0 / ((double)n0 * n1));
  const int64_t P = f->P;
  auto hbuf = [&](int l) { return train ? f->H[l] : f->H[l & 1]; };
  hipLaunchKernelGGL(k_lift, dim3((unsigned)((P * (C / 8) + 255) / 256)), dim3(256), 0, s, u,
                     f->params + f->L.wl, hbuf(0), P);
  for (int l = 0; l < LAYERS; ++l) {
    const uint16_t *h = hbuf(l);
    float2 *Xl = train ? f->X[l] : f->X[0];
    hipLaunchKernelGGL((k_dftx<uint16_t>), dim3(n1), dim3(256), 0, s, h, f->tf, f->A, n0, m2);
    dfty(f, s, Xl, 1.0f, 1.0f);
    hipLaunchKernelGGL(k_mix, dim3(m2 * m / 4), dim3(256), 0, s, Xl,
                       (const float2 *)(f->params + f->L.R(l)), f->Y);
    idfty(f, s, f->Y, sinv0, sinv1);
    float *Zl = train ? f->Z[l] : nullptr;
    if (l < LAYERS - 1) {
      combine<CMB_FWD, uint16_t>(f, s, h, f->Wf + l * C * C, f->params + f->L.b(l), nullptr,
                                 nullptr, Zl, hbuf(l + 1), nullptr);
    } else {
      combine<CMB_LAST, uint16_t>(f, s, h, f->Wf + l * C * C, f->params + f->L.b(l),
                                  f->params + f->L.wq, nullptr, Zl, nullptr, f->y);
    }
  }
  FLAUNCH();
  return MCPM_OK;
}

// Backward pass after forward(train = true, u) and the loss (f->dy).
// Writes the full gradient into f->grad.
int backward(mcpm_fno *f, hipStream_t s, const float *u) {
  const int m = f->m, m2 = 2 * m, n0 = f->n0, n1 = f->n1;
  const int64_t P = f->P;
  const int nb = nchunks(f);
  const float sadj0 = (float)(1.0 / ((double)n0 * n1)), sadj1 = (float)(2.0 / ((double)n0 * n1));
  const unsigned gPC = (unsigned)((P * C + 255) / 256);
  // projection: wq, bq
  const int ncs = ncolsum(f);
  hipLaunchKernelGGL(k_colsum, dim3(ncs), dim3(256), 0, s, f->Z[LAYERS - 1], f->dy, f->part, P);
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, s, f->part, ncs, 2 * C + 1, 0, C,
                     f->grad + f->L.wq);
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, s, f->part, ncs, 2 * C + 1, 2 * C, 1,
                     f->grad + f->L.wq + C);
  hipLaunchKernelGGL(k_proj_bwd, dim3(gPC), dim3(256), 0, s, f->dy, f->params + f->L.wq, f->dZ[0],
                     P);
  int cur = 0;
  for (int l = LAYERS - 1; l >= 0; --l) {
    float *dz = f->dZ[cur], *dzn = f->dZ[cur ^ 1];
    hipLaunchKernelGGL(k_wgrad, dim3(nb), dim3(256), 0, s, dz, f->H[l], f->part, P);
    hipLaunchKernelGGL(k_reduce, dim3((C * C + C + 63) / 64), dim3(1024), 0, s, f->part, nb,
                       C * C + C, 0, C * C + C, f->grad + f->L.W(l));
    hipLaunchKernelGGL((k_dftx<float>), dim3(n1), dim3(256), 0, s, dz, f->tf, f->A, n0, m2);
    dfty(f, s, f->Y, sadj0, sadj1);
    hipLaunchKernelGGL(k_mix_dr, dim3(m2 * m / 4), dim3(256), 0, s, f->X[l], f->Y,
                       (float2 *)(f->grad + f->L.R(l)), 0);
    hipLaunchKernelGGL(k_mix_dx, dim3(m2 * m / 4), dim3(256), 0, s, f->Y,
                       (const float2 *)(f->params + f->L.R(l)), f->dX);
    idfty(f, s, f->dX, 1.0f, 1.0f);
    combine<CMB_BWD, float>(f, s, dz, f->Wb + l * C * C, nullptr, nullptr,
                            l > 0 ? f->Z[l - 1] : nullptr, dzn, nullptr, nullptr);
    cur ^= 1;
  }
  // lift: wl, bl from dH0 = dZ[cur]
  hipLaunchKernelGGL(k_colsum, dim3(ncs), dim3(256), 0, s, f->dZ[cur], u, f->part, P);
  hipLaunchKernelGGL(k_reduce, dim3(2), dim3(1024), 0, s, f->part, ncs, 2 * C + 1, 0, 2 * C,
                     f->grad + f->L.wl);
  FLAUNCH();
  return MCPM_OK;
}

int loss(mcpm_fno *f, hipStream_t s, const float *t, const double *D, float w0, float *out) {
  const unsigned nb = (unsigned)((f->P + 255) / 256);
  hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, s, f->y, t, D, f->dy, f->lpart, f->n0, f->n1,
                     w0);
  hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(1024), 0, s, f->lpart, (int)nb, D, out);
  FLAUNCH();
  return MCPM_OK;
}

int upload_params(mcpm_fno *f, const float *params) {
  FHIP(hipMemcpyAsync(f->params, params, f->L.n * sizeof(float), hipMemcpyHostToDevice, f->s));
  hipLaunchKernelGGL(k_cast, dim3((LAYERS * C * C + 255) / 256), dim3(256), 0, f->s, f->params,
                     f->L.W(0), f->L.lsz, f->Wf, f->Wb);
  FLAUNCH();
  FHIP(hipStreamSynchronize(f->s));
  return MCPM_OK;
}

float bf16_round(float v) {
  uint32_t u;
  memcpy(&u, &v, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  u &= 0xFFFF0000u;
  float r;
  memcpy(&r, &u, 4);
  return r;
}

uint16_t bf16_bits(float v) {
  uint32_t u;
  memcpy(&u, &v, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

}  // namespace

extern "C" {

int mcpm_fno_param_count(int32_t modes, int64_t *count) {
  if (!count || modes < 8 || modes % 8) return MCPM_ERR_ARG;
  *count = make_layout(modes).n;
  return MCPM_OK;
}

int mcpm_fno_create(int32_t n0, int32_t n1, int32_t modes, const float *params, mcpm_fno **out) {
  if (!out || !params) return MCPM_ERR_ARG;
  *out = nullptr;
  if (modes < 8 || modes % 8 || modes > 32) return MCPM_ERR_ARG;
  if (n0 < 64 || n0 % 64 || n1 < 2 * modes || n1 % 16 || 2 * modes > n0) return MCPM_ERR_ARG;
  mcpm_fno *f = new mcpm_fno;
  f->n0 = n0;
  f->n1 = n1;
  f->m = modes;
  f->P = (int64_t)n0 * n1;
  f->L = make_layout(modes);
  auto fail = [&](int st) {
    mcpm_fno_destroy(f);
    return st;
  };
  if (hipStreamCreate(&f->s) != hipSuccess) return fail(MCPM_ERR_HIP);
  const int m2 = 2 * modes;
  const int64_t P = f->P, np = f->L.n;
  int st = 0;
#define FA(ptr, bytes)                                        \
  if ((st = falloc(f, (void **)&(ptr), (size_t)(bytes)))) return fail(st);
  FA(f->params, np * 4);
  FA(f->grad, np * 4);
  FA(f->am, np * 4);
  FA(f->av, np * 4);
  FA(f->Wf, LAYERS * C * C * 2);
  FA(f->Wb, LAYERS * C * C * 2);
  FA(f->tf, (int64_t)m2 * n0 * 2);
  FA(f->tinv, (int64_t)m2 * n0 * 2);
  FA(f->twy, (int64_t)m2 * n1 * 8);
  FA(f->u, P * 4);
  FA(f->t, P * 4);
  FA(f->y, P * 4);
  FA(f->dy, P * 4);
  for (int l = 0; l < LAYERS; ++l) {
    FA(f->H[l], P * C * 2);
    FA(f->Z[l], P * C * 4);
    FA(f->X[l], (int64_t)m2 * modes * C * 8);
  }
  FA(f->dZ[0], P * C * 4);
  FA(f->dZ[1], P * C * 4);
  FA(f->A, (int64_t)n1 * m2 * C * 4);
  FA(f->Y, (int64_t)m2 * modes * C * 8);
  FA(f->dX, (int64_t)m2 * modes * C * 8);
  FA(f->Cp, (int64_t)n1 * C * m2 * 2);
  FA(f->part, std::max((int64_t)((P + CHUNK - 1) / CHUNK) * (C * C + C),
                        (int64_t)((P + COLSUM_NODES - 1) / COLSUM_NODES) * (2 * C + 1)) * 4);
  FA(f->lpart, ((P + 255) / 256) * 8);
  FA(f->dscr, 64);
#undef FA
  // tables: angles reduced exactly in integers, FP64 cos/sin, rounded
  std::vector<uint16_t> tf((size_t)m2 * n0), ti((size_t)m2 * n0);
  for (int k = 0; k < modes; ++k)
    for (int x = 0; x < n0; ++x) {
      const double ang = 2.0 * M_PI * (double)(((int64_t)k * x) % n0) / n0;
      const uint16_t c = bf16_bits((float)std::cos(ang)), sn = bf16_bits((float)(-std::sin(ang)));
      tf[(size_t)k * n0 + x] = c;
      tf[(size_t)(modes + k) * n0 + x] = sn;
      ti[(size_t)x * m2 + k] = c;
      ti[(size_t)x * m2 + modes + k] = sn;
    }
  std::vector<float2> tw((size_t)m2 * n1);
  for (int a = 0; a < m2; ++a) {
    const int64_t ky = a < modes ? a : n1 - m2 + a;
    for (int y = 0; y < n1; ++y) {
      const double ang = 2.0 * M_PI * (double)((ky * y) % n1) / n1;
      tw[(size_t)a * n1 + y] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    }
  }
  if (hipMemcpy(f->tf, tf.data(), tf.size() * 2, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(f->tinv, ti.data(), ti.size() * 2, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(f->twy, tw.data(), tw.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(f->am, 0, np * 4) != hipSuccess || hipMemset(f->av, 0, np * 4) != hipSuccess)
    return fail(MCPM_ERR_HIP);
  if ((st = upload_params(f, params))) return fail(st);
  *out = f;
  return MCPM_OK;
}

int mcpm_fno_destroy(mcpm_fno *f) {
  if (!f) return MCPM_OK;
  settle(f);
  if (f->s) (void)hipStreamSynchronize(f->s);
  for (void *p : f->allocs) (void)hipFree(p);
  if (f->s) (void)hipStreamDestroy(f->s);
  delete f;
  return MCPM_OK;
}

int mcpm_fno_set_params(mcpm_fno *f, const float *params) {
  if (!f || !params) return MCPM_ERR_ARG;
  settle(f);
  return upload_params(f, params);
}

int mcpm_fno_get_params(mcpm_fno *f, float *params) {
  if (!f || !params) return MCPM_ERR_ARG;
  settle(f);
  FHIP(hipMemcpyAsync(params, f->params, f->L.n * 4, hipMemcpyDeviceToHost, f->s));
  FHIP(hipStreamSynchronize(f->s));
  return MCPM_OK;
}

int mcpm_fno_set_norm(mcpm_fno *f, double in_shift, double in_scale, double out_shift,
                      double out_scale) {
  if (!f || !(in_scale != 0.0) || !(out_scale != 0.0)) return MCPM_ERR_ARG;
  f->in_shift = in_shift;
  f->in_scale = in_scale;
  f->out_shift = out_shift;
  f->out_scale = out_scale;
  return MCPM_OK;
}

int mcpm_fno_forward(mcpm_fno *f, const float *u, float *y) {
  if (!f || !u || !y) return MCPM_ERR_ARG;
  settle(f);
  FHIP(hipMemcpyAsync(f->u, u, f->P * 4, hipMemcpyHostToDevice, f->s));
  int st = forward(f, f->s, false, f->u);
  if (st) return st;
  FHIP(hipMemcpyAsync(y, f->y, f->P * 4, hipMemcpyDeviceToHost, f->s));
  FHIP(hipStreamSynchronize(f->s));
  return MCPM_OK;
}

int mcpm_fno_loss_grad(mcpm_fno *f, const float *u, const float *t, float w0, float *loss_out,
                       float *grad, float *y) {
  if (!f || !u || !t) return MCPM_ERR_ARG;
  settle(f);
  FHIP(hipMemcpyAsync(f->u, u, f->P * 4, hipMemcpyHostToDevice, f->s));
  FHIP(hipMemcpyAsync(f->t, t, f->P * 4, hipMemcpyHostToDevice, f->s));
  tden(f, f->s, f->t, f->dscr, w0);
  int st = forward(f, f->s, true, f->u);
  if (!st) st = loss(f, f->s, f->t, f->dscr, w0, (float *)(f->dscr + 1));
  if (!st) st = backward(f, f->s, f->u);
  if (st) return st;
  float lv = 0.f;
  FHIP(hipMemcpyAsync(&lv, f->dscr + 1, 4, hipMemcpyDeviceToHost, f->s));
  if (grad) FHIP(hipMemcpyAsync(grad, f->grad, f->L.n * 4, hipMemcpyDeviceToHost, f->s));
  if (y) FHIP(hipMemcpyAsync(y, f->y, f->P * 4, hipMemcpyDeviceToHost, f->s));
  FHIP(hipStreamSynchronize(f->s));
  if (loss_out) *loss_out = lv;
  return MCPM_OK;
}

int mcpm_fno_reserve_snapshots(mcpm_fno *f, int32_t cap) {
  if (!f || cap < 1) return MCPM_ERR_ARG;
  if (f->snap_cap) return MCPM_ERR_STATE;
  int st;
  if ((st = falloc(f, (void **)&f->su, (size_t)cap * f->P * 4))) return st;
  if ((st = falloc(f, (void **)&f->st, (size_t)cap * f->P * 4))) return st;
  if ((st = falloc(f, (void **)&f->sD, (size_t)cap * 8))) return st;
  f->snap_cap = cap;
  return MCPM_OK;
}

int mcpm_fno_add_snapshot_host(mcpm_fno *f, const float *u, const float *t, float w0) {
  if (!f || !u || !t || !(w0 >= 0.f)) return MCPM_ERR_ARG;
  if (f->snap_n && w0 != f->w0) return MCPM_ERR_ARG;
  if (f->snap_n >= f->snap_cap) return MCPM_ERR_CAPACITY;
  f->w0 = w0;
  settle(f);
  const size_t off = (size_t)f->snap_n * f->P;
  FHIP(hipMemcpyAsync(f->su + off, u, f->P * 4, hipMemcpyHostToDevice, f->s));
  FHIP(hipMemcpyAsync(f->st + off, t, f->P * 4, hipMemcpyHostToDevice, f->s));
  tden(f, f->s, f->st + off, f->sD + f->snap_n, w0);
  FLAUNCH();
  FHIP(hipStreamSynchronize(f->s));
  ++f->snap_n;
  return MCPM_OK;
}

int mcpm_fno_snapshot_count(mcpm_fno *f, int32_t *n) {
  if (!f || !n) return MCPM_ERR_ARG;
  *n = f->snap_n;
  return MCPM_OK;
}

int mcpm_fno_train(mcpm_fno *f, int32_t iters, float lr, uint64_t seed, int32_t n_snap,
                   float *losses) {
  if (!f || iters < 0 || !(lr > 0.f)) return MCPM_ERR_ARG;
  const int32_t ns = n_snap > 0 ? std::min(n_snap, f->snap_n) : f->snap_n;
  if (ns < 1) return MCPM_ERR_STATE;
  settle(f);
  if (iters > f->lhist_cap) {
    int st;
    if ((st = falloc(f, (void **)&f->lhist, (size_t)iters * 4))) return st;
    f->lhist_cap = iters;
  }
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const int64_t np = f->L.n;
  const float w0 = f->w0;  // the weight the snapshot denominators were made with
  uint64_t x = seed ^ 0x9E3779B97F4A7C15ull;
  for (int32_t it = 0; it < iters; ++it) {
    x += 0x9E3779B97F4A7C15ull;  // splitmix64 picks the snapshot
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const int32_t k = (int32_t)(z % (uint64_t)ns);
    const size_t off = (size_t)k * f->P;
    const float *u = f->su + off;
    int st = forward(f, f->s, true, u);
    if (!st) st = loss(f, f->s, f->st + off, f->sD + k, w0, f->lhist + it);
    if (!st) st = backward(f, f->s, u);
    if (st) return st;
    ++f->adam_t;
    const float bc1 = (float)(1.0 - std::pow((double)b1, (double)f->adam_t));
    const float bc2 = (float)(1.0 - std::pow((double)b2, (double)f->adam_t));
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, f->s, f->params,
                       f->grad, f->am, f->av, np, lr, b1, b2, eps, bc1, bc2);
    hipLaunchKernelGGL(k_cast, dim3((LAYERS * C * C + 255) / 256), dim3(256), 0, f->s, f->params,
                       f->L.W(0), f->L.lsz, f->Wf, f->Wb);
    FLAUNCH();
  }
  if (losses && iters)
    FHIP(hipMemcpyAsync(losses, f->lhist, (size_t)iters * 4, hipMemcpyDeviceToHost, f->s));
  FHIP(hipStreamSynchronize(f->s));
  return MCPM_OK;
}

}  // extern "C"

// ---- the surrogate on a plane context ----

int fno_check_plane(const mcpm_fno *f, const FemParams &fem) {
  if (!fem.per_x || !fem.per_y || fem.nx != f->n0 || fem.ny != f->n1) return MCPM_ERR_ARG;
  return MCPM_OK;
}

// Network field of the context's current nodal density, written as the
// pre-core FEM solution into the context's potential array.
int fno_field_guess(mcpm_fno *f, mcpm_ctx *c) {
  PlaneField pf{};
  int st = ctx_plane_nodes(c, &pf);
  if (st) return st;
  if ((st = fno_check_plane(f, pf.fem))) return st;
  settle(f);
  FHIP(hipStreamSynchronize(f->s));
  const hipStream_t s = pf.s;
  hipLaunchKernelGGL(k_from_nodes, dim3((unsigned)((f->P + 255) / 256)), dim3(256), 0, s,
                     pf.n_node, f->u, f->n0, f->n1, f->in_shift, f->in_scale);
  FLAUNCH();
  if ((st = forward(f, s, false, f->u))) return st;
  const int64_t nn = (int64_t)(f->n0 + 1) * (f->n1 + 1);
  hipLaunchKernelGGL(k_to_nodes, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, f->y,
                     pf.pot, f->n0, f->n1, f->out_shift, 1.0 / f->out_scale);
  FLAUNCH();
  f->pending = s;
  return MCPM_OK;
}

extern "C" {

int mcpm_fno_add_snapshot(mcpm_fno *f, mcpm_ctx *c, float w0) {
  if (!f || !c || !(w0 >= 0.f)) return MCPM_ERR_ARG;
  if (f->snap_n && w0 != f->w0) return MCPM_ERR_ARG;
  if (f->snap_n >= f->snap_cap) return MCPM_ERR_CAPACITY;
  PlaneField pf{};
  int st = ctx_plane_nodes(c, &pf);
  if (st) return st;
  if ((st = fno_check_plane(f, pf.fem))) return st;
  int iters = 0;
  double res0 = 0.0;
  int solved = 0;
  if ((st = ctx_plane_cg(c, 0, 0, 0.0, &res0, &iters, &solved))) return st;
  settle(f);
  f->w0 = w0;
  const hipStream_t s = pf.s;
  const size_t off = (size_t)f->snap_n * f->P;
  const unsigned g = (unsigned)((f->P + 255) / 256);
  hipLaunchKernelGGL(k_from_nodes, dim3(g), dim3(256), 0, s, pf.n_node, f->su + off, f->n0, f->n1,
                     f->in_shift, f->in_scale);
  hipLaunchKernelGGL(k_from_nodes, dim3(g), dim3(256), 0, s, pf.pot, f->st + off, f->n0, f->n1,
                     f->out_shift, f->out_scale);
  tden(f, s, f->st + off, f->sD + f->snap_n, w0);
  FLAUNCH();
  if ((st = ctx_plane_finish(c))) return st;
  f->pending = s;
  ++f->snap_n;
  return MCPM_OK;
}

int mcpm_fno_field(mcpm_fno *f, mcpm_ctx *c, double accept_residual, int32_t exact,
                   mcpm_fno_field_stats *stats) {
  if (!f || !c) return MCPM_ERR_ARG;
  int st = fno_field_guess(f, c);
  if (st) return st;
  double res0 = 0.0;
  int iters = 0;
  // exact: solve from the network's field; accept_residual >= 0: keep the
  // network's field when its residual is at most that, else solve; < 0: keep it
  const int mode = exact ? 0 : (accept_residual >= 0.0 ? 1 : 2);
  int solved = 0;
  if ((st = ctx_plane_cg(c, 1, mode, accept_residual, &res0, &iters, &solved))) return st;
  if ((st = ctx_plane_finish(c))) return st;
  if (stats) {
    stats->residual = res0;
    stats->cg_iterations = iters;
    stats->used_network = solved ? 0 : 1;
  }
  return MCPM_OK;
}

}  // extern "C"

// The field of step `step` under mcpm_set_field_surrogate: every `every`
// steps (step % every == 0; never when every is 0) the exact solve runs,
// started from the network's field; otherwise the network's field is kept
// when its residual is at most `accept` (no check when accept < 0).
int fno_field_step(mcpm_fno *f, mcpm_ctx *c, int64_t step, double accept, int64_t every) {
  mcpm_fno_field_stats stx{};
  const int exact = every > 0 && step % every == 0;
  int st = mcpm_fno_field(f, c, accept, exact, &stx);
  if (st) return st;
  return ctx_set_surrogate_stats(c, &stx);
}
