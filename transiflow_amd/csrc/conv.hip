// Convolutions of the learned preconditioner on the cell grid (BF16 MFMA,
// FP32 accumulation): forward, backward-data, backward-weights, plus the
// elementwise pieces of training (MSE gradient, ReLU backward, Adam).
//
// Layouts.  Activations are cell-major [ncell][C] (the state vector layout:
// index c + C * cell, cell = i + nx (j + ny k)), BF16 stored as uint16.  A
// layer has taps = 3^dim neighbour offsets, tap = (dx+1) + 3 (dy+1) [+ 9 (dz+1)
// in 3D; dz = 0 in 2D], and zero padding outside the grid:
//
//   out[cell][co] = b[co] + sum_{tap, ci} W[tap][ci][co] in[cell + off(tap)][ci]
//
// which is a GEMM with K = taps * Cin (row k = tap * Cin + ci).  The MFMA
// copy of W is [Kpad][Npad] BF16 with K and Cout padded to multiples of 16.
//
// MFMA v_mfma_f32_16x16x16bf16_1k (one wave, lane l, measured on gfx950):
//   a[i] = A[l % 16][4 (l / 16) + i], b[i] = B[4 (l / 16) + i][l % 16],
//   D[4 (l / 16) + v][l % 16].
#include "tf_common.h"

#include <cstdint>

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((unsigned)h << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

struct Grid3 {
  int nx, ny, nz, dim;
};

// in-grid neighbour of cell (x, y, z) through tap, or -1
__device__ __forceinline__ long neighbour(const Grid3& g, int x, int y, int z, int tap) {
  const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = g.dim == 3 ? tap / 9 - 1 : 0;
  const int X = x + dx, Y = y + dy, Z = z + dz;
  if (X < 0 || X >= g.nx || Y < 0 || Y >= g.ny || Z < 0 || Z >= g.nz) return -1;
  return (long)X + (long)g.nx * (Y + (long)g.ny * Z);
}

// One wave: 16 cells x all output tiles.  Block: 4 waves = 64 cells.
// W (MFMA layout, BF16) is staged in LDS; A fragments are gathered from the
// neighbours' activations.
constexpr int CONV_LDS = 28 * 1024;  // uint16 elements: K x N up to 864 x 32

__global__ __launch_bounds__(256) void k_conv_fwd(Grid3 g, long ncell, int cin, int cout, int kpad, int npad,
                                                  const uint16_t* __restrict__ in, const uint16_t* __restrict__ W,
                                                  const float* __restrict__ bias, int relu, float* __restrict__ out32,
                                                  uint16_t* __restrict__ out16) {
  __shared__ uint16_t Ws[CONV_LDS];
  const int K = (g.dim == 3 ? 27 : 9) * cin;
  for (int e = threadIdx.x; e < kpad * npad; e += 256) Ws[e] = W[e];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long cell0 = ((long)blockIdx.x * 4 + wave) * 16;
  const long cell = cell0 + (lane & 15);
  const int q = lane >> 4;
  int x = 0, y = 0, z = 0;
  const bool live = cell < ncell;
  if (live) {
    x = (int)(cell % g.nx);
    y = (int)((cell / g.nx) % g.ny);
    z = (int)(cell / ((long)g.nx * g.ny));
  }
  const int ntiles = npad / 16;
  f32x4 acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0};
  for (int k0 = 0; k0 < kpad; k0 += 16) {
    bf16x4 a;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int k = k0 + 4 * q + i;
      uint16_t v = 0;
      if (live && k < K) {
        const int tap = k / cin, ci = k - tap * cin;
        const long nb = neighbour(g, x, y, z, tap);
        if (nb >= 0) v = in[nb * cin + ci];
      }
      a[i] = (short)v;
    }
    bf16x4 b0, b1;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      b0[i] = (short)Ws[(k0 + 4 * q + i) * npad + (lane & 15)];
      b1[i] = ntiles > 1 ? (short)Ws[(k0 + 4 * q + i) * npad + 16 + (lane & 15)] : (short)0;
    }
    acc0 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b0, acc0, 0, 0, 0);
    if (ntiles > 1) acc1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b1, acc1, 0, 0, 0);
  }
  // D[4 q + v][lane % 16]: row = cell0 + 4 q + v, column = output channel
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    if (t >= ntiles) break;
    const int co = t * 16 + (lane & 15);
    if (co >= cout) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const long c = cell0 + 4 * q + v;
      if (c >= ncell) continue;
      float s = (t == 0 ? acc0[v] : acc1[v]) + (bias ? bias[co] : 0.f);
      if (relu) s = s > 0.f ? s : 0.f;
      if (out32) out32[c * cout + co] = s;
      if (out16) out16[c * cout + co] = f2bf(s);
    }
  }
}

// Partial weight gradients over a chunk of cells per block:
//   dWp[block][k][co] = sum_{cell in chunk} patch[cell][k] g[cell][co]
// MFMA with A = patch^T (k x cells), B = g (cells x co).  One wave per
// (k-tile, n-tile) pair, waves stride over the pairs.
constexpr int WG_CHUNK = 256;

__global__ __launch_bounds__(256) void k_conv_wgrad(Grid3 g, long ncell, int cin, int cout, int kpad, int npad,
                                                    const uint16_t* __restrict__ in, const uint16_t* __restrict__ gout,
                                                    float* __restrict__ partial) {
  const int K = (g.dim == 3 ? 27 : 9) * cin;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int q = lane >> 4, r = lane & 15;
  const long c0 = (long)blockIdx.x * WG_CHUNK;
  const int ktiles = kpad / 16, ntiles = npad / 16;
  float* P = partial + (long)blockIdx.x * kpad * npad;
  for (int tile = wave; tile < ktiles * ntiles; tile += 4) {
    const int kt = tile / ntiles, nt = tile % ntiles;
    const int k = kt * 16 + r;  // this lane's A row
    int tap = 0, ci = 0;
    if (k < K) {
      tap = k / cin;
      ci = k - tap * cin;
    }
    const int co = nt * 16 + r;  // this lane's B column
    f32x4 acc = {0, 0, 0, 0};
    for (int cc = 0; cc < WG_CHUNK; cc += 16) {
      bf16x4 a, b;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const long cell = c0 + cc + 4 * q + i;
        uint16_t av = 0, bv = 0;
        if (cell < ncell) {
          if (k < K) {
            const int x = (int)(cell % g.nx), y = (int)((cell / g.nx) % g.ny), z = (int)(cell / ((long)g.nx * g.ny));
            const long nb = neighbour(g, x, y, z, tap);
            if (nb >= 0) av = in[nb * cin + ci];
          }
          if (co < cout) bv = gout[cell * cout + co];
        }
        a[i] = (short)av;
        b[i] = (short)bv;
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) P[(long)(kt * 16 + 4 * q + v) * npad + nt * 16 + r] = acc[v];
  }
}

// out[e] = sum_b partial[b][e]  (fixed order: deterministic)
__global__ __launch_bounds__(256) void k_sum_partials(int nb, long len, const float* __restrict__ partial,
                                                      float* __restrict__ out) {
  const long e = blockIdx.x * 256L + threadIdx.x;
  if (e >= len) return;
  float s = 0.f;
  for (int b = 0; b < nb; ++b) s += partial[(long)b * len + e];
  out[e] = s;
}

// bias partials: bp[block][co] = sum over the block's chunk of g[cell][co] (FP32 gradient)
__global__ __launch_bounds__(256) void k_bias_partial(long ncell, int cout, const float* __restrict__ g,
                                                      float* __restrict__ bp) {
  __shared__ float sh[256];
  const long c0 = (long)blockIdx.x * WG_CHUNK;
  for (int co = 0; co < cout; ++co) {
    const long cell = c0 + threadIdx.x;
    sh[threadIdx.x] = cell < ncell ? g[cell * cout + co] : 0.f;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) bp[(long)blockIdx.x * cout + co] = sh[0];
    __syncthreads();
  }
}

// g = 2 (y - t) / count ; partial[block] = sum (y - t)^2 over the block
__global__ __launch_bounds__(256) void k_mse(long n, const float* __restrict__ y, const float* __restrict__ t,
                                             float inv_count, float* __restrict__ g, float* __restrict__ partial) {
  __shared__ float sh[256];
  const long e = blockIdx.x * 256L + threadIdx.x;
  float d = 0.f;
  if (e < n) {
    d = y[e] - t[e];
    g[e] = 2.f * d * inv_count;
  }
  sh[threadIdx.x] = d * d;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// gradient through ReLU given the layer's (post-ReLU) output act: g *= (act > 0);
// writes FP32 (bias gradient) and BF16 (GEMM operand) copies
__global__ __launch_bounds__(256) void k_relu_bwd(long n, const float* __restrict__ gin, const uint16_t* __restrict__ act,
                                                  int relu, float* __restrict__ g32, uint16_t* __restrict__ g16) {
  const long e = blockIdx.x * 256L + threadIdx.x;
  if (e >= n) return;
  float v = gin[e];
  if (relu && !(bf2f(act[e]) > 0.f)) v = 0.f;
  g32[e] = v;
  g16[e] = f2bf(v);
}

// FP32 master W [kpad][npad] (padded MFMA layout, zero padding) -> BF16 forward
// copy Wf [kpad][npad] and the backward-data copy Wt [kpad_t][npad_t] with
// Wt[tap' * cout + co][ci] = W[(taps-1-tap') * cin + ci][co] (flipped taps,
// transposed channels).
__global__ __launch_bounds__(256) void k_pack(int taps, int cin, int cout, int kpad, int npad, int kpad_t, int npad_t,
                                              const float* __restrict__ W, uint16_t* __restrict__ Wf,
                                              uint16_t* __restrict__ Wt) {
  const long e = blockIdx.x * 256L + threadIdx.x;
  if (e < (long)kpad * npad) {
    const int k = (int)(e / npad), co = (int)(e % npad);
    Wf[e] = (k < taps * cin && co < cout) ? f2bf(W[e]) : (uint16_t)0;
  }
  if (Wt && e < (long)kpad_t * npad_t) {
    const int k = (int)(e / npad_t), ci = (int)(e % npad_t);
    uint16_t v = 0;
    if (k < taps * cout && ci < cin) {
      const int tp = k / cout, co = k - tp * cout;
      v = f2bf(W[(long)((taps - 1 - tp) * cin + ci) * npad + co]);
    }
    Wt[e] = v;
  }
}

__global__ __launch_bounds__(256) void k_adam(long n, float* __restrict__ w, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                                              float b2, float eps, float c1, float c2) {
  const long e = blockIdx.x * 256L + threadIdx.x;
  if (e >= n) return;
  const float gi = g[e];
  const float mi = b1 * m[e] + (1.f - b1) * gi;
  const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
  m[e] = mi;
  v[e] = vi;
  w[e] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ __launch_bounds__(256) void k_to_bf16(long n, const float* __restrict__ x, float scale,
                                                 uint16_t* __restrict__ y) {
  const long e = blockIdx.x * 256L + threadIdx.x;
  if (e < n) y[e] = f2bf(x[e] * scale);
}

int check_layer(int dim, int cin, int cout, int kpad, int npad) {
  const int taps = dim == 3 ? 27 : 9;
  if (dim != 2 && dim != 3) return tf::fail("conv: dim must be 2 or 3");
  if (cin < 1 || cout < 1 || cout > 32) return tf::fail("conv: 1 <= cout <= 32 required");
  if (kpad % 16 || npad % 16 || kpad < taps * cin || npad < cout || npad > 32) return tf::fail("conv: bad padding");
  if ((long)kpad * npad > CONV_LDS) return tf::fail("conv: weights exceed the LDS stage");
  return 0;
}

}  // namespace

extern "C" {

// Forward of one layer.  in: BF16 [ncell][cin]; W: BF16 [kpad][npad]; bias FP32
// (may be null); out32 / out16 optional FP32 / BF16 [ncell][cout].
int tf_conv_fwd(int nx, int ny, int nz, int dim, int cin, int cout, int kpad, int npad, const void* in,
                const void* W, const float* bias, int relu, float* out32, void* out16, void* stream) {
  if (int e = check_layer(dim, cin, cout, kpad, npad)) return e;
  const long ncell = (long)nx * ny * nz;
  Grid3 g{nx, ny, nz, dim};
  hipLaunchKernelGGL(k_conv_fwd, dim3((unsigned)((ncell + 63) / 64)), dim3(256), 0, (hipStream_t)stream, g, ncell,
                     cin, cout, kpad, npad, (const uint16_t*)in, (const uint16_t*)W, bias, relu, out32,
                     (uint16_t*)out16);
  TF_PROF_WORK(tf::PROF_CONV, 2.0 * ncell * (dim == 3 ? 27 : 9) * cin * cout);
  TF_LAUNCH_CHECK("k_conv_fwd");
  return 0;
}

int tf_conv_wgrad_work_size(long ncell, int kpad, int npad) {
  const long blocks = (ncell + WG_CHUNK - 1) / WG_CHUNK;
  const long n = blocks * kpad * npad + blocks * 32;
  return n > 2147483647L ? -1 : (int)n;
}

// dW [kpad][npad] FP32 = sum over cells of patch(in)^T gout; db[cout] = sum of g32.
// in: BF16 [ncell][cin]; gout16 / g32: BF16 / FP32 [ncell][cout]; work: FP32 of
// tf_conv_wgrad_work_size() elements.
int tf_conv_wgrad(int nx, int ny, int nz, int dim, int cin, int cout, int kpad, int npad, const void* in,
                  const void* gout16, const float* g32, float* work, float* dW, float* db, void* stream) {
  if (int e = check_layer(dim, cin, cout, kpad, npad)) return e;
  const long ncell = (long)nx * ny * nz;
  const long blocks = (ncell + WG_CHUNK - 1) / WG_CHUNK;
  hipStream_t st = (hipStream_t)stream;
  Grid3 g{nx, ny, nz, dim};
  hipLaunchKernelGGL(k_conv_wgrad, dim3((unsigned)blocks), dim3(256), 0, st, g, ncell, cin, cout, kpad, npad,
                     (const uint16_t*)in, (const uint16_t*)gout16, work);
  const long len = (long)kpad * npad;
  hipLaunchKernelGGL(k_sum_partials, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st, (int)blocks, len, work,
                     dW);
  float* bp = work + blocks * len;
  if (db) {
    hipLaunchKernelGGL(k_bias_partial, dim3((unsigned)blocks), dim3(256), 0, st, ncell, cout, g32, bp);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, (int)blocks, (long)cout, bp, db);
  }
  TF_PROF_WORK(tf::PROF_CONV, 2.0 * ncell * (dim == 3 ? 27 : 9) * cin * cout);
  TF_LAUNCH_CHECK("conv wgrad");
  return 0;
}

// Loss work: partial sums per 256 elements -> loss_partial[(n+255)/256].
int tf_mse_grad(long n, const float* y, const float* t, float* g, float* loss_partial, void* stream) {
  hipLaunchKernelGGL(k_mse, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, n, y, t,
                     1.f / (float)n, g, loss_partial);
  TF_LAUNCH_CHECK("k_mse");
  return 0;
}

int tf_relu_bwd(long n, const float* gin, const void* act16, int relu, float* g32, void* g16, void* stream) {
  hipLaunchKernelGGL(k_relu_bwd, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, n, gin,
                     (const uint16_t*)act16, relu, g32, (uint16_t*)g16);
  TF_LAUNCH_CHECK("k_relu_bwd");
  return 0;
}

int tf_conv_pack(int dim, int cin, int cout, int kpad, int npad, int kpad_t, int npad_t, const float* W, void* Wf,
                 void* Wt, void* stream) {
  const int taps = dim == 3 ? 27 : 9;
  const long n = (long)kpad * npad > (long)kpad_t * npad_t ? (long)kpad * npad : (long)kpad_t * npad_t;
  hipLaunchKernelGGL(k_pack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, taps, cin, cout,
                     kpad, npad, kpad_t, npad_t, W, (uint16_t*)Wf, (uint16_t*)Wt);
  TF_LAUNCH_CHECK("k_pack");
  return 0;
}

int tf_adam(long n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2, float eps, int step,
            void* stream) {
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, n, w, g, m, v, lr,
                     b1, b2, eps, c1, c2);
  TF_LAUNCH_CHECK("k_adam");
  return 0;
}

int tf_to_bf16(long n, const float* x, float scale, void* y, void* stream) {
  hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, n, x, scale,
                     (uint16_t*)y);
  TF_LAUNCH_CHECK("k_to_bf16");
  return 0;
}

}  // extern "C"
