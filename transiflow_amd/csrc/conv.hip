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

// The same layer when cin is a multiple of 4 (every layer of the 3D network and
// its data gradients): a lane's four K entries of one MFMA step are four
// consecutive channels of one tap, so the A fragment is one 8-byte load of the
// neighbour's activations; W is staged transposed ([n][k]) so the B fragment
// is one 8-byte LDS read.  Each wave computes CF_MT tiles of 16 cells, reusing
// each B fragment CF_MT times, and a block stages W once for 64 CF_MT cells.
// The K order of the accumulation is that of k_conv_fwd.
constexpr int CF_MT = 4;

template <int NT>
__global__ __launch_bounds__(256) void k_conv_fwd_v(Grid3 g, long ncell, int cin, int cout, int kpad, int npad,
                                                    const uint16_t* __restrict__ in, const uint16_t* __restrict__ W,
                                                    const float* __restrict__ bias, int relu,
                                                    float* __restrict__ out32, uint16_t* __restrict__ out16) {
  extern __shared__ __align__(16) uint16_t Ws[];   // kpad * npad entries (dynamic: small layers keep occupancy)
  for (int e = threadIdx.x; e < kpad * npad; e += 256) {
    const int k = e / npad, n = e - k * npad;
    Ws[n * kpad + k] = W[e];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int q = lane >> 4, r = lane & 15;
  const int K = (g.dim == 3 ? 27 : 9) * cin;
  const long base = ((long)blockIdx.x * 4 + wave) * (CF_MT * 16);
  int xm[CF_MT], ym[CF_MT], zm[CF_MT];
  bool livem[CF_MT];
#pragma unroll
  for (int m = 0; m < CF_MT; ++m) {
    const long cell = base + m * 16 + r;
    livem[m] = cell < ncell;
    const long c = livem[m] ? cell : 0;
    xm[m] = (int)(c % g.nx);
    ym[m] = (int)((c / g.nx) % g.ny);
    zm[m] = (int)(c / ((long)g.nx * g.ny));
  }
  f32x4 acc[CF_MT][NT];
#pragma unroll
  for (int m = 0; m < CF_MT; ++m)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[m][t] = f32x4{0, 0, 0, 0};
  for (int k0 = 0; k0 < kpad; k0 += 16) {
    const int k = k0 + 4 * q;  // this lane's four K entries: channels ci..ci+3 of one tap
    const bool kin = k < K;
    const int tap = kin ? k / cin : 0;
    const int ci = k - tap * cin;
    const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = g.dim == 3 ? tap / 9 - 1 : 0;
    bf16x4 b[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) b[t] = *reinterpret_cast<const bf16x4*>(&Ws[(t * 16 + r) * kpad + k]);
    bf16x4 a[CF_MT];
#pragma unroll
    for (int m = 0; m < CF_MT; ++m) {
      a[m] = bf16x4{0, 0, 0, 0};
      const int X = xm[m] + dx, Y = ym[m] + dy, Z = zm[m] + dz;
      if (kin && livem[m] && X >= 0 && X < g.nx && Y >= 0 && Y < g.ny && Z >= 0 && Z < g.nz) {
        const long nb = (long)X + (long)g.nx * (Y + (long)g.ny * Z);
        a[m] = *reinterpret_cast<const bf16x4*>(in + nb * cin + ci);
      }
    }
#pragma unroll
    for (int m = 0; m < CF_MT; ++m)
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[m], b[t], acc[m][t], 0, 0, 0);
  }
  // D[4 q + v][r]: cell base + 16 m + 4 q + v, output channel 16 t + r
#pragma unroll
  for (int m = 0; m < CF_MT; ++m)
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int co = t * 16 + r;
      if (co >= cout) continue;
      const float bco = bias ? bias[co] : 0.f;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const long c = base + m * 16 + 4 * q + v;
        if (c >= ncell) continue;
        float sv = acc[m][t][v] + bco;
        if (relu) sv = sv > 0.f ? sv : 0.f;
        if (out32) out32[c * cout + co] = sv;
        if (out16) out16[c * cout + co] = f2bf(sv);
      }
    }
}

// Partial weight gradients, one block per WG_CHUNK cells:
//   partial[block][k][co] = sum over the block's cells of patch[cell][k] g[cell][co]
// as MFMAs with M = k, N = co and the cells as the reduction dimension.  The
// block walks its cells WG_SUB at a time: the patch (all K values of WG_SUB
// cells, zero outside the grid) and the gradient are staged transposed in LDS
// ([k][cell], [co][cell]), so a lane's four reduction entries are one 8-byte
// LDS read for A and for B.  The staging reads each (cell, tap) neighbour's
// channels in 8- or 16-byte pieces.  Wave w of the WG_WAVES owns the output
// tiles w, w + WG_WAVES, ... and keeps their accumulators in registers across
// the whole chunk, so a block writes its partial once.
constexpr int WG_CHUNK = 1024;
constexpr int WG_SUB = 32;
constexpr int WG_WAVES = 8;
constexpr int WG_MAXT = 14;   // output tiles per wave: ceil(56 k-tiles x 2 n-tiles / 8)
constexpr int WG_KMAX = 896;  // staged K rows

template <int VEC>   // channels per staging load: 8 (16 bytes), 4 (8 bytes) or 1
__global__ __launch_bounds__(64 * WG_WAVES) void k_conv_wgrad(Grid3 g, long ncell, int cin, int cout, int kpad,
                                                              int npad, const uint16_t* __restrict__ in,
                                                              const uint16_t* __restrict__ gout,
                                                              float* __restrict__ partial) {
  __shared__ __align__(16) uint16_t pt[WG_KMAX * WG_SUB];   // [kpad][WG_SUB]
  __shared__ __align__(16) uint16_t gt[32 * WG_SUB];        // [npad][WG_SUB]
  constexpr int NT = 64 * WG_WAVES;
  const int taps = g.dim == 3 ? 27 : 9;
  const int K = taps * cin;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int q = lane >> 4, r = lane & 15;
  const int ktiles = kpad / 16, ntiles = npad / 16, ntot = ktiles * ntiles;
  const int groups = cin / VEC;                 // channel pieces per (cell, tap)
  const int staged_taps = (kpad + cin - 1) / cin;
  const long c0 = (long)blockIdx.x * WG_CHUNK;
  // K padding rows beyond the last staged tap stay zero for the whole chunk
  for (int e = staged_taps * cin * WG_SUB + threadIdx.x; e < kpad * WG_SUB; e += NT) pt[e] = 0;
  f32x4 acc[WG_MAXT];
#pragma unroll
  for (int t = 0; t < WG_MAXT; ++t) acc[t] = f32x4{0, 0, 0, 0};
  for (int s0 = 0; s0 < WG_CHUNK && c0 + s0 < ncell; s0 += WG_SUB) {
    // (piece, cell, tap) items: consecutive threads take consecutive pieces of one neighbour
    for (int p = threadIdx.x; p < WG_SUB * staged_taps * groups; p += NT) {
      const int gi = p % groups, rest = p / groups;
      const int cl = rest % WG_SUB, tap = rest / WG_SUB;
      const long cell = c0 + s0 + cl;
      long nb = -1;
      if (tap < taps && cell < ncell) {
        const int ic = (int)cell;   // ncell < 2^31 (checked by the caller)
        const int x = ic % g.nx, y = (ic / g.nx) % g.ny, z = ic / (g.nx * g.ny);
        const int X = x + tap % 3 - 1, Y = y + (tap / 3) % 3 - 1, Z = z + (g.dim == 3 ? tap / 9 - 1 : 0);
        if (X >= 0 && X < g.nx && Y >= 0 && Y < g.ny && Z >= 0 && Z < g.nz)
          nb = (long)X + (long)g.nx * (Y + (long)g.ny * Z);
      }
      const int k0 = tap * cin + gi * VEC;
      uint16_t v[VEC];
      if (nb >= 0) {
        const uint16_t* src = in + nb * cin + gi * VEC;
        if constexpr (VEC == 8) {
          const uint4 w = *reinterpret_cast<const uint4*>(src);
          const unsigned u[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            v[2 * j] = (uint16_t)(u[j] & 0xffffu);
            v[2 * j + 1] = (uint16_t)(u[j] >> 16);
          }
        } else if constexpr (VEC == 4) {
          const uint2 w = *reinterpret_cast<const uint2*>(src);
          v[0] = (uint16_t)(w.x & 0xffffu);
          v[1] = (uint16_t)(w.x >> 16);
          v[2] = (uint16_t)(w.y & 0xffffu);
          v[3] = (uint16_t)(w.y >> 16);
        } else {
          v[0] = src[0];
        }
      } else {
#pragma unroll
        for (int j = 0; j < VEC; ++j) v[j] = 0;
      }
#pragma unroll
      for (int j = 0; j < VEC; ++j) {
        const int k = k0 + j;
        if (k < kpad) pt[k * WG_SUB + cl] = k < K ? v[j] : (uint16_t)0;
      }
    }
    for (int e = threadIdx.x; e < npad * WG_SUB; e += NT) {
      const int co = e / WG_SUB, cl = e % WG_SUB;
      const long cell = c0 + s0 + cl;
      gt[co * WG_SUB + cl] = (co < cout && cell < ncell) ? gout[cell * cout + co] : (uint16_t)0;
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < WG_MAXT; ++t) {
      const int tile = wave + WG_WAVES * t;
      if (tile < ntot) {
        const int kt = tile / ntiles, nt = tile - kt * ntiles;
#pragma unroll
        for (int h = 0; h < WG_SUB; h += 16) {
          const bf16x4 a = *reinterpret_cast<const bf16x4*>(&pt[(kt * 16 + r) * WG_SUB + h + 4 * q]);
          const bf16x4 b = *reinterpret_cast<const bf16x4*>(&gt[(nt * 16 + r) * WG_SUB + h + 4 * q]);
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
  float* P = partial + (long)blockIdx.x * kpad * npad;
#pragma unroll
  for (int t = 0; t < WG_MAXT; ++t) {
    const int tile = wave + WG_WAVES * t;
    if (tile < ntot) {
      const int kt = tile / ntiles, nt = tile - kt * ntiles;
#pragma unroll
      for (int v = 0; v < 4; ++v) P[(long)(kt * 16 + 4 * q + v) * npad + nt * 16 + r] = acc[t][v];
    }
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

// bias partials: bp[block][co] = sum over the block's WG_CHUNK cells of g[cell][co] (FP32 gradient)
__global__ __launch_bounds__(256) void k_bias_partial(long ncell, int cout, const float* __restrict__ g,
                                                      float* __restrict__ bp) {
  __shared__ float sh[256];
  const long c0 = (long)blockIdx.x * WG_CHUNK;
  for (int co = 0; co < cout; ++co) {
    float acc = 0.f;
    for (int cl = threadIdx.x; cl < WG_CHUNK; cl += 256) {
      const long cell = c0 + cl;
      if (cell < ncell) acc += g[cell * cout + co];
    }
    sh[threadIdx.x] = acc;
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
  hipStream_t st = (hipStream_t)stream;
  const uint16_t* in16 = (const uint16_t*)in;
  const uint16_t* W16 = (const uint16_t*)W;
  if (cin % 4 == 0 && ((uintptr_t)in % 8) == 0) {
    const dim3 grid((unsigned)((ncell + 64 * CF_MT - 1) / (64 * CF_MT)));
    const size_t lds = sizeof(uint16_t) * kpad * npad;
    if (npad == 16)
      hipLaunchKernelGGL(k_conv_fwd_v<1>, grid, dim3(256), lds, st, g, ncell, cin, cout, kpad, npad, in16, W16, bias,
                         relu, out32, (uint16_t*)out16);
    else
      hipLaunchKernelGGL(k_conv_fwd_v<2>, grid, dim3(256), lds, st, g, ncell, cin, cout, kpad, npad, in16, W16, bias,
                         relu, out32, (uint16_t*)out16);
  } else {
    hipLaunchKernelGGL(k_conv_fwd, dim3((unsigned)((ncell + 63) / 64)), dim3(256), 0, st, g, ncell, cin, cout, kpad,
                       npad, in16, W16, bias, relu, out32, (uint16_t*)out16);
  }
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
  if (kpad > WG_KMAX || ncell >= (1L << 31)) return tf::fail("conv wgrad: K or the grid too large");
  const long blocks = (ncell + WG_CHUNK - 1) / WG_CHUNK;
  hipStream_t st = (hipStream_t)stream;
  Grid3 g{nx, ny, nz, dim};
  const dim3 wg_grid((unsigned)blocks), wg_block(64 * WG_WAVES);
  const uint16_t* in16 = (const uint16_t*)in;
  const uint16_t* g16 = (const uint16_t*)gout16;
  if (cin % 8 == 0 && ((uintptr_t)in % 16) == 0)
    hipLaunchKernelGGL(k_conv_wgrad<8>, wg_grid, wg_block, 0, st, g, ncell, cin, cout, kpad, npad, in16, g16, work);
  else if (cin % 4 == 0 && ((uintptr_t)in % 8) == 0)
    hipLaunchKernelGGL(k_conv_wgrad<4>, wg_grid, wg_block, 0, st, g, ncell, cin, cout, kpad, npad, in16, g16, work);
  else
    hipLaunchKernelGGL(k_conv_wgrad<1>, wg_grid, wg_block, 0, st, g, ncell, cin, cout, kpad, npad, in16, g16, work);
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
