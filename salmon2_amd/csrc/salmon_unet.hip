// Learned potential predictor: a 3-level periodic U-Net with BF16 MFMA inference and
// training kernels (C ABI in include/salmon_unet.h).
//
// Tensors are channels-last (point-major, channel fastest), p = ix + nx*(iy + ny*iz).
// Activations and MFMA operands are BF16 (round to nearest even), sums FP32.
//
//   x0 (16 ch: 4 inputs + zero pad)            level 0 (n)
//   e1a = relu(conv(x0)) 16, e1b = relu(conv(e1a)) 16
//   p1 = avgpool(e1b)                          level 1 (n/2)
//   e2a = relu(conv(p1)) 32, e2b = relu(conv(e2a)) 32
//   p2 = avgpool(e2b)                          level 2 (n/4)
//   e3a = relu(conv(p2)) 64, e3b = relu(conv(e3a)) 64
//   d2a = relu(conv([up(e3b), e2b])) 32, d2b = relu(conv(d2a)) 32      level 1
//   d1a = relu(conv([up(d2b), e1b])) 16, d1b = relu(conv(d1a)) 16      level 0
//   out = head(d1b) (1x1x1, FP32)
// conv = 3x3x3 periodic, weights [27][co][ci]; `up` = nearest upsampling, read on the fly.
//
// Kernels:
//   k_conv<NT>     implicit GEMM on gfx950's double-K v_mfma_f32_16x16x32_bf16: M = 16
//                  points per wave, N = NT x 16 output channels, K = 32 of the flattened
//                  (tap, input channel) index (two taps per instruction at 16 channels);
//                  also the data gradient (flipped, transposed weights, FP32 output)
//   k_wgrad        weight gradient per (tap, co tile, ci tile) as MFMA over point chunks
//                  (M = co, N = ci, K = 32 points), FP32 partials reduced in fixed order
//   small kernels  pooling, gradient routing (mask, down-sum, pool-back), head, MSE, Adam
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "salmon_hip.h"
#include "salmon_unet.h"

void salmon_hip_internal_error(const char *msg);  // salmon_hip.hip: sets salmon_hip_last_error()

namespace {

typedef unsigned short bf16_t;
typedef short s8v __attribute__((ext_vector_type(8)));
typedef __bf16 b8v __attribute__((ext_vector_type(8)));
typedef float f4v __attribute__((ext_vector_type(4)));

// v_mfma_f32_16x16x32_bf16 (gfx950): lane L holds A[L%16][8(L/16) + e] and
// B[8(L/16) + e][L%16], e = 0..7; D[4(L/16) + r][L%16] lands in register r.
__device__ __forceinline__ f4v mfma32(s8v a, s8v b, f4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8v, a), __builtin_bit_cast(b8v, b), c, 0, 0, 0);
}

int ufail(int code, const char *fmt, ...);

#define UHIP(x)                                                                               \
  do {                                                                                        \
    hipError_t e_ = (x);                                                                      \
    if (e_ != hipSuccess) return ufail(SALMON_HIP_EHIP, "%s: %s", #x, hipGetErrorString(e_)); \
  } while (0)
#define ULAUNCH(name)                                                                                   \
  do {                                                                                                  \
    hipError_t e_ = hipGetLastError();                                                                  \
    if (e_ != hipSuccess) return ufail(SALMON_HIP_EHIP, "launch %s: %s", name, hipGetErrorString(e_)); \
  } while (0)

__device__ __forceinline__ bf16_t f2bf(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float((unsigned)h << 16); }

// ---------------------------------------------------------------------------
// 3x3x3 periodic convolution as implicit GEMM.  Wave w of block b owns points
// p0 = 16 (4b + w) .. p0 + 15; A row = point, B col = output channel, K = the
// flattened index f = tap * ci + input channel.  One MFMA takes 32 consecutive f:
// lane group q = L/16 supplies f0 + 8q .. f0 + 8q + 7, which lie in one tap and one
// 16-channel chunk (ci is a multiple of 16), i.e. one 16-byte load of activations at
// the tap's neighbour and one of weights.  The input is [up(src1), src2] (channel
// concatenation, src1 at half resolution when up1).
// ---------------------------------------------------------------------------
constexpr size_t kConvZero = 128 * 128 + 64;  // bf16 zeros read by k_conv (co x ci up to 128 x 128)

struct ConvArgs {
  const bf16_t *src1, *src2;
  int c1, c2, up1;
  const bf16_t *w;      // [27][co][c1 + c2]
  const float *bias;    // [co] or null
  bf16_t *out_bf;       // [npts][co] (bias, optional relu), or
  float *out_f;         // [npts][co] raw
  int co, relu;
  int nx, ny, nz, npts;
  const bf16_t *zero;   // zero-filled, at least co * (c1 + c2) + 8 elements
};

// SK = false: each wave takes its own 16 points and the whole K; SK = true (coarse levels, few
// points): the four waves of a block share 16 points and take every fourth K step each, and wave
// 0 adds the four partial sums in a fixed order (deterministic).
template <int NT, bool SK>
__global__ __launch_bounds__(256) void k_conv(ConvArgs a) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int p0 = (SK ? blockIdx.x : blockIdx.x * 4 + wave) * 16;
  if (p0 >= a.npts) return;
  const int row = lane & 15, kq = lane >> 4;
  const int p = p0 + row;
  const bool pv = p < a.npts;
  const int pp = pv ? p : 0;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int x = pp % nx, y = (pp / nx) % ny, z = pp / (nx * ny);
  const int ci = a.c1 + a.c2;
  const int K = 27 * ci;
  f4v acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = f4v{0.f, 0.f, 0.f, 0.f};
  // the operands of K step f0 (32 wide: lane group kq takes f0 + 8 kq .. + 7)
  // (branch-free: what lies outside the point range or past K is read from a zero buffer, so
  // every lane issues the same loads and each wait can sit at the MFMA that uses the data)
  auto load = [&](int f0, s8v &av, s8v (&bv)[NT]) __attribute__((always_inline)) {
    const int f = f0 + 8 * kq;
    const bool fv = f < K;
    const int fc = fv ? f : 0;
    const int tap = fc / ci, c = fc - tap * ci;
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    const int xn = (x + dx + nx) % nx, yn = (y + dy + ny) % ny, zn = (z + dz + nz) % nz;
    const size_t pn = (size_t)xn + (size_t)nx * ((size_t)yn + (size_t)ny * zn);
    const size_t ph = a.up1 ? (size_t)(xn >> 1) + (size_t)(nx >> 1) * ((size_t)(yn >> 1) + (size_t)(ny >> 1) * (zn >> 1))
                            : pn;
    const bf16_t *src = c < a.c1 ? a.src1 + ph * a.c1 + c : a.src2 + pn * a.c2 + (c - a.c1);
    av = *(const s8v *)((fv && pv) ? src : a.zero);
    const bf16_t *wt = fv ? a.w + (size_t)tap * a.co * ci + c : a.zero;
#pragma unroll
    for (int t = 0; t < NT; ++t) bv[t] = *(const s8v *)(wt + (size_t)(t * 16 + row) * ci);
  };
  // two operand sets: the loads of step f0 + 32 are in flight while the MFMAs of step f0 run
  // (the accumulation order over f0 is unchanged)
  constexpr int kst = SK ? 4 * 32 : 32;  // this wave's K steps: fs, fs + kst, ...
  const int fs = SK ? 32 * wave : 0;
  s8v avA, avB, bvA[NT], bvB[NT];
  load(fs, avA, bvA);
  for (int f0 = fs; f0 < K; f0 += 2 * kst) {
    if (f0 + kst < K) load(f0 + kst, avB, bvB);
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = mfma32(avA, bvA[t], acc[t]);
    if (f0 + kst >= K) break;
    if (f0 + 2 * kst < K) load(f0 + 2 * kst, avA, bvA);
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = mfma32(avB, bvB[t], acc[t]);
  }
  if constexpr (SK) {
    __shared__ f4v red[3][NT][64];
    if (wave > 0)
#pragma unroll
      for (int t = 0; t < NT; ++t) red[wave - 1][t][lane] = acc[t];
    __syncthreads();
    if (wave != 0) return;
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = ((acc[t] + red[0][t][lane]) + red[1][t][lane]) + red[2][t][lane];
  }
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int pr = p0 + 4 * kq + r;
      const int ch = t * 16 + row;
      if (pr >= a.npts) continue;
      float v = acc[t][r];
      if (a.out_bf) {
        if (a.bias) v += a.bias[ch];
        if (a.relu) v = fmaxf(v, 0.f);
        a.out_bf[(size_t)pr * a.co + ch] = f2bf(v);
      } else {
        a.out_f[(size_t)pr * a.co + ch] = v;
      }
    }
}

// ---------------------------------------------------------------------------
// Weight gradient dW[tap][co][ci] = sum_p dy[p][co] x[p + d(tap)][ci] over one chunk of
// points per grid row y; wave = (tap, co tile, ci tile); A = dy^T (M = co, K = points),
// B = shifted input (K = points, N = ci), 32 points per MFMA (lane group q takes points
// pb + 8q .. pb + 8q + 7).  part[(chunk*27 + tap)*co + co_][ci_].
// ---------------------------------------------------------------------------
struct WgArgs {
  const bf16_t *dy;
  int co;
  const bf16_t *src1, *src2;
  int c1, c2, up1;
  float *part;
  int nx, ny, nz, npts, chunk;
};

__global__ __launch_bounds__(256) void k_wgrad(WgArgs a) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ci = a.c1 + a.c2;
  const int nco = a.co / 16, nci = ci / 16;
  const int task = blockIdx.x * 4 + wave;
  if (task >= 27 * nco * nci) return;
  const int cit = task % nci, cot = (task / nci) % nco, tap = task / (nci * nco);
  const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
  const int row = lane & 15, kq = lane >> 4;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int pbeg = blockIdx.y * a.chunk, pend = min(pbeg + a.chunk, a.npts);
  const int cc = cit * 16 + row;
  const bool first = cit * 16 < a.c1;
  f4v acc = f4v{0.f, 0.f, 0.f, 0.f};
  for (int pb = pbeg; pb < pend; pb += 32) {
    s8v av, bv;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int pk = pb + 8 * kq + e;
      short va = 0, vb = 0;
      if (pk < pend) {
        va = (short)a.dy[(size_t)pk * a.co + cot * 16 + row];
        const int x = pk % nx, y = (pk / nx) % ny, z = pk / (nx * ny);
        const int xn = (x + dx + nx) % nx, yn = (y + dy + ny) % ny, zn = (z + dz + nz) % nz;
        if (first) {
          const size_t ph = a.up1 ? (size_t)(xn >> 1) + (size_t)(nx >> 1) * ((size_t)(yn >> 1) + (size_t)(ny >> 1) * (zn >> 1))
                                  : (size_t)xn + (size_t)nx * ((size_t)yn + (size_t)ny * zn);
          vb = (short)a.src1[ph * a.c1 + cc];
        } else {
          const size_t pn = (size_t)xn + (size_t)nx * ((size_t)yn + (size_t)ny * zn);
          vb = (short)a.src2[pn * a.c2 + (cc - a.c1)];
        }
      }
      av[e] = va;
      bv[e] = vb;
    }
    acc = mfma32(av, bv, acc);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r)
    a.part[((size_t)(blockIdx.y * 27 + tap) * a.co + cot * 16 + 4 * kq + r) * ci + cc] = acc[r];
}

// grad[i] += sum_k part[k][i], k ascending
__global__ __launch_bounds__(256) void k_reduce_add(const float *part, int nk, size_t n, float *grad) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int k = 0; k < nk; ++k) s += part[(size_t)k * n + i];
  grad[i] += s;
}

// grad[c] += sum_p g[p][c] (bias gradient), one block per channel, fixed-order tree
__global__ __launch_bounds__(256) void k_colsum(const bf16_t *g, int npts, int C, float *grad) {
  __shared__ float sh[256];
  const int c = blockIdx.x;
  float s = 0.f;
  for (int p = threadIdx.x; p < npts; p += 256) s += bf2f(g[(size_t)p * C + c]);
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) grad[c] += sh[0];
}

// 2x2x2 average pool: out (hx, hy, hz) = mean of its 8 children
__global__ __launch_bounds__(256) void k_pool(const bf16_t *in, bf16_t *out, int hx, int hy, int hz, int C) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  const size_t n = (size_t)hx * hy * hz * C;
  if (i >= n) return;
  const int c = i % C;
  const size_t q = i / C;
  const int x = q % hx, y = (q / hx) % hy, z = q / ((size_t)hx * hy);
  const int nx = 2 * hx, ny = 2 * hy;
  float s = 0.f;
  for (int k = 0; k < 8; ++k) {
    const int xx = 2 * x + (k & 1), yy = 2 * y + ((k >> 1) & 1), zz = 2 * z + (k >> 2);
    s += bf2f(in[((size_t)xx + (size_t)nx * ((size_t)yy + (size_t)ny * zz)) * C + c]);
  }
  out[i] = f2bf(s * 0.125f);
}

// out[q][c] = sum of the 8 children of q of gin[.][off + c] (gradient of nearest upsampling)
__global__ __launch_bounds__(256) void k_downsum(const float *gin, int Cs, int off, int C, float *out, int hx, int hy,
                                                 int hz) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  const size_t n = (size_t)hx * hy * hz * C;
  if (i >= n) return;
  const int c = i % C;
  const size_t q = i / C;
  const int x = q % hx, y = (q / hx) % hy, z = q / ((size_t)hx * hy);
  const int nx = 2 * hx, ny = 2 * hy;
  float s = 0.f;
  for (int k = 0; k < 8; ++k) {
    const int xx = 2 * x + (k & 1), yy = 2 * y + ((k >> 1) & 1), zz = 2 * z + (k >> 2);
    s += gin[((size_t)xx + (size_t)nx * ((size_t)yy + (size_t)ny * zz)) * Cs + off + c];
  }
  out[i] = s;
}

// out[p][c] = bf16(gin[p][off + c] * (act[p][c] > 0)), act null = no mask
__global__ __launch_bounds__(256) void k_mask(const float *gin, int Cs, int off, int C, const bf16_t *act, size_t npts,
                                              bf16_t *out) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= npts * C) return;
  const size_t p = i / C;
  const int c = i % C;
  float v = gin[p * Cs + off + c];
  if (act && !(bf2f(act[i]) > 0.f)) v = 0.f;
  out[i] = f2bf(v);
}

// out[p][c] = bf16((gp[parent(p)][c] / 8 + gskip[p][off + c]) * (act[p][c] > 0))
__global__ __launch_bounds__(256) void k_poolback(const float *gp, const float *gskip, int Cs, int off, int C,
                                                  const bf16_t *act, int nx, int ny, int nz, bf16_t *out) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  const size_t n = (size_t)nx * ny * nz * C;
  if (i >= n) return;
  const int c = i % C;
  const size_t p = i / C;
  const int x = p % nx, y = (p / nx) % ny, z = p / ((size_t)nx * ny);
  const size_t q = (size_t)(x >> 1) + (size_t)(nx >> 1) * ((size_t)(y >> 1) + (size_t)(ny >> 1) * (z >> 1));
  float v = gp[q * C + c] * 0.125f;
  if (gskip) v += gskip[p * Cs + off + c];
  if (!(bf2f(act[i]) > 0.f)) v = 0.f;
  out[i] = f2bf(v);
}

__global__ __launch_bounds__(256) void k_pack_input(const float *in, size_t npts, bf16_t *x0) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= npts * 16) return;
  const size_t p = i / 16;
  const int c = i % 16;
  x0[i] = c < 4 ? f2bf(in[p * 4 + c]) : (bf16_t)0;
}

// out[p] = sum_c x[p][c] wh[c] + bh
__global__ __launch_bounds__(256) void k_head(const bf16_t *x, const float *wh, size_t npts, float *out) {
  const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= npts) return;
  float s = 0.f;
  for (int c = 0; c < 16; ++c) s += bf2f(x[p * 16 + c]) * wh[c];
  out[p] = s + wh[16];
}

// dout[p] = 2 (out - t) / n; part[block] = sum (out - t)^2 over the block
__global__ __launch_bounds__(256) void k_mse(const float *out, const float *t, size_t n, float *dout, float *part) {
  __shared__ float sh[256];
  const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
  float e2 = 0.f;
  if (p < n) {
    const float d = out[p] - t[p];
    dout[p] = 2.f * d / (float)n;
    e2 = d * d;
  }
  sh[threadIdx.x] = e2;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(256) void k_loss(const float *part, int nb, size_t n, float *loss) {
  __shared__ float sh[256];
  float s = 0.f;
  for (int k = threadIdx.x; k < nb; k += 256) s += part[k];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = sh[0] / (float)n;
}

// head gradients: block c < 16: grad[c] += sum_p dout[p] x[p][c]; block 16: grad[16] += sum dout
__global__ __launch_bounds__(256) void k_head_grad(const float *dout, const bf16_t *x, size_t npts, float *grad) {
  __shared__ float sh[256];
  const int c = blockIdx.x;
  float s = 0.f;
  for (size_t p = threadIdx.x; p < npts; p += 256) s += dout[p] * (c < 16 ? bf2f(x[p * 16 + c]) : 1.f);
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) grad[c] += sh[0];
}

// g[p][c] = bf16(dout[p] wh[c] (act[p][c] > 0))
__global__ __launch_bounds__(256) void k_head_back(const float *dout, const float *wh, const bf16_t *act, size_t npts,
                                                   bf16_t *g) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= npts * 16) return;
  const size_t p = i / 16;
  const int c = i % 16;
  g[i] = f2bf(bf2f(act[i]) > 0.f ? dout[p] * wh[c] : 0.f);
}

__global__ __launch_bounds__(256) void k_adam(size_t n, float *w, float *g, float *m, float *v, float lr, float c1,
                                              float c2, float inv_ns) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i] * inv_ns;
  const float mi = 0.9f * m[i] + 0.1f * gi;
  const float vi = 0.999f * v[i] + 0.001f * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
  g[i] = 0.f;
}

// wt[tap][co][ci] = bf16(w), wb[26 - tap][ci][co] = bf16(w) (data-gradient weights)
__global__ __launch_bounds__(256) void k_pack_w(const float *w, int co, int ci, bf16_t *wt, bf16_t *wb) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (size_t)27 * co * ci) return;
  const int c = i % ci;
  const int o = (i / ci) % co;
  const int tap = i / ((size_t)co * ci);
  const bf16_t h = f2bf(w[i]);
  wt[i] = h;
  wb[((size_t)(26 - tap) * ci + c) * co + o] = h;
}

// ---------------------------------------------------------------------------
// Predictor glue on FP64 potentials V_local(ix,iy,iz), history v0 = V(t), v1 = V(t-dt), ...
// ---------------------------------------------------------------------------
// in[p] = ((v0 - v1)/sd, (v1 - v2)/sd, (v2 - v3)/sd, (v0 - vpsl)/sa)
__global__ __launch_bounds__(256) void k_features(const double *v0, const double *v1, const double *v2,
                                                  const double *v3, const double *vpsl, double sd, double sa,
                                                  size_t n, float *in) {
  const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  const double a = v0[p], b = v1[p], c = v2[p], d = v3[p];
  in[4 * p + 0] = (float)((a - b) / sd);
  in[4 * p + 1] = (float)((b - c) / sd);
  in[4 * p + 2] = (float)((c - d) / sd);
  in[4 * p + 3] = (float)((a - vpsl[p]) / sa);
}

// vpred = v0 + (v0 - v1)/2 + so * nn   (linear extrapolation to t + dt/2 plus the learned correction)
__global__ __launch_bounds__(256) void k_predict(const double *v0, const double *v1, const float *nn, double so,
                                                 size_t n, double *vpred) {
  const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  const double a = v0[p];
  vpred[p] = a + 0.5 * (a - v1[p]) + so * (double)nn[p];
}

// t = (vmid - v0 - (v0 - v1)/2) / so
__global__ __launch_bounds__(256) void k_target(const double *vmid, const double *v0, const double *v1, double so,
                                                size_t n, float *t) {
  const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  const double a = v0[p];
  t[p] = (float)((vmid[p] - a - 0.5 * (a - v1[p])) / so);
}

// part[block] = max_p |vpred - (v0 + vnew)/2| over the block; out = max over blocks
__global__ __launch_bounds__(256) void k_resid_part(const double *vpred, const double *v0, const double *vnew, size_t n,
                                                    double *part) {
  __shared__ double sh[256];
  double m = 0.0;
  for (size_t p = (size_t)blockIdx.x * 256 + threadIdx.x; p < n; p += (size_t)gridDim.x * 256)
    m = fmax(m, fabs(vpred[p] - 0.5 * (v0[p] + vnew[p])));
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + k]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(256) void k_resid_final(const double *part, int nb, double *out) {
  __shared__ double sh[256];
  double m = 0.0;
  for (int k = threadIdx.x; k < nb; k += 256) m = fmax(m, part[k]);
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + k]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

// The gate of a step propagated with a predicted potential, decided on the device: flag =
// (resid <= tol); stats[0] counts accepted, stats[1] attempted steps; log[stats[1]] = resid while
// the log has room.
__global__ void k_gate(const double *resid, double tol, int *flag, int *stats, double *log, int log_cap) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double r = *resid;
  const int ok = r <= tol ? 1 : 0;
  const int n = stats[1];
  if (log && n < log_cap) log[n] = r;
  *flag = ok;
  stats[0] += ok;
  stats[1] = n + 1;
}

int ufail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  salmon_hip_internal_error(buf);
  return code;
}

unsigned blocks(size_t n) { return (unsigned)((n + 255) / 256); }

struct Layer {
  int co, c1, c2, up1, level;  // input [up(src1) (c1), src2 (c2)] at `level`
  size_t woff, boff;           // offsets in the flat parameter array
  bf16_t *wt = nullptr, *wb = nullptr;
};

}  // namespace

struct salmon_unet {
  int n[3][3];   // grid extents per level
  size_t np[3];  // points per level
  hipStream_t stream = nullptr;
  std::vector<Layer> L;
  size_t nparam = 0, head_off = 0;
  float *d_w = nullptr, *d_g = nullptr, *d_m = nullptr, *d_v = nullptr;
  int adam_t = 0;
  bool packed = false;
  // activations (bf16): x0, e1a, e1b, p1, e2a, e2b, p2, e3a, e3b, d2a, d2b, d1a, d1b
  bf16_t *x0 = nullptr, *act[12] = {};
  bf16_t *d_zero = nullptr;  // kConvZero zeros: the operands of k_conv outside the grid or past K
  // backward work
  float *d_out = nullptr, *d_dout = nullptr, *d_part = nullptr, *d_fx = nullptr, *d_fy = nullptr, *d_wpart = nullptr;
  bf16_t *g_a = nullptr, *g_b = nullptr;
  size_t wpart_n = 0, fx_n = 0;
  int chunk = 1024;
  double *d_rpart = nullptr;  // residual partial maxima
  std::vector<void *> allocs;
};

namespace {

enum { E1A = 0, E1B, P1, E2A, E2B, P2, E3A, E3B, D2A, D2B, D1A, D1B };

template <typename T>
int ualloc(salmon_unet *u, T **p, size_t n) {
  *p = nullptr;
  if (hipMalloc((void **)p, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess)
    return ufail(SALMON_HIP_ENOMEM, "hipMalloc(%zu bytes) failed", sizeof(T) * n);
  u->allocs.push_back((void *)*p);
  return 0;
}

int launch_conv(salmon_unet *u, const bf16_t *s1, int c1, int up1, const bf16_t *s2, int c2, const bf16_t *w,
                const float *bias, int co, int relu, int level, bf16_t *out_bf, float *out_f) {
  ConvArgs a;
  a.src1 = s1; a.src2 = s2; a.c1 = c1; a.c2 = c2; a.up1 = up1;
  a.w = w; a.bias = bias; a.out_bf = out_bf; a.out_f = out_f; a.co = co; a.relu = relu;
  a.nx = u->n[level][0]; a.ny = u->n[level][1]; a.nz = u->n[level][2];
  a.npts = (int)u->np[level];
  a.zero = u->d_zero;
  if ((size_t)co * (c1 + c2) + 8 > kConvZero)
    return ufail(SALMON_HIP_EUNSUPPORTED, "conv %d -> %d channels above the zero operand", c1 + c2, co);
  // fewer 64-point blocks than CUs: split K over the four waves of 16-point blocks
  const bool sk = (a.npts + 63) / 64 < 256;
  const unsigned nb = (unsigned)(sk ? (a.npts + 15) / 16 : (a.npts + 63) / 64);
#define CONV_LAUNCH(NTv)                                                                   \
  do {                                                                                     \
    if (sk) hipLaunchKernelGGL((k_conv<NTv, true>), dim3(nb), dim3(256), 0, u->stream, a); \
    else hipLaunchKernelGGL((k_conv<NTv, false>), dim3(nb), dim3(256), 0, u->stream, a);   \
  } while (0)
  switch (co / 16) {
    case 1: CONV_LAUNCH(1); break;
    case 2: CONV_LAUNCH(2); break;
    case 3: CONV_LAUNCH(3); break;
    case 4: CONV_LAUNCH(4); break;
    case 6: CONV_LAUNCH(6); break;
    default: return ufail(SALMON_HIP_EUNSUPPORTED, "conv with %d output channels", co);
  }
#undef CONV_LAUNCH
  ULAUNCH("k_conv");
  return 0;
}

const bf16_t *layer_src1(salmon_unet *u, int i) {
  static const int src[12] = {-1, E1A, -2, P1, E2A, -2, P2, E3A, E3B, D2A, D2B, D1A};
  return src[i] == -1 ? u->x0 : u->act[src[i]];
}
const bf16_t *layer_src2(salmon_unet *u, int i) {
  if (i == D2A) return u->act[E2B];
  if (i == D1A) return u->act[E1B];
  return nullptr;
}

int pack_weights(salmon_unet *u) {
  for (Layer &l : u->L) {
    const int ci = l.c1 + l.c2;
    hipLaunchKernelGGL(k_pack_w, dim3(blocks((size_t)27 * l.co * ci)), dim3(256), 0, u->stream, u->d_w + l.woff, l.co,
                       ci, l.wt, l.wb);
    ULAUNCH("k_pack_w");
  }
  u->packed = true;
  return 0;
}

int forward(salmon_unet *u, const float *in, float *out) {
  if (!u->packed)
    if (int rc = pack_weights(u)) return rc;
  hipLaunchKernelGGL(k_pack_input, dim3(blocks(u->np[0] * 16)), dim3(256), 0, u->stream, in, u->np[0], u->x0);
  ULAUNCH("k_pack_input");
  int li = 0;
  for (int i = 0; i < 12; ++i) {
    if (i == P1 || i == P2) {
      const int lv = i == P1 ? 1 : 2;
      const int C = i == P1 ? 16 : 32;
      hipLaunchKernelGGL(k_pool, dim3(blocks(u->np[lv] * C)), dim3(256), 0, u->stream, u->act[i - 1], u->act[i],
                         u->n[lv][0], u->n[lv][1], u->n[lv][2], C);
      ULAUNCH("k_pool");
      continue;
    }
    Layer &l = u->L[li++];
    if (int rc = launch_conv(u, layer_src1(u, i), l.c1, l.up1, layer_src2(u, i), l.c2, l.wt, u->d_w + l.boff, l.co, 1,
                             l.level, u->act[i], nullptr))
      return rc;
  }
  hipLaunchKernelGGL(k_head, dim3(blocks(u->np[0])), dim3(256), 0, u->stream, u->act[D1B], u->d_w + u->head_off,
                     u->np[0], out);
  ULAUNCH("k_head");
  return 0;
}

// weight and bias gradient of conv layer l (act index i) from its pre-activation gradient g
int layer_wgrad(salmon_unet *u, int i, const Layer &l, const bf16_t *g) {
  const int ci = l.c1 + l.c2;
  WgArgs a;
  a.dy = g; a.co = l.co;
  a.src1 = layer_src1(u, i); a.src2 = layer_src2(u, i); a.c1 = l.c1; a.c2 = l.c2; a.up1 = l.up1;
  a.part = u->d_wpart;
  a.nx = u->n[l.level][0]; a.ny = u->n[l.level][1]; a.nz = u->n[l.level][2];
  a.npts = (int)u->np[l.level];
  a.chunk = u->chunk;
  const int nchunk = (a.npts + a.chunk - 1) / a.chunk;
  const int ntask = 27 * (l.co / 16) * (ci / 16);
  hipLaunchKernelGGL(k_wgrad, dim3((ntask + 3) / 4, nchunk), dim3(256), 0, u->stream, a);
  ULAUNCH("k_wgrad");
  const size_t nw = (size_t)27 * l.co * ci;
  hipLaunchKernelGGL(k_reduce_add, dim3(blocks(nw)), dim3(256), 0, u->stream, u->d_wpart, nchunk, nw, u->d_g + l.woff);
  ULAUNCH("k_reduce_add");
  hipLaunchKernelGGL(k_colsum, dim3(l.co), dim3(256), 0, u->stream, g, a.npts, l.co, u->d_g + l.boff);
  ULAUNCH("k_colsum");
  return 0;
}

// FP32 data gradient of conv layer l's input into u->d_fx ([np][c1 + c2])
int layer_dgrad(salmon_unet *u, const Layer &l, const bf16_t *g) {
  const int ci = l.c1 + l.c2;
  return launch_conv(u, g, l.co, 0, nullptr, 0, l.wb, nullptr, ci, 0, l.level, nullptr, u->d_fx);
}

int backward(salmon_unet *u, const float *target, float *loss) {
  const size_t n0 = u->np[0];
  const unsigned nb0 = blocks(n0);
  hipLaunchKernelGGL(k_mse, dim3(nb0), dim3(256), 0, u->stream, u->d_out, target, n0, u->d_dout, u->d_part);
  ULAUNCH("k_mse");
  if (loss) {
    hipLaunchKernelGGL(k_loss, dim3(1), dim3(256), 0, u->stream, u->d_part, (int)nb0, n0, loss);
    ULAUNCH("k_loss");
  }
  hipLaunchKernelGGL(k_head_grad, dim3(17), dim3(256), 0, u->stream, u->d_dout, u->act[D1B], n0, u->d_g + u->head_off);
  ULAUNCH("k_head_grad");
  bf16_t *g = u->g_a, *gn = u->g_b;
  hipLaunchKernelGGL(k_head_back, dim3(blocks(n0 * 16)), dim3(256), 0, u->stream, u->d_dout, u->d_w + u->head_off,
                     u->act[D1B], n0, g);
  ULAUNCH("k_head_back");
  // conv layers in reverse; layer index li <-> act index
  static const int conv_act[10] = {E1A, E1B, E2A, E2B, E3A, E3B, D2A, D2B, D1A, D1B};
  for (int li = 9; li >= 0; --li) {
    const Layer &l = u->L[li];
    const int i = conv_act[li];
    if (int rc = layer_wgrad(u, i, l, g)) return rc;
    if (li == 0) break;
    if (int rc = layer_dgrad(u, l, g)) return rc;
    const int lv = l.level;
    const int ci = l.c1 + l.c2;
    switch (i) {
      case D1B: case D2B: case E3B: case E2B: case E1B: {
        // input is the previous activation: mask by it
        hipLaunchKernelGGL(k_mask, dim3(blocks(u->np[lv] * ci)), dim3(256), 0, u->stream, u->d_fx, ci, 0, ci,
                           u->act[i - 1], u->np[lv], gn);
        ULAUNCH("k_mask");
        break;
      }
      case D1A: case D2A: {
        // input [up(d2b | e3b), e1b | e2b]: the up part goes to the coarse activation, the skip part is kept
        const int up = i == D1A ? D2B : E3B;
        const int C1 = l.c1;
        float *fy = u->d_fy;
        hipLaunchKernelGGL(k_downsum, dim3(blocks(u->np[lv + 1] * C1)), dim3(256), 0, u->stream, u->d_fx, ci, 0, C1,
                           fy, u->n[lv + 1][0], u->n[lv + 1][1], u->n[lv + 1][2]);
        ULAUNCH("k_downsum");
        hipLaunchKernelGGL(k_mask, dim3(blocks(u->np[lv + 1] * C1)), dim3(256), 0, u->stream, fy, C1, 0, C1,
                           u->act[up], u->np[lv + 1], gn);
        ULAUNCH("k_mask");
        // keep the skip gradient: copy the FP32 tensor aside (fx is reused by the next layers)
        float *keep = i == D1A ? u->d_fx + u->fx_n : u->d_fx + u->fx_n + n0 * 48;
        UHIP(hipMemcpyAsync(keep, u->d_fx, sizeof(float) * u->np[lv] * ci, hipMemcpyDeviceToDevice, u->stream));
        break;
      }
      case E3A: case E2A: {
        // input is a pooled tensor: gradient goes back through the pool into e2b | e1b, plus the skip part
        const int dst = i == E3A ? E2B : E1B;
        const int lvf = lv - 1;
        const int Cs = i == E3A ? 96 : 48;
        const int off = i == E3A ? 64 : 32;
        const float *skip = i == E3A ? u->d_fx + u->fx_n + n0 * 48 : u->d_fx + u->fx_n;
        hipLaunchKernelGGL(k_poolback, dim3(blocks(u->np[lvf] * ci)), dim3(256), 0, u->stream, u->d_fx, skip, Cs, off,
                           ci, u->act[dst], u->n[lvf][0], u->n[lvf][1], u->n[lvf][2], gn);
        ULAUNCH("k_poolback");
        break;
      }
      default:
        return ufail(SALMON_HIP_EINVAL, "internal: layer %d", i);
    }
    std::swap(g, gn);
  }
  return 0;
}

int ucheck(salmon_unet_t u) { return u ? 0 : ufail(SALMON_HIP_EINVAL, "null predictor handle"); }

}  // namespace

extern "C" {

int salmon_unet_create(int nx, int ny, int nz, unsigned long long seed, salmon_unet_t *out) {
  if (!out) return ufail(SALMON_HIP_EINVAL, "out is null");
  *out = nullptr;
  if (nx % 4 || ny % 4 || nz % 4 || nx < 8 || ny < 8 || nz < 8)
    return ufail(SALMON_HIP_EUNSUPPORTED, "grid %dx%dx%d: extents must be multiples of 4, >= 8", nx, ny, nz);
  salmon_unet *u = new salmon_unet();
  for (int lv = 0; lv < 3; ++lv) {
    u->n[lv][0] = nx >> lv; u->n[lv][1] = ny >> lv; u->n[lv][2] = nz >> lv;
    u->np[lv] = (size_t)u->n[lv][0] * u->n[lv][1] * u->n[lv][2];
  }
  // conv layers: co, c1, c2, up1, level
  const int spec[10][5] = {{16, 16, 0, 0, 0}, {16, 16, 0, 0, 0}, {32, 16, 0, 0, 1}, {32, 32, 0, 0, 1},
                           {64, 32, 0, 0, 2}, {64, 64, 0, 0, 2}, {32, 64, 32, 1, 1}, {32, 32, 0, 0, 1},
                           {16, 32, 16, 1, 0}, {16, 16, 0, 0, 0}};
  size_t off = 0;
  for (auto &sp : spec) {
    Layer l;
    l.co = sp[0]; l.c1 = sp[1]; l.c2 = sp[2]; l.up1 = sp[3]; l.level = sp[4];
    l.woff = off;
    off += (size_t)27 * l.co * (l.c1 + l.c2);
    l.boff = off;
    off += l.co;
    u->L.push_back(l);
  }
  u->head_off = off;
  off += 17;
  u->nparam = off;
  auto bail = [&](int rc) { salmon_unet_destroy(u); return rc; };
  int rc;
  if ((rc = ualloc(u, &u->d_w, off)) || (rc = ualloc(u, &u->d_g, off)) || (rc = ualloc(u, &u->d_m, off)) ||
      (rc = ualloc(u, &u->d_v, off)))
    return bail(rc);
  for (Layer &l : u->L) {
    const size_t nw = (size_t)27 * l.co * (l.c1 + l.c2);
    if ((rc = ualloc(u, &l.wt, nw)) || (rc = ualloc(u, &l.wb, nw))) return bail(rc);
  }
  const size_t n0 = u->np[0], n1 = u->np[1], n2 = u->np[2];
  const size_t act_n[12] = {n0 * 16, n0 * 16, n1 * 16, n1 * 32, n1 * 32, n2 * 32, n2 * 64, n2 * 64,
                            n1 * 32, n1 * 32, n0 * 16, n0 * 16};
  if ((rc = ualloc(u, &u->x0, n0 * 16)) || (rc = ualloc(u, &u->d_zero, kConvZero))) return bail(rc);
  for (int i = 0; i < 12; ++i)
    if ((rc = ualloc(u, &u->act[i], act_n[i]))) return bail(rc);
  // d_fx: data gradient (max np * ci) followed by the two kept skip gradients (n0 x 48, n1 x 96)
  u->fx_n = std::max({n0 * 48, n1 * 96, n2 * 64, n1 * 32, n0 * 16});
  size_t wp = 0;
  for (Layer &l : u->L) {
    const size_t nch = (u->np[l.level] + u->chunk - 1) / u->chunk;
    wp = std::max(wp, nch * 27 * l.co * (l.c1 + l.c2));
  }
  u->wpart_n = wp;
  if ((rc = ualloc(u, &u->d_out, n0)) || (rc = ualloc(u, &u->d_dout, n0)) || (rc = ualloc(u, &u->d_part, blocks(n0))) ||
      (rc = ualloc(u, &u->d_fx, u->fx_n + n0 * 48 + n1 * 96)) || (rc = ualloc(u, &u->d_fy, n1 * 64)) ||
      (rc = ualloc(u, &u->d_wpart, wp)) || (rc = ualloc(u, &u->g_a, n0 * 64)) || (rc = ualloc(u, &u->g_b, n0 * 64)) ||
      (rc = ualloc(u, &u->d_rpart, 1024)))
    return bail(rc);
  // He-normal weights, zero biases; head: small normal weights, zero bias
  std::vector<float> w(off, 0.f);
  std::mt19937_64 gen(seed);
  for (Layer &l : u->L) {
    const int fan_in = 27 * (l.c1 + l.c2);
    std::normal_distribution<float> nd(0.f, std::sqrt(2.f / fan_in));
    for (size_t k = 0; k < (size_t)27 * l.co * (l.c1 + l.c2); ++k) w[l.woff + k] = nd(gen);
  }
  std::normal_distribution<float> nh(0.f, 0.25f);
  for (int c = 0; c < 16; ++c) w[u->head_off + c] = nh(gen);
  if (hipMemcpy(u->d_w, w.data(), sizeof(float) * off, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(u->d_g, 0, sizeof(float) * off) != hipSuccess || hipMemset(u->d_m, 0, sizeof(float) * off) != hipSuccess ||
      hipMemset(u->d_v, 0, sizeof(float) * off) != hipSuccess || hipMemset(u->x0, 0, sizeof(bf16_t) * n0 * 16) != hipSuccess ||
      hipMemset(u->d_zero, 0, sizeof(bf16_t) * kConvZero) != hipSuccess)
    return bail(ufail(SALMON_HIP_EHIP, "predictor initialisation copies failed"));
  *out = u;
  return 0;
}

int salmon_unet_destroy(salmon_unet_t u) {
  if (!u) return 0;
  if (u->stream) (void)hipStreamSynchronize(u->stream);
  else (void)hipDeviceSynchronize();
  for (void *p : u->allocs) (void)hipFree(p);
  delete u;
  return 0;
}

int salmon_unet_set_stream(salmon_unet_t u, void *stream) {
  if (int rc = ucheck(u)) return rc;
  u->stream = (hipStream_t)stream;
  return 0;
}

int salmon_unet_num_params(salmon_unet_t u, long *n) {
  if (int rc = ucheck(u)) return rc;
  if (!n) return ufail(SALMON_HIP_EINVAL, "null pointer");
  *n = (long)u->nparam;
  return 0;
}

int salmon_unet_get_params(salmon_unet_t u, float *host) {
  if (int rc = ucheck(u)) return rc;
  if (!host) return ufail(SALMON_HIP_EINVAL, "null pointer");
  UHIP(hipMemcpyAsync(host, u->d_w, sizeof(float) * u->nparam, hipMemcpyDeviceToHost, u->stream));
  UHIP(hipStreamSynchronize(u->stream));
  return 0;
}

int salmon_unet_set_params(salmon_unet_t u, const float *host) {
  if (int rc = ucheck(u)) return rc;
  if (!host) return ufail(SALMON_HIP_EINVAL, "null pointer");
  UHIP(hipMemcpyAsync(u->d_w, host, sizeof(float) * u->nparam, hipMemcpyHostToDevice, u->stream));
  UHIP(hipStreamSynchronize(u->stream));
  u->packed = false;
  return 0;
}

int salmon_unet_get_grads(salmon_unet_t u, float *host) {
  if (int rc = ucheck(u)) return rc;
  if (!host) return ufail(SALMON_HIP_EINVAL, "null pointer");
  UHIP(hipMemcpyAsync(host, u->d_g, sizeof(float) * u->nparam, hipMemcpyDeviceToHost, u->stream));
  UHIP(hipStreamSynchronize(u->stream));
  return 0;
}

int salmon_unet_get_activation(salmon_unet_t u, int index, float *host, long n) {
  if (int rc = ucheck(u)) return rc;
  if (!host || index < -1 || index >= 12) return ufail(SALMON_HIP_EINVAL, "activation %d", index);
  static const int ch[12] = {16, 16, 16, 32, 32, 32, 64, 64, 32, 32, 16, 16};
  static const int lv[12] = {0, 0, 1, 1, 1, 2, 2, 2, 1, 1, 0, 0};
  const size_t cnt = index < 0 ? u->np[0] * 16 : u->np[lv[index]] * ch[index];
  if ((size_t)n != cnt) return ufail(SALMON_HIP_EINVAL, "activation %d has %zu values, not %ld", index, cnt, n);
  std::vector<bf16_t> tmp(cnt);
  UHIP(hipMemcpyAsync(tmp.data(), index < 0 ? u->x0 : u->act[index], sizeof(bf16_t) * cnt, hipMemcpyDeviceToHost,
                      u->stream));
  UHIP(hipStreamSynchronize(u->stream));
  for (size_t i = 0; i < cnt; ++i) {
    const unsigned v = (unsigned)tmp[i] << 16;
    memcpy(host + i, &v, 4);
  }
  return 0;
}

int salmon_unet_forward(salmon_unet_t u, const float *in_dev, float *out_dev) {
  if (int rc = ucheck(u)) return rc;
  if (!in_dev || !out_dev) return ufail(SALMON_HIP_EINVAL, "null pointer");
  return forward(u, in_dev, out_dev);
}

int salmon_unet_train_sample(salmon_unet_t u, const float *in_dev, const float *target_dev, float *loss_dev) {
  if (int rc = ucheck(u)) return rc;
  if (!in_dev || !target_dev) return ufail(SALMON_HIP_EINVAL, "null pointer");
  if (int rc = forward(u, in_dev, u->d_out)) return rc;
  return backward(u, target_dev, loss_dev);
}

int salmon_unet_features(salmon_unet_t u, const double *v0, const double *v1, const double *v2, const double *v3,
                         const double *vpsl, double s_diff, double s_abs, float *in_dev) {
  if (int rc = ucheck(u)) return rc;
  if (!v0 || !v1 || !v2 || !v3 || !vpsl || !in_dev) return ufail(SALMON_HIP_EINVAL, "null pointer");
  if (!(s_diff > 0.0) || !(s_abs > 0.0)) return ufail(SALMON_HIP_EINVAL, "scales must be positive");
  hipLaunchKernelGGL(k_features, dim3(blocks(u->np[0])), dim3(256), 0, u->stream, v0, v1, v2, v3, vpsl, s_diff, s_abs,
                     u->np[0], in_dev);
  ULAUNCH("k_features");
  return 0;
}

int salmon_unet_predict(salmon_unet_t u, const double *v0, const double *v1, const float *nn_dev, double s_out,
                        double *vpred_dev) {
  if (int rc = ucheck(u)) return rc;
  if (!v0 || !v1 || !nn_dev || !vpred_dev) return ufail(SALMON_HIP_EINVAL, "null pointer");
  hipLaunchKernelGGL(k_predict, dim3(blocks(u->np[0])), dim3(256), 0, u->stream, v0, v1, nn_dev, s_out, u->np[0],
                     vpred_dev);
  ULAUNCH("k_predict");
  return 0;
}

int salmon_unet_target(salmon_unet_t u, const double *vmid, const double *v0, const double *v1, double s_out,
                       float *t_dev) {
  if (int rc = ucheck(u)) return rc;
  if (!vmid || !v0 || !v1 || !t_dev) return ufail(SALMON_HIP_EINVAL, "null pointer");
  if (!(s_out > 0.0)) return ufail(SALMON_HIP_EINVAL, "scale must be positive");
  hipLaunchKernelGGL(k_target, dim3(blocks(u->np[0])), dim3(256), 0, u->stream, vmid, v0, v1, s_out, u->np[0], t_dev);
  ULAUNCH("k_target");
  return 0;
}

int salmon_unet_residual(salmon_unet_t u, const double *vpred, const double *v0, const double *vnew, double *out_dev) {
  if (int rc = ucheck(u)) return rc;
  if (!vpred || !v0 || !vnew || !out_dev) return ufail(SALMON_HIP_EINVAL, "null pointer");
  const int nb = (int)std::min<size_t>(1024, blocks(u->np[0]));
  hipLaunchKernelGGL(k_resid_part, dim3(nb), dim3(256), 0, u->stream, vpred, v0, vnew, u->np[0], u->d_rpart);
  ULAUNCH("k_resid_part");
  hipLaunchKernelGGL(k_resid_final, dim3(1), dim3(256), 0, u->stream, u->d_rpart, nb, out_dev);
  ULAUNCH("k_resid_final");
  return 0;
}

int salmon_unet_gate(salmon_unet_t u, const double *resid_dev, double tol, int *flag_dev, int *stats_dev,
                     double *log_dev, int log_cap) {
  if (int rc = ucheck(u)) return rc;
  if (!resid_dev || !flag_dev || !stats_dev || (log_cap > 0 && !log_dev))
    return ufail(SALMON_HIP_EINVAL, "null pointer");
  hipLaunchKernelGGL(k_gate, dim3(1), dim3(64), 0, u->stream, resid_dev, tol, flag_dev, stats_dev,
                     log_cap > 0 ? log_dev : nullptr, log_cap);
  ULAUNCH("k_gate");
  return 0;
}

int salmon_unet_adam_step(salmon_unet_t u, float lr, int nsamples) {
  if (int rc = ucheck(u)) return rc;
  if (nsamples < 1) return ufail(SALMON_HIP_EINVAL, "nsamples=%d", nsamples);
  u->adam_t += 1;
  const float c1 = 1.f - std::pow(0.9f, (float)u->adam_t);
  const float c2 = 1.f - std::pow(0.999f, (float)u->adam_t);
  hipLaunchKernelGGL(k_adam, dim3(blocks(u->nparam)), dim3(256), 0, u->stream, u->nparam, u->d_w, u->d_g, u->d_m,
                     u->d_v, lr, c1, c2, 1.f / (float)nsamples);
  ULAUNCH("k_adam");
  u->packed = false;
  return 0;
}

}  // extern "C"
