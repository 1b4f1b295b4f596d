/*
 * warmstart.hip -- learned warm start for the 2-D Poisson solve on gfx950.
 *
 * A convolutional encoder-decoder maps the right-hand side f = dx^2 rho/eps0
 * of a square grid with phi = 0 on all four walls to phi at the interior
 * nodes; its output seeds the iterative solver (PCG or multigrid), which then
 * iterates to the same FP64 residual as from any other start, so the network
 * can only change the iteration count, never the answer.  Trained with the
 * energy objective (the A-norm of the error left by the network plus one
 * damped Jacobi sweep), the same network preconditions every iteration of a
 * flexible CG on the FP64 residual (iskra_ws_solve).  There is no
 * counterpart in the reference (its Poisson solve is a dense direct solve,
 * FiniteDifferenceMethod/src/generalized_poisson.jl:201-203,372-378).
 *
 * Network (L levels, channels c_l = base * 2^min(l, 3), 3x3 convolutions, zero
 * padding, no biases, so the map is positively homogeneous and inputs are
 * normalised by their RMS):
 *   encoder  l = 0..L-1:  Ea_l = relu(conv(P_l)),  E_l = relu(conv(Ea_l)),
 *                         P_0 = input (8 channels, channel 0 = f / s),
 *                         P_l = avgpool2(E_{l-1})
 *   decoder  l = L-2..0:  U_l = upsample2(D_{l+1}) (D_{L-1} = E_{L-1}),
 *                         Da_l = relu(conv(U_l) + E_l),  D_l = relu(conv(Da_l))
 *   output   o = conv(D_0) channel 0;  phi = o * gain * s.
 * Tensors are NHWC bfloat16 (interior nodes padded to side S = nodes - 1);
 * convolutions are implicit GEMMs on v_mfma_f32_16x16x32_bf16 with FP32
 * accumulation: forward and input-gradient kernels read 8 consecutive
 * channels per lane (16-byte loads) for both operands, the weight-gradient
 * kernel stages pixel tiles in LDS and reduces over pixels in fixed order
 * (split into chunks, summed chunk by chunk), so training is deterministic.
 * Gradients of activations are FP32; Adam keeps FP32 master weights.
 */
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "isk_internal.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

#define WS_MAX_LEVELS 8
#define WS_IN_CH 8       /* input channels (channel 0 carries f) */
#define WS_OUT_CH 16     /* output channels computed (channel 0 is used) */
#define WS_PART_BYTES (64ll << 20)
#define EW_BLOCK 256
#define WS_MAX_SWEEPS 16
#ifndef WS_TAP_UNROLL
#define WS_TAP_UNROLL 1
#endif
#ifndef WS_CONV_LDS
#define WS_CONV_LDS 1 /* the LDS-staged convolution where its shape fits */
#endif
#ifndef WS_CONV_MIN_WG
#define WS_CONV_MIN_WG 256 /* work-groups a conv tile should give (1 per CU) */
#endif

__host__ __device__ inline uint16_t ws_f2bf(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float ws_bf2f(uint16_t h) {
  return __builtin_bit_cast(float, (uint32_t)h << 16);
}

#define WRC(x)             \
  do {                     \
    int _rc = (x);         \
    if (_rc) return _rc;   \
  } while (0)

/* ------------------------------------------------------------------------ */
/* convolution: forward (mode 0, bf16 out, optional skip + relu), input       */
/* gradient (mode 1, fp32 accumulate), network output (mode 2, channel 0)     */

struct ConvArgs {
  const uint16_t *X;     /* NHWC, Cin channels */
  const uint16_t *Wt;    /* [Cout][Kpad], k = tap * Cin + ci, zero padded */
  const uint16_t *skip;  /* NHWC, Cout channels, or null */
  uint16_t *Y;           /* mode 0 */
  float *Yf;             /* mode 1: NHWC fp32 (+=); mode 2: plane [npix] */
  int H, W, Cin, Cout, Kpad, relu, mode;
};

/* A wave computes TM output channels x NP groups of 16 pixels: per 32-wide
 * K block it loads TM/16 weight fragments once and reuses each for the NP
 * pixel groups (TM/16 x NP MFMAs per block).  Work-group = 4 waves =
 * 64 NP pixels. */
template <int TM, int NP>
__global__ __launch_bounds__(256) void k_ws_conv(ConvArgs a) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int col = lane & 15, grp = lane >> 4;
  const int HW = a.H * a.W;
  long long p[NP];
  int n[NP], h[NP], w[NP];
#pragma unroll
  for (int g = 0; g < NP; ++g) {
    p[g] = (long long)blockIdx.x * (64 * NP) + wv * (16 * NP) + g * 16 + col;
    n[g] = (int)(p[g] / HW);
    const int rem = (int)(p[g] - (long long)n[g] * HW);
    h[g] = rem / a.W;
    w[g] = rem - h[g] * a.W;
  }
  const int co0 = blockIdx.y * TM;
  f32x4 acc[NP][TM / 16];
#pragma unroll
  for (int g = 0; g < NP; ++g)
#pragma unroll
    for (int j = 0; j < TM / 16; ++j) acc[g][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
  const int nkb = a.Kpad >> 5;
  for (int kb = 0; kb < nkb; ++kb) {
    const int k0 = kb * 32 + grp * 8;
    const int tap = k0 / a.Cin;
    const int ci0 = k0 - tap * a.Cin;
    u32x4 av[TM / 16];
#pragma unroll
    for (int j = 0; j < TM / 16; ++j)
      av[j] = *(const u32x4 *)(a.Wt + (long long)(co0 + j * 16 + col) * a.Kpad + k0);
    u32x4 bv[NP];
#pragma unroll
    for (int g = 0; g < NP; ++g) {
      bv[g] = (u32x4){0u, 0u, 0u, 0u};
      if (tap < 9) {
        const int hh = h[g] + tap / 3 - 1, ww = w[g] + tap % 3 - 1;
        if (hh >= 0 && hh < a.H && ww >= 0 && ww < a.W)
          bv[g] = *(const u32x4 *)(a.X + ((long long)(n[g] * a.H + hh) * a.W + ww) * a.Cin + ci0);
      }
    }
#pragma unroll
    for (int g = 0; g < NP; ++g)
#pragma unroll
      for (int j = 0; j < TM / 16; ++j)
        acc[g][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            __builtin_bit_cast(bf16x8, av[j]), __builtin_bit_cast(bf16x8, bv[g]), acc[g][j], 0, 0,
            0);
  }
#pragma unroll
  for (int g = 0; g < NP; ++g)
#pragma unroll
    for (int j = 0; j < TM / 16; ++j) {
      const int cb = co0 + j * 16 + grp * 4;
      float v[4] = {acc[g][j][0], acc[g][j][1], acc[g][j][2], acc[g][j][3]};
      if (a.mode == 0) {
        if (a.skip) {
          const uint2 s = *(const uint2 *)(a.skip + p[g] * a.Cout + cb);
          v[0] += ws_bf2f((uint16_t)(s.x & 0xffffu));
          v[1] += ws_bf2f((uint16_t)(s.x >> 16));
          v[2] += ws_bf2f((uint16_t)(s.y & 0xffffu));
          v[3] += ws_bf2f((uint16_t)(s.y >> 16));
        }
        if (a.relu) {
#pragma unroll
          for (int r = 0; r < 4; ++r) v[r] = fmaxf(v[r], 0.f);
        }
        uint2 o;
        o.x = (uint32_t)ws_f2bf(v[0]) | ((uint32_t)ws_f2bf(v[1]) << 16);
        o.y = (uint32_t)ws_f2bf(v[2]) | ((uint32_t)ws_f2bf(v[3]) << 16);
        *(uint2 *)(a.Y + p[g] * a.Cout + cb) = o;
      } else if (a.mode == 1) {
        float4 *dst = (float4 *)(a.Yf + p[g] * a.Cout + cb);
        float4 d = *dst;
        d.x += v[0];
        d.y += v[1];
        d.z += v[2];
        d.w += v[3];
        *dst = d;
      } else if (cb == 0) {
        a.Yf[p[g]] = v[0];
      }
    }
}

/* The same convolution with the input staged in LDS: a work-group computes
 * a 16 x 16 pixel tile x TM output channels (wave wv: rows 4 wv .. 4 wv + 3,
 * one 16-pixel row per MFMA column block).  Per chunk of 32 input channels
 * the 18 x 18 halo of the tile is staged once (pixel stride 40 channels, so
 * the 16 pixels a quarter-wave reads fall in distinct banks), then the 9
 * taps read their fragments from LDS.  Needs Cin % 32 == 0 and H, W
 * multiples of 16; K is ordered (tap, ci) as in Wt. */
#define WS_LDS_PIX 40
template <int TM>
__global__ __launch_bounds__(256) void k_ws_conv_lds(ConvArgs a) {
  __shared__ __align__(16) uint16_t tile[18 * 18 * WS_LDS_PIX];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int col = lane & 15, grp = lane >> 4;
  const int tw = a.W >> 4, th = a.H >> 4;
  int t = blockIdx.x;
  const int tc = t % tw;
  t /= tw;
  const int tr = t % th;
  const int n = t / th;
  const int h0 = tr * 16, w0 = tc * 16;
  const int co0 = blockIdx.y * TM;
  f32x4 acc[4][TM / 16];
#pragma unroll
  for (int g = 0; g < 4; ++g)
#pragma unroll
    for (int j = 0; j < TM / 16; ++j) acc[g][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
  for (int c0 = 0; c0 < a.Cin; c0 += 32) {
    __syncthreads(); /* the previous chunk's fragments have been read */
    for (int v = threadIdx.x; v < 18 * 18 * 4; v += 256) {
      const int px = v >> 2, q = v & 3;
      const int hy = px / 18, hx = px - hy * 18;
      const int hh = h0 + hy - 1, ww = w0 + hx - 1;
      u32x4 val = (u32x4){0u, 0u, 0u, 0u};
      if (hh >= 0 && hh < a.H && ww >= 0 && ww < a.W)
        val = *(const u32x4 *)(a.X + ((long long)(n * a.H + hh) * a.W + ww) * a.Cin + c0 + q * 8);
      *(u32x4 *)(tile + px * WS_LDS_PIX + q * 8) = val;
    }
    __syncthreads();
#pragma unroll WS_TAP_UNROLL
    for (int tap = 0; tap < 9; ++tap) {
      const int dy = tap / 3, dx = tap - (tap / 3) * 3;
      const int k0 = tap * a.Cin + c0 + grp * 8;
      u32x4 av[TM / 16];
#pragma unroll
      for (int j = 0; j < TM / 16; ++j)
        av[j] = *(const u32x4 *)(a.Wt + (long long)(co0 + j * 16 + col) * a.Kpad + k0);
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int row = wv * 4 + g;
        const u32x4 bv =
            *(const u32x4 *)(tile + ((row + dy) * 18 + col + dx) * WS_LDS_PIX + grp * 8);
#pragma unroll
        for (int j = 0; j < TM / 16; ++j)
          acc[g][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
              __builtin_bit_cast(bf16x8, av[j]), __builtin_bit_cast(bf16x8, bv), acc[g][j], 0, 0,
              0);
      }
    }
  }
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const long long p = ((long long)(n * a.H + h0 + wv * 4 + g)) * a.W + w0 + col;
#pragma unroll
    for (int j = 0; j < TM / 16; ++j) {
      const int cb = co0 + j * 16 + grp * 4;
      float v[4] = {acc[g][j][0], acc[g][j][1], acc[g][j][2], acc[g][j][3]};
      if (a.mode == 0) {
        if (a.skip) {
          const uint2 sk = *(const uint2 *)(a.skip + p * a.Cout + cb);
          v[0] += ws_bf2f((uint16_t)(sk.x & 0xffffu));
          v[1] += ws_bf2f((uint16_t)(sk.x >> 16));
          v[2] += ws_bf2f((uint16_t)(sk.y & 0xffffu));
          v[3] += ws_bf2f((uint16_t)(sk.y >> 16));
        }
        if (a.relu) {
#pragma unroll
          for (int r = 0; r < 4; ++r) v[r] = fmaxf(v[r], 0.f);
        }
        uint2 o;
        o.x = (uint32_t)ws_f2bf(v[0]) | ((uint32_t)ws_f2bf(v[1]) << 16);
        o.y = (uint32_t)ws_f2bf(v[2]) | ((uint32_t)ws_f2bf(v[3]) << 16);
        *(uint2 *)(a.Y + p * a.Cout + cb) = o;
      } else if (a.mode == 1) {
        float4 *dst = (float4 *)(a.Yf + p * a.Cout + cb);
        float4 d = *dst;
        d.x += v[0];
        d.y += v[1];
        d.z += v[2];
        d.w += v[3];
        *dst = d;
      } else if (cb == 0) {
        a.Yf[p] = v[0];
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* weight gradient: part[chunk][co][tap*Cin+ci] = sum over the chunk's pixels */
/* of G[p][co] * X[p + tap][ci]                                               */

struct WgradArgs {
  const uint16_t *X;  /* NHWC Cin */
  const uint16_t *G;  /* NHWC Cout (masked output gradient) */
  float *part;
  int H, W, Cin, Cout;
  long long nseg;     /* pixel segments of 64 */
  int seg_per_chunk;
};

__global__ __launch_bounds__(256) void k_ws_wgrad(WgradArgs a) {
  __shared__ __attribute__((aligned(16))) uint16_t sG[64 * 32];
  __shared__ __attribute__((aligned(16))) uint16_t sX[9 * 64 * 16];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int col = lane & 15, grp = lane >> 4;
  const int co0 = blockIdx.y * 32, ci0 = blockIdx.z * 16;
  const int cb = wv & 1;
  const int t0 = (wv >> 1) * 5;
  const int t1 = t0 + 5 < 9 ? t0 + 5 : 9;
  f32x4 acc[5];
#pragma unroll
  for (int t = 0; t < 5; ++t) acc[t] = (f32x4){0.f, 0.f, 0.f, 0.f};
  const int HW = a.H * a.W;
  const long long s0 = (long long)blockIdx.x * a.seg_per_chunk;
  const long long s1 = s0 + a.seg_per_chunk < a.nseg ? s0 + a.seg_per_chunk : a.nseg;
  for (long long s = s0; s < s1; ++s) {
    const long long pbase = s * 64;
    __syncthreads();
    {
      const int px = threadIdx.x >> 2, c8 = (threadIdx.x & 3) * 8;
      u32x4 v = (u32x4){0u, 0u, 0u, 0u};
      if (co0 + c8 < a.Cout) v = *(const u32x4 *)(a.G + (pbase + px) * a.Cout + co0 + c8);
      *(u32x4 *)(sG + px * 32 + c8) = v;
    }
    for (int idx = threadIdx.x; idx < 9 * 128; idx += 256) {
      const int tap = idx >> 7, r = idx & 127, px = r >> 1, c8 = (r & 1) * 8;
      const long long p = pbase + px;
      const int n = (int)(p / HW);
      const int rem = (int)(p - (long long)n * HW);
      const int h = rem / a.W, w = rem - h * a.W;
      const int hh = h + tap / 3 - 1, ww = w + tap % 3 - 1;
      u32x4 v = (u32x4){0u, 0u, 0u, 0u};
      if (ci0 + c8 < a.Cin && hh >= 0 && hh < a.H && ww >= 0 && ww < a.W)
        v = *(const u32x4 *)(a.X + ((long long)(n * a.H + hh) * a.W + ww) * a.Cin + ci0 + c8);
      *(u32x4 *)(sX + (tap * 64 + px) * 16 + c8) = v;
    }
    __syncthreads();
#pragma unroll
    for (int kc = 0; kc < 2; ++kc) {
      u16x8 av;
#pragma unroll
      for (int q = 0; q < 8; ++q) av[q] = sG[(kc * 32 + grp * 8 + q) * 32 + cb * 16 + col];
      const bf16x8 A = __builtin_bit_cast(bf16x8, av);
#pragma unroll
      for (int t = 0; t < 5; ++t) {
        if (t0 + t < t1) {
          u16x8 bv;
#pragma unroll
          for (int q = 0; q < 8; ++q)
            bv[q] = sX[((t0 + t) * 64 + kc * 32 + grp * 8 + q) * 16 + col];
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, __builtin_bit_cast(bf16x8, bv),
                                                           acc[t], 0, 0, 0);
        }
      }
    }
  }
  const long long K = 9ll * a.Cin;
  float *out = a.part + (long long)blockIdx.x * a.Cout * K;
  const int ci = ci0 + col;
#pragma unroll
  for (int t = 0; t < 5; ++t) {
    if (t0 + t >= t1) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int co = co0 + cb * 16 + grp * 4 + r;
      if (co < a.Cout && ci < a.Cin) out[co * K + (t0 + t) * a.Cin + ci] = acc[t][r];
    }
  }
}

/* grads[i] = sum_c part[c][i], chunks in order */
__global__ void k_ws_reduce(const float *__restrict__ part, float *__restrict__ out, long long n,
                            int nchunk) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int c = 0; c < nchunk; ++c) s += part[(long long)c * n + i];
    out[i] = s;
  }
}

/* ------------------------------------------------------------------------ */
/* elementwise kernels (grid-stride, one element per thread)                  */

__global__ void k_ws_fill(float *p, long long n, float v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

/* Pooling and upsampling move 8 bf16 channels (16 bytes) or 4 fp32
 * gradient channels per thread; C is a multiple of 8.  Vector index i covers
 * channels 8 (or 4) i .. of the flat NHWC array. */

/* P[n,h2,w2,c] = bf16(0.25 ((X00 + X01) + (X10 + X11))) over the 2x2 block */
__global__ void k_ws_pool(const uint16_t *X, uint16_t *P, int H, int W, int C, long long n_out) {
  const int H2 = H >> 1, W2 = W >> 1, CV = C >> 3;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < (n_out >> 3);
       i += (long long)gridDim.x * blockDim.x) {
    const int cv = (int)(i % CV);
    const long long pix = i / CV;
    const int w2 = (int)(pix % W2);
    const long long r = pix / W2;
    const int h2 = (int)(r % H2);
    const int n = (int)(r / H2);
    const long long b = ((long long)(n * H + 2 * h2) * W + 2 * w2) * C + 8 * cv;
    const u16x8 x00 = *(const u16x8 *)(X + b), x01 = *(const u16x8 *)(X + b + C);
    const u16x8 x10 = *(const u16x8 *)(X + b + (long long)W * C);
    const u16x8 x11 = *(const u16x8 *)(X + b + (long long)W * C + C);
    u16x8 o;
#pragma unroll
    for (int k = 0; k < 8; ++k)
      o[k] = ws_f2bf(0.25f * ((ws_bf2f(x00[k]) + ws_bf2f(x01[k])) +
                              (ws_bf2f(x10[k]) + ws_bf2f(x11[k]))));
    *(u16x8 *)(P + 8 * i) = o;
  }
}

/* GX[n,h,w,c] += 0.25 * GP[n,h/2,w/2,c] */
__global__ void k_ws_pool_bwd(const float *GP, float *GX, int H, int W, int C, long long n_in) {
  const int H2 = H >> 1, W2 = W >> 1, CV = C >> 2;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < (n_in >> 2);
       i += (long long)gridDim.x * blockDim.x) {
    const int cv = (int)(i % CV);
    const long long pix = i / CV;
    const int w = (int)(pix % W);
    const long long r = pix / W;
    const int h = (int)(r % H);
    const int n = (int)(r / H);
    const float4 g = *(const float4 *)(GP + ((long long)(n * H2 + (h >> 1)) * W2 + (w >> 1)) * C +
                                       4 * cv);
    float4 *d = (float4 *)(GX + 4 * i);
    float4 v = *d;
    v.x += 0.25f * g.x;
    v.y += 0.25f * g.y;
    v.z += 0.25f * g.z;
    v.w += 0.25f * g.w;
    *d = v;
  }
}

/* U[n,h,w,c] = X[n,h/2,w/2,c] ; U is H x W */
__global__ void k_ws_up(const uint16_t *X, uint16_t *U, int H, int W, int C, long long n_out) {
  const int H2 = H >> 1, W2 = W >> 1, CV = C >> 3;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < (n_out >> 3);
       i += (long long)gridDim.x * blockDim.x) {
    const int cv = (int)(i % CV);
    const long long pix = i / CV;
    const int w = (int)(pix % W);
    const long long r = pix / W;
    const int h = (int)(r % H);
    const int n = (int)(r / H);
    *(u16x8 *)(U + 8 * i) =
        *(const u16x8 *)(X + ((long long)(n * H2 + (h >> 1)) * W2 + (w >> 1)) * C + 8 * cv);
  }
}

/* GX[n,h2,w2,c] += sum of GU over the 2x2 block ; GU is H x W */
__global__ void k_ws_up_bwd(const float *GU, float *GX, int H, int W, int C, long long n_in) {
  const int H2 = H >> 1, W2 = W >> 1, CV = C >> 2;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < (n_in >> 2);
       i += (long long)gridDim.x * blockDim.x) {
    const int cv = (int)(i % CV);
    const long long pix = i / CV;
    const int w2 = (int)(pix % W2);
    const long long r = pix / W2;
    const int h2 = (int)(r % H2);
    const int n = (int)(r / H2);
    const long long b = ((long long)(n * H + 2 * h2) * W + 2 * w2) * C + 4 * cv;
    const float4 g00 = *(const float4 *)(GU + b), g01 = *(const float4 *)(GU + b + C);
    const float4 g10 = *(const float4 *)(GU + b + (long long)W * C);
    const float4 g11 = *(const float4 *)(GU + b + (long long)W * C + C);
    float4 *d = (float4 *)(GX + 4 * i);
    float4 v = *d;
    v.x += (g00.x + g01.x) + (g10.x + g11.x);
    v.y += (g00.y + g01.y) + (g10.y + g11.y);
    v.z += (g00.z + g01.z) + (g10.z + g11.z);
    v.w += (g00.w + g01.w) + (g10.w + g11.w);
    *d = v;
  }
}

/* Gm = bf16(G * [Y > 0]); skip (optional) += G * [Y > 0] */
__global__ void k_ws_mask(const float *G, const uint16_t *Y, uint16_t *Gm, float *skip,
                          long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float g = ws_bf2f(Y[i]) > 0.f ? G[i] : 0.f;
    Gm[i] = ws_f2bf(g);
    if (skip) skip[i] += g;
  }
}

/* loss over interior pixels (h, w < S-1): sum (o - t)^2 per block -> part;
 * Gout[p][0] = bf16(2 (o - t) * inv_count), other output channels 0 */
__global__ __launch_bounds__(EW_BLOCK) void k_ws_loss(const float *o, const float *t,
                                                      uint16_t *Gout, double *part, int S,
                                                      long long npix, float inv_count) {
  __shared__ double red[EW_BLOCK];
  double acc = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < npix;
       i += (long long)gridDim.x * blockDim.x) {
    const int pix = (int)(i % ((long long)S * S));
    const int h = pix / S, w = pix - (pix / S) * S;
    float d = 0.f;
    if (h < S - 1 && w < S - 1) d = o[i] - t[i];
    acc += (double)d * (double)d;
    if (Gout) {
      uint16_t *g = Gout + i * WS_OUT_CH;
      g[0] = ws_f2bf(2.f * d * inv_count);
      for (int c = 1; c < WS_OUT_CH; ++c) g[c] = 0;
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = EW_BLOCK / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_ws_adam(float *p, const float *g, float *m, float *v, long long n, float lr,
                          float b1, float b2, float eps, float c1, float c2) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  }
}

/* bf16 weights of one layer: Wt[co][k] (forward) and Wtt[ci][k'] (input
 * gradient: taps mirrored, in/out channels swapped) */
__global__ void k_ws_wprep(const float *p, uint16_t *Wt, uint16_t *Wtt, int Cin, int Cout,
                           int Kpad, int KpadT) {
  const long long n1 = (long long)Cout * Kpad, n2 = Wtt ? (long long)Cin * KpadT : 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n1 + n2;
       i += (long long)gridDim.x * blockDim.x) {
    if (i < n1) {
      const int co = (int)(i / Kpad), k = (int)(i - (long long)co * Kpad);
      float v = 0.f;
      if (k < 9 * Cin) v = p[(long long)co * 9 * Cin + k];
      Wt[i] = ws_f2bf(v);
    } else {
      const long long j = i - n1;
      const int ci = (int)(j / KpadT), k = (int)(j - (long long)ci * KpadT);
      float v = 0.f;
      if (k < 9 * Cout) {
        const int tp = k / Cout, co = k - tp * Cout;
        v = p[(long long)co * 9 * Cin + (8 - tp) * Cin + ci];
      }
      Wtt[j] = ws_f2bf(v);
    }
  }
}

/* training batch: A0[b][pix][0..7] = (ds_in[idx[b]][pix], 0, ...);
 * tgt[b][pix] = ds_out[idx[b]][pix] * inv_gain */
__global__ void k_ws_batch(const uint16_t *ds_in, const float *ds_out, const long long *idx,
                           uint16_t *A0, float *tgt, int S2, int nb, float inv_gain) {
  const long long n = (long long)nb * S2;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int b = (int)(i / S2);
    const int pix = (int)(i - (long long)b * S2);
    const long long s = idx[b];
    u16x8 v = (u16x8){0, 0, 0, 0, 0, 0, 0, 0};
    v[0] = ds_in[s * S2 + pix];
    *(u16x8 *)(A0 + i * WS_IN_CH) = v;
    if (tgt) tgt[i] = ds_out ? ds_out[s * S2 + pix] * inv_gain : 0.f;
  }
}

/* host-supplied (already normalised) inputs: A0[i][0] = bf16(x[i]) */
__global__ void k_ws_input(const float *x, uint16_t *A0, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    u16x8 v = (u16x8){0, 0, 0, 0, 0, 0, 0, 0};
    v[0] = ws_f2bf(x[i]);
    *(u16x8 *)(A0 + i * WS_IN_CH) = v;
  }
}

/* right-hand side of the context's rho: sum of f^2 over interior nodes per
 * block, f = rho * fscale */
__global__ __launch_bounds__(EW_BLOCK) void k_ws_rhs_sumsq(const double *rho, int nx,
                                                           double fscale, double *part) {
  __shared__ double red[EW_BLOCK];
  const int S = nx - 1;
  double acc = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
       i < (long long)(S - 1) * (S - 1); i += (long long)gridDim.x * blockDim.x) {
    const int h = (int)(i / (S - 1)), w = (int)(i - (long long)h * (S - 1));
    const double f = rho[(long long)(h + 1) * nx + (w + 1)] * fscale;
    acc += f * f;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = EW_BLOCK / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

/* scal[0] = s = sqrt(mean f^2) (1 if f == 0), scal[1] = 1/s */
/* sum of part[stride b + v] over b < nb, one work-group of EW_BLOCK threads
 * (thread t takes b = t, t + EW_BLOCK, ... in order, then a fixed tree):
 * deterministic; valid in thread 0 */
__device__ __forceinline__ double ws_group_sum(const double *part, int nb, int stride, int v,
                                               double *red) {
  double t = 0.0;
  for (int b = threadIdx.x; b < nb; b += EW_BLOCK) t += part[(long long)stride * b + v];
  red[threadIdx.x] = t;
  __syncthreads();
  for (int s = EW_BLOCK / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

/* launched with one work-group of EW_BLOCK threads */
__global__ __launch_bounds__(EW_BLOCK) void k_ws_rhs_scale(const double *part, int nb,
                                                           long long count, double *scal) {
  __shared__ double red[EW_BLOCK];
  const double t = ws_group_sum(part, nb, 1, 0, red);
  if (threadIdx.x == 0) {
    double s = sqrt(t / (double)count);
    if (!(s > 0.0)) s = 1.0;
    scal[0] = s;
    scal[1] = 1.0 / s;
  }
}

/* normalised input (and, for capture, the normalised target phi / s) */
__global__ void k_ws_rhs_store(const double *rho, const double *phi, int nx, double fscale,
                               const double *scal, uint16_t *in_plane, uint16_t *A0,
                               float *out_plane) {
  const int S = nx - 1;
  const double inv = scal[1];
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < (long long)S * S;
       i += (long long)gridDim.x * blockDim.x) {
    const int h = (int)(i / S), w = (int)(i - (long long)h * S);
    const bool in = h < S - 1 && w < S - 1;
    const long long node = (long long)(h + 1) * nx + (w + 1);
    const float x = in ? (float)(rho[node] * fscale * inv) : 0.f;
    if (in_plane) in_plane[i] = ws_f2bf(x);
    if (A0) {
      u16x8 v = (u16x8){0, 0, 0, 0, 0, 0, 0, 0};
      v[0] = ws_f2bf(x);
      *(u16x8 *)(A0 + i * WS_IN_CH) = v;
    }
    if (out_plane) out_plane[i] = in ? (float)(phi[node] * inv) : 0.f;
  }
}

/* phi (interior nodes) = o * gain * s */
__global__ void k_ws_phi_out(const float *o, double *phi, int nx, double gain,
                             const double *scal) {
  const int S = nx - 1;
  const double sc = gain * scal[0];
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
       i < (long long)(S - 1) * (S - 1); i += (long long)gridDim.x * blockDim.x) {
    const int h = (int)(i / (S - 1)), w = (int)(i - (long long)h * (S - 1));
    phi[(long long)(h + 1) * nx + (w + 1)] = (double)o[(long long)h * S + w] * sc;
  }
}

/* sum of squares of a float range, per block */
__global__ __launch_bounds__(EW_BLOCK) void k_ws_sumsq(const float *x, long long n,
                                                       double *part) {
  __shared__ double red[EW_BLOCK];
  double acc = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    acc += (double)x[i] * (double)x[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = EW_BLOCK / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

/* ------------------------------------------------------------------------ */
/* energy objective: the network (times gain) followed by k damped Jacobi     */
/* sweeps, scored in the A-norm.  Plane pixels (h, w) < S-1 are the interior  */
/* nodes; A v = 4 v - (4 neighbours), neighbours outside the interior are 0.  */

__device__ __forceinline__ float eg_nb(const float *v, long long i, int h, int w, int S) {
  float s = 0.f;
  if (h > 0) s += v[i - S];
  if (h < S - 2) s += v[i + S];
  if (w > 0) s += v[i - 1];
  if (w < S - 2) s += v[i + 1];
  return s;
}

/* one sweep dst = sc src + om4 (x - A sc src), x = the network input (A0
 * channel 0).  last: dst -= gain t (the error e) and part[nb + b] = the
 * block's sum of (gain t).A(gain t) */
__global__ __launch_bounds__(EW_BLOCK) void k_ws_eg_sweep(const float *src, float *dst,
                                                          const uint16_t *A0, const float *t,
                                                          int S, long long npix, float sc,
                                                          float om4, float gain, int last,
                                                          double *part, int nb) {
  __shared__ double red[EW_BLOCK];
  double acc = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < npix;
       i += (long long)gridDim.x * blockDim.x) {
    const int pix = (int)(i % ((long long)S * S));
    const int h = pix / S, w = pix - (pix / S) * S;
    float v = 0.f;
    if (h < S - 1 && w < S - 1) {
      const float y = sc * src[i];
      const float ay = 4.f * y - sc * eg_nb(src, i, h, w, S);
      const float x = ws_bf2f(A0[i * WS_IN_CH]);
      v = y + om4 * (x - ay);
      if (last) {
        const float tv = gain * t[i];
        v -= tv;
        acc += (double)tv * (double)(4.f * tv - gain * eg_nb(t, i, h, w, S));
      }
    }
    dst[i] = v;
  }
  if (!last) return;
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = EW_BLOCK / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[nb + blockIdx.x] = red[0];
}

/* ae = A e; part[b] = sum e.Ae */
__global__ __launch_bounds__(EW_BLOCK) void k_ws_eg_ae(const float *e, float *ae, int S,
                                                       long long npix, double *part) {
  __shared__ double red[EW_BLOCK];
  double acc = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < npix;
       i += (long long)gridDim.x * blockDim.x) {
    const int pix = (int)(i % ((long long)S * S));
    const int h = pix / S, w = pix - (pix / S) * S;
    float a = 0.f;
    if (h < S - 1 && w < S - 1) {
      a = 4.f * e[i] - eg_nb(e, i, h, w, S);
      acc += (double)e[i] * (double)a;
    }
    ae[i] = a;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = EW_BLOCK / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

/* back through one sweep: h' = coef (h - om4 A h); Gout (non-null, the last
 * sweep) receives bf16(h') in channel 0 and 0 in the other channels, else
 * dst = h' */
__global__ void k_ws_eg_back(const float *hsrc, float *dst, uint16_t *Gout, int S,
                             long long npix, float om4, float coef) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < npix;
       i += (long long)gridDim.x * blockDim.x) {
    const int pix = (int)(i % ((long long)S * S));
    const int h = pix / S, w = pix - (pix / S) * S;
    float g = 0.f;
    if (h < S - 1 && w < S - 1) {
      const float a = hsrc[i];
      g = coef * (a - om4 * (4.f * a - eg_nb(hsrc, i, h, w, S)));
    }
    if (Gout) {
      uint16_t *gp = Gout + i * WS_OUT_CH;
      gp[0] = ws_f2bf(g);
      for (int c = 1; c < WS_OUT_CH; ++c) gp[c] = 0;
    } else {
      dst[i] = g;
    }
  }
}

/* dst = src - om4 A src on the interior (0 outside): one factor of the
 * pre-smoothing polynomial P = (I - om4 A)^k that maps a training pair (x, t)
 * to the pair the network meets after k pre-sweeps, (P x, P t) */
__global__ void k_ws_eg_poly(const float *src, float *dst, int S, long long npix, float om4) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < npix;
       i += (long long)gridDim.x * blockDim.x) {
    const int pix = (int)(i % ((long long)S * S));
    const int h = pix / S, w = pix - (pix / S) * S;
    float v = 0.f;
    if (h < S - 1 && w < S - 1) {
      const float a = src[i];
      v = a - om4 * (4.f * a - eg_nb(src, i, h, w, S));
    }
    dst[i] = v;
  }
}

/* plane <-> channel 0 of the network input */
__global__ void k_ws_a0_get(const uint16_t *A0, float *x, long long npix) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < npix;
       i += (long long)gridDim.x * blockDim.x)
    x[i] = ws_bf2f(A0[i * WS_IN_CH]);
}

/* ------------------------------------------------------------------------ */
/* the network inside the iteration: flexible CG on the interior nodes of     */
/* the network's grid, A x = 4 x - (4 neighbours) = f = dx^2 rho / eps0, walls */
/* 0 (the reference system with its Dirichlet columns eliminated and the rows  */
/* multiplied by -dx^2).  Node k = j nx + i; vectors keep 0 on the walls.      */

#define LS_BLOCK 256
enum { LS_FF = 0, LS_RR = 1, LS_ZQ = 2, LS_PQ = 3, LS_RP = 4, LS_NS = 8 };

__device__ __forceinline__ double ls_Ax(const double *x, int nx, long long k) {
  return 4.0 * x[k] - x[k - 1] - x[k + 1] - x[k - nx] - x[k + nx];
}

/* interior node of flat index q < (nx-2)^2 */
__device__ __forceinline__ long long ls_node(long long q, int nx) {
  const int m = nx - 2;
  const long long j = q / m;
  return (j + 1) * nx + (q - j * m) + 1;
}

/* block sum, valid in thread 0 (wave shuffles, then the waves in order) */
__device__ __forceinline__ double ls_block_sum(double v, double *red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int q = 0; q < LS_BLOCK / 64; ++q) s += red[q];
  return s;
}

/* x = phi (or 0), r = f - A x; partials of f.f and r.r */
__global__ __launch_bounds__(LS_BLOCK) void k_ls_init(const double *rho, const double *phi,
                                                      int nx, double fscale, double *x,
                                                      double *r, double *part) {
  __shared__ double red[LS_BLOCK / 64];
  const long long n = (long long)(nx - 2) * (nx - 2);
  double ff = 0.0, rr = 0.0;
  for (long long q = blockIdx.x * (long long)LS_BLOCK + threadIdx.x; q < n;
       q += (long long)gridDim.x * LS_BLOCK) {
    const long long k = ls_node(q, nx);
    const double f = rho[k] * fscale;
    const double xv = phi ? phi[k] : 0.0;
    const double rv = phi ? f - ls_Ax(phi, nx, k) : f;
    x[k] = xv;
    r[k] = rv;
    ff += f * f;
    rr += rv * rv;
  }
  ff = ls_block_sum(ff, red);
  rr = ls_block_sum(rr, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = ff;
    part[2 * blockIdx.x + 1] = rr;
  }
}

/* out[v] = sum over blocks of part[nv b + v], v < nv (one work-group of
 * EW_BLOCK threads, fixed order) */
__global__ __launch_bounds__(EW_BLOCK) void k_ls_reduce(const double *part, int nb, int nv,
                                                        double *out) {
  __shared__ double red[EW_BLOCK];
  for (int v = 0; v < nv; ++v) {
    const double t = ws_group_sum(part, nb, nv, v, red);
    if (threadIdx.x == 0) out[v] = t;
  }
}

/* zs = z + om4 (r - A z): one damped Jacobi sweep (z null: from z = 0) */
__global__ __launch_bounds__(LS_BLOCK) void k_ls_smooth(const double *r, const double *z,
                                                        double *zs, int nx, double om4) {
  const long long n = (long long)(nx - 2) * (nx - 2);
  for (long long q = blockIdx.x * (long long)LS_BLOCK + threadIdx.x; q < n;
       q += (long long)gridDim.x * LS_BLOCK) {
    const long long k = ls_node(q, nx);
    zs[k] = z ? z[k] + om4 * (r[k] - ls_Ax(z, nx, k)) : om4 * r[k];
  }
}

/* rp = r - A z: what the pre-sweeps left for the network */
__global__ __launch_bounds__(LS_BLOCK) void k_ls_resid(const double *r, const double *z,
                                                       double *rp, int nx) {
  const long long n = (long long)(nx - 2) * (nx - 2);
  for (long long q = blockIdx.x * (long long)LS_BLOCK + threadIdx.x; q < n;
       q += (long long)gridDim.x * LS_BLOCK) {
    const long long k = ls_node(q, nx);
    rp[k] = r[k] - ls_Ax(z, nx, k);
  }
}

/* z += c: the network's correction added to the pre-sweeps' iterate */
__global__ __launch_bounds__(LS_BLOCK) void k_ls_add(double *z, const double *c, int nx) {
  const long long n = (long long)(nx - 2) * (nx - 2);
  for (long long q = blockIdx.x * (long long)LS_BLOCK + threadIdx.x; q < n;
       q += (long long)gridDim.x * LS_BLOCK) {
    const long long k = ls_node(q, nx);
    z[k] += c[k];
  }
}

/* partials of a.b */
__global__ __launch_bounds__(LS_BLOCK) void k_ls_dot(const double *a, const double *b, int nx,
                                                     double *part) {
  __shared__ double red[LS_BLOCK / 64];
  const long long n = (long long)(nx - 2) * (nx - 2);
  double s = 0.0;
  for (long long q = blockIdx.x * (long long)LS_BLOCK + threadIdx.x; q < n;
       q += (long long)gridDim.x * LS_BLOCK) {
    const long long k = ls_node(q, nx);
    s += a[k] * b[k];
  }
  s = ls_block_sum(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

/* p = z + beta p_prev, beta = -(z.q_prev)/(p_prev.q_prev) (none in the first
 * iteration): the new direction A-orthogonal to the previous one */
__global__ __launch_bounds__(LS_BLOCK) void k_ls_dir(const double *z, const double *pprev,
                                                     double *p, int nx, const double *s,
                                                     int first) {
  const long long n = (long long)(nx - 2) * (nx - 2);
  const double beta = first ? 0.0 : -s[LS_ZQ] / s[LS_PQ];
  for (long long q = blockIdx.x * (long long)LS_BLOCK + threadIdx.x; q < n;
       q += (long long)gridDim.x * LS_BLOCK) {
    const long long k = ls_node(q, nx);
    p[k] = first ? z[k] : z[k] + beta * pprev[k];
  }
}

/* q = A p; partials of p.q and r.p */
__global__ __launch_bounds__(LS_BLOCK) void k_ls_Ap(const double *p, const double *r, double *qv,
                                                    int nx, double *part) {
  __shared__ double red[LS_BLOCK / 64];
  const long long n = (long long)(nx - 2) * (nx - 2);
  double pq = 0.0, rp = 0.0;
  for (long long q = blockIdx.x * (long long)LS_BLOCK + threadIdx.x; q < n;
       q += (long long)gridDim.x * LS_BLOCK) {
    const long long k = ls_node(q, nx);
    const double a = ls_Ax(p, nx, k);
    qv[k] = a;
    pq += p[k] * a;
    rp += r[k] * p[k];
  }
  pq = ls_block_sum(pq, red);
  rp = ls_block_sum(rp, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = pq;
    part[2 * blockIdx.x + 1] = rp;
  }
}

/* alpha = (r.p)/(p.q); x += alpha p, r -= alpha q; partials of r.r */
__global__ __launch_bounds__(LS_BLOCK) void k_ls_xr(double *x, double *r, const double *p,
                                                    const double *qv, int nx, const double *s,
                                                    double *part) {
  __shared__ double red[LS_BLOCK / 64];
  const long long n = (long long)(nx - 2) * (nx - 2);
  const double alpha = s[LS_RP] / s[LS_PQ];
  double rr = 0.0;
  for (long long q = blockIdx.x * (long long)LS_BLOCK + threadIdx.x; q < n;
       q += (long long)gridDim.x * LS_BLOCK) {
    const long long k = ls_node(q, nx);
    x[k] = __builtin_fma(alpha, p[k], x[k]);
    const double rv = __builtin_fma(-alpha, qv[k], r[k]);
    r[k] = rv;
    rr += rv * rv;
  }
  rr = ls_block_sum(rr, red);
  if (threadIdx.x == 0) part[blockIdx.x] = rr;
}

/* phi (interior nodes) = x */
__global__ __launch_bounds__(LS_BLOCK) void k_ls_out(const double *x, double *phi, int nx) {
  const long long n = (long long)(nx - 2) * (nx - 2);
  for (long long q = blockIdx.x * (long long)LS_BLOCK + threadIdx.x; q < n;
       q += (long long)gridDim.x * LS_BLOCK) {
    const long long k = ls_node(q, nx);
    phi[k] = x[k];
  }
}

/* a training pair from an FCG iterate: input r / s, target (phi - x) / s at
 * the interior nodes (s = scal[0], the RMS of r), plane layout of capture */
__global__ void k_ws_pair_store(const double *r, const double *phi, const double *x, int nx,
                                const double *scal, uint16_t *in_plane, float *out_plane) {
  const int S = nx - 1;
  const double inv = scal[1];
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < (long long)S * S;
       i += (long long)gridDim.x * blockDim.x) {
    const int h = (int)(i / S), w = (int)(i - (long long)h * S);
    const bool in = h < S - 1 && w < S - 1;
    const long long node = (long long)(h + 1) * nx + (w + 1);
    in_plane[i] = ws_f2bf(in ? (float)(r[node] * inv) : 0.f);
    out_plane[i] = in ? (float)((phi[node] - x[node]) * inv) : 0.f;
  }
}

/* ------------------------------------------------------------------------ */
/* host side                                                                  */

struct WsConv {
  int cin, cout, kpad, kpadt, level, relu;
  long long woff;  /* offset in the parameter vector ([co][tap][ci]) */
  uint16_t *wt = nullptr, *wtt = nullptr;
};

struct WsTensor {
  int level = 0, C = 0;
  uint16_t *a = nullptr; /* bf16 activations */
  float *g = nullptr;    /* fp32 gradient */
};

struct iskra_ws {
  iskra_ctx *ctx = nullptr;
  int L = 0, base = 0, S = 0, maxb = 0;
  std::vector<WsConv> conv;
  long long nparam = 0;
  float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr;
  long long adam_t = 0;
  double gain = 1.0;
  WsTensor A0, Ea[WS_MAX_LEVELS], E[WS_MAX_LEVELS], P[WS_MAX_LEVELS], U[WS_MAX_LEVELS],
      Da[WS_MAX_LEVELS], D[WS_MAX_LEVELS];
  float *out = nullptr, *tgt = nullptr;
  uint16_t *gout = nullptr, *gm = nullptr;
  long long gm_elems = 0;
  float *part = nullptr;
  double *dpart = nullptr, *scal = nullptr;
  int dpart_n = 0;
  long long *idx = nullptr;
  float *xin = nullptr;  /* host-input staging */
  /* dataset */
  uint16_t *ds_in = nullptr;
  float *ds_out = nullptr;
  long long ds_cap = 0, ds_n = 0;
  /* training objective (ISKRA_WS_*), its Jacobi weight / 4 and planes */
  int obj = ISKRA_WS_PHI_MSE;
  double om4 = 0.0;
  int sweeps = 1;
  float *eg_e = nullptr, *eg_ae = nullptr;
  /* the network-preconditioned solve (node grid, FP64; walls stay 0) */
  double *lx = nullptr, *lr = nullptr, *lz = nullptr, *lzs = nullptr, *lp[2] = {nullptr, nullptr},
         *lq[2] = {nullptr, nullptr}, *lpart = nullptr, *ls = nullptr, *lpre[2] = {nullptr, nullptr},
         *lrp = nullptr;
};

static int ws_side(const iskra_ws *w, int l) { return w->S >> l; }
static long long ws_npix(const iskra_ws *w, int l, int b) {
  return (long long)b * ws_side(w, l) * ws_side(w, l);
}
static unsigned ws_grid(long long n) {
  long long b = (n + EW_BLOCK - 1) / EW_BLOCK;
  return (unsigned)std::max(1ll, std::min(b, 8192ll));
}

static int ws_tensor(iskra_ws *w, WsTensor *t, int level, int C, bool grad) {
  t->level = level;
  t->C = C;
  const long long n = ws_npix(w, level, w->maxb) * C;
  WRC(isk_alloc((void **)&t->a, n * 2));
  if (grad) WRC(isk_alloc((void **)&t->g, n * 4));
  w->gm_elems = std::max(w->gm_elems, n);
  return ISKRA_OK;
}

static void ws_free_tensor(WsTensor *t) {
  isk_free(t->a);
  isk_free(t->g);
  t->a = nullptr;
  t->g = nullptr;
}

/* channels of level l: base * 2^l, held at base * 8 from level 3 on (so a
 * deep network keeps its coarse levels affordable) */
static int ws_chan(const iskra_ws *w, int l) { return w->base << std::min(l, 3); }

/* layer order: enc0A, enc0B, ..., enc(L-1)B, dec(L-2)A, dec(L-2)B, ..., dec0B, out */
static int ws_enc(int l, int b) { return 2 * l + b; }
static int ws_dec(const iskra_ws *w, int l, int b) { return 2 * w->L + 2 * (w->L - 2 - l) + b; }
static int ws_out(const iskra_ws *w) { return 2 * w->L + 2 * (w->L - 1); }

static void ws_add_conv(iskra_ws *w, int cin, int cout, int level, int relu) {
  WsConv c;
  c.cin = cin;
  c.cout = cout;
  c.kpad = (9 * cin + 31) / 32 * 32;
  c.kpadt = (9 * cout + 31) / 32 * 32;
  c.level = level;
  c.relu = relu;
  c.woff = w->nparam;
  w->nparam += (long long)cout * 9 * cin;
  w->conv.push_back(c);
}

static int ws_launch_conv(iskra_ws *w, const WsConv &c, int b, const uint16_t *X,
                          const uint16_t *Wt, int cin, int cout, int kpad, const uint16_t *skip,
                          uint16_t *Y, float *Yf, int relu, int mode) {
  ConvArgs a;
  a.X = X;
  a.Wt = Wt;
  a.skip = skip;
  a.Y = Y;
  a.Yf = Yf;
  a.H = a.W = ws_side(w, c.level);
  a.Cin = cin;
  a.Cout = cout;
  a.Kpad = kpad;
  a.relu = relu;
  a.mode = mode;
  const long long npix = ws_npix(w, c.level, b);
  if (npix % 64 || cin % 8 || cout % 16 || kpad % 32) {
    isk_set_error("warm start: conv shape %d->%d on %lld pixels not supported", cin, cout, npix);
    return ISKRA_E_ARG;
  }
  /* tile (TM output channels, NP pixel groups per wave): the first of the
   * candidates, in order of fewer loads per MFMA, that gives at least
   * WS_CONV_MIN_WG work-groups; else the one with the most */
  static const int cand[6][2] = {{64, 4}, {32, 4}, {16, 4}, {64, 1}, {32, 1}, {16, 1}};
  int tm = 16, np = 1;
  long long best = -1;
  for (int k = 0; k < 6; ++k) {
    const int ct = cand[k][0], cn = cand[k][1];
    if (cout % ct || npix % (64 * cn)) continue;
    const long long wg = (npix / (64 * cn)) * (cout / ct);
    if (wg >= WS_CONV_MIN_WG) {
      tm = ct;
      np = cn;
      break;
    }
    if (wg > best) {
      best = wg;
      tm = ct;
      np = cn;
    }
  }
  dim3 grid((unsigned)(npix / (64 * np)), (unsigned)(cout / tm));
  hipStream_t st = w->ctx->stream;
  if (WS_CONV_LDS && np == 4 && cin % 32 == 0 && a.H % 16 == 0 && a.W % 16 == 0) {
    /* the LDS-staged kernel: one 16 x 16 tile per work-group (npix / 256) */
    if (tm == 64) hipLaunchKernelGGL((k_ws_conv_lds<64>), grid, dim3(256), 0, st, a);
    else if (tm == 32) hipLaunchKernelGGL((k_ws_conv_lds<32>), grid, dim3(256), 0, st, a);
    else hipLaunchKernelGGL((k_ws_conv_lds<16>), grid, dim3(256), 0, st, a);
  } else if (np == 4) {
    if (tm == 64) hipLaunchKernelGGL((k_ws_conv<64, 4>), grid, dim3(256), 0, st, a);
    else if (tm == 32) hipLaunchKernelGGL((k_ws_conv<32, 4>), grid, dim3(256), 0, st, a);
    else hipLaunchKernelGGL((k_ws_conv<16, 4>), grid, dim3(256), 0, st, a);
  } else {
    if (tm == 64) hipLaunchKernelGGL((k_ws_conv<64, 1>), grid, dim3(256), 0, st, a);
    else if (tm == 32) hipLaunchKernelGGL((k_ws_conv<32, 1>), grid, dim3(256), 0, st, a);
    else hipLaunchKernelGGL((k_ws_conv<16, 1>), grid, dim3(256), 0, st, a);
  }
  ISK_CHECK_LAUNCH();
  return ISKRA_OK;
}

static int ws_fwd_conv(iskra_ws *w, int layer, int b, const WsTensor &in, const uint16_t *skip,
                       uint16_t *Y, float *Yf, int mode) {
  const WsConv &c = w->conv[layer];
  return ws_launch_conv(w, c, b, in.a, c.wt, c.cin, c.cout, c.kpad, skip, Y, Yf, c.relu, mode);
}

/* input gradient: GX += conv(Gm, W mirrored) */
static int ws_dgrad(iskra_ws *w, int layer, int b, float *GX) {
  const WsConv &c = w->conv[layer];
  return ws_launch_conv(w, c, b, w->gm, c.wtt, c.cout, c.cin, c.kpadt, nullptr, nullptr, GX, 0,
                        1);
}

static int ws_wgrad(iskra_ws *w, int layer, int b, const uint16_t *X) {
  const WsConv &c = w->conv[layer];
  const long long npix = ws_npix(w, c.level, b);
  const long long nseg = npix / 64;
  const long long K = 9ll * c.cin;
  const long long per_chunk = (long long)c.cout * K;
  const int gy = (c.cout + 31) / 32, gz = (c.cin + 15) / 16;
  long long max_chunks = std::max(1ll, (WS_PART_BYTES / 4) / per_chunk);
  long long want = std::max(1ll, (long long)(1024 / (gy * gz)));
  long long nchunk = std::min(std::min(max_chunks, want), nseg);
  const int spc = (int)((nseg + nchunk - 1) / nchunk);
  nchunk = (nseg + spc - 1) / spc;
  WgradArgs a;
  a.X = X;
  a.G = w->gm;
  a.part = w->part;
  a.H = a.W = ws_side(w, c.level);
  a.Cin = c.cin;
  a.Cout = c.cout;
  a.nseg = nseg;
  a.seg_per_chunk = spc;
  hipLaunchKernelGGL(k_ws_wgrad, dim3((unsigned)nchunk, gy, gz), dim3(256), 0, w->ctx->stream,
                     a);
  ISK_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_ws_reduce, dim3(ws_grid(per_chunk)), dim3(EW_BLOCK), 0, w->ctx->stream,
                     (const float *)w->part, w->grads + c.woff, per_chunk, (int)nchunk);
  ISK_CHECK_LAUNCH();
  return ISKRA_OK;
}

static int ws_mask(iskra_ws *w, const WsTensor &t, int b, float *skip) {
  const long long n = ws_npix(w, t.level, b) * t.C;
  hipLaunchKernelGGL(k_ws_mask, dim3(ws_grid(n)), dim3(EW_BLOCK), 0, w->ctx->stream,
                     (const float *)t.g, (const uint16_t *)t.a, w->gm, skip, n);
  ISK_CHECK_LAUNCH();
  return ISKRA_OK;
}

static int ws_zero(iskra_ws *w, float *p, long long n) {
  hipLaunchKernelGGL(k_ws_fill, dim3(ws_grid(n)), dim3(EW_BLOCK), 0, w->ctx->stream, p, n, 0.f);
  ISK_CHECK_LAUNCH();
  return ISKRA_OK;
}

static const WsTensor &ws_dec_out(const iskra_ws *w, int l) {
  return l == w->L - 1 ? w->E[l] : w->D[l];
}

/* forward of a batch already in A0; output plane in w->out */
static int ws_forward(iskra_ws *w, int b) {
  hipStream_t st = w->ctx->stream;
  const int L = w->L;
  for (int l = 0; l < L; ++l) {
    if (l > 0) {
      const WsTensor &src = w->E[l - 1];
      const long long n = ws_npix(w, l, b) * src.C;
      const int Hs = ws_side(w, l - 1);
      hipLaunchKernelGGL(k_ws_pool, dim3(ws_grid(n >> 3)), dim3(EW_BLOCK), 0, st,
                         (const uint16_t *)src.a, w->P[l].a, Hs, Hs, src.C, n);
      ISK_CHECK_LAUNCH();
    }
    WRC(ws_fwd_conv(w, ws_enc(l, 0), b, l == 0 ? w->A0 : w->P[l], nullptr, w->Ea[l].a, nullptr,
                    0));
    WRC(ws_fwd_conv(w, ws_enc(l, 1), b, w->Ea[l], nullptr, w->E[l].a, nullptr, 0));
  }
  for (int l = L - 2; l >= 0; --l) {
    const WsTensor &src = ws_dec_out(w, l + 1);
    const long long n = ws_npix(w, l, b) * src.C;
    const int H = ws_side(w, l);
    hipLaunchKernelGGL(k_ws_up, dim3(ws_grid(n >> 3)), dim3(EW_BLOCK), 0, st, (const uint16_t *)src.a,
                       w->U[l].a, H, H, src.C, n);
    ISK_CHECK_LAUNCH();
    WRC(ws_fwd_conv(w, ws_dec(w, l, 0), b, w->U[l], w->E[l].a, w->Da[l].a, nullptr, 0));
    WRC(ws_fwd_conv(w, ws_dec(w, l, 1), b, w->Da[l], nullptr, w->D[l].a, nullptr, 0));
  }
  const WsTensor &last = L > 1 ? w->D[0] : w->E[0];
  return ws_fwd_conv(w, ws_out(w), b, last, nullptr, nullptr, w->out, 2);
}

static int ws_pre(iskra_ws *w, int b);

/* energy objective on a batch (x, t) already mapped to (P x, P t) by
 * ws_pre: z = k post-sweeps from gain * out, e = z - gain t (= P (gain out
 * - gain t), the error of the whole two-level cycle), loss = sum e.Ae /
 * sum (gain t).A(gain t) over the batch; with_grad: Gout = its gradient with
 * respect to the output plane, 2 gain (I - om4 A)^k A e / denominator.
 * Planes ping-pong between eg_e and eg_ae. */
static int ws_energy_loss(iskra_ws *w, int b, bool with_grad, double *loss) {
  hipStream_t st = w->ctx->stream;
  const long long npix = ws_npix(w, 0, b);
  const unsigned nb = std::min(ws_grid(npix), (unsigned)(w->dpart_n / 2));
  /* no post-sweeps (sweeps = 0, or the network's own error scored):
   * one pass with om4 = 0 forms e = gain (o - t) */
  const bool post = w->obj == ISKRA_WS_ENERGY && w->sweeps > 0;
  const int ks = post ? w->sweeps : 1;
  const float gain = (float)w->gain, om4 = post ? (float)w->om4 : 0.f;
  float *buf[2] = {w->eg_e, w->eg_ae};
  const float *src = w->out;
  int cur = 0;
  for (int k = 0; k < ks; ++k) {
    const int last = k == ks - 1;
    hipLaunchKernelGGL(k_ws_eg_sweep, dim3(nb), dim3(EW_BLOCK), 0, st, src, buf[cur],
                       (const uint16_t *)w->A0.a, (const float *)w->tgt, w->S, npix,
                       k == 0 ? gain : 1.f, om4, gain, last, w->dpart, (int)nb);
    ISK_CHECK_LAUNCH();
    src = buf[cur];
    cur ^= 1;
  }
  /* e in buf[cur ^ 1]; A e into buf[cur] */
  hipLaunchKernelGGL(k_ws_eg_ae, dim3(nb), dim3(EW_BLOCK), 0, st, src, buf[cur], w->S, npix,
                     w->dpart);
  ISK_CHECK_LAUNCH();
  std::vector<double> h(2 * nb);
  ISK_HIP(hipMemcpyAsync(h.data(), w->dpart, 2 * nb * 8, hipMemcpyDeviceToHost, st));
  ISK_HIP(hipStreamSynchronize(st));
  double num = 0.0, den = 0.0;
  for (unsigned i = 0; i < nb; ++i) {
    num += h[i];
    den += h[nb + i];
  }
  if (!(den > 0.0)) den = 1.0;
  if (loss) *loss = num / den;
  if (with_grad) {
    /* the sweeps' adjoint, applied k times (A is symmetric); the first
     * carries the factor 2 gain / den */
    for (int k = 0; k < ks; ++k) {
      const int last = k == ks - 1;
      hipLaunchKernelGGL(k_ws_eg_back, dim3(ws_grid(npix)), dim3(EW_BLOCK), 0, st,
                         (const float *)buf[cur], buf[cur ^ 1], last ? w->gout : nullptr, w->S,
                         npix, om4, k == 0 ? (float)(2.0 * w->gain / den) : 1.f);
      ISK_CHECK_LAUNCH();
      cur ^= 1;
    }
  }
  return ISKRA_OK;
}

/* loss of w->out against w->tgt; with_grad: also Gout; returns the mean */
static int ws_loss(iskra_ws *w, int b, bool with_grad, double *loss) {
  if (w->obj != ISKRA_WS_PHI_MSE) return ws_energy_loss(w, b, with_grad, loss);
  const long long npix = ws_npix(w, 0, b);
  const long long count = (long long)b * (w->S - 1) * (w->S - 1);
  const unsigned nb = std::min(ws_grid(npix), (unsigned)w->dpart_n);
  hipLaunchKernelGGL(k_ws_loss, dim3(nb), dim3(EW_BLOCK), 0, w->ctx->stream,
                     (const float *)w->out, (const float *)w->tgt, with_grad ? w->gout : nullptr,
                     w->dpart, w->S, npix, (float)(1.0 / (double)count));
  ISK_CHECK_LAUNCH();
  std::vector<double> h(nb);
  ISK_HIP(hipMemcpyAsync(h.data(), w->dpart, nb * 8, hipMemcpyDeviceToHost, w->ctx->stream));
  ISK_HIP(hipStreamSynchronize(w->ctx->stream));
  double s = 0.0;
  for (unsigned i = 0; i < nb; ++i) s += h[i];
  if (loss) *loss = s / (double)count;
  return ISKRA_OK;
}

/* backward after ws_forward + ws_loss(with_grad): parameter gradients in w->grads */
static int ws_backward(iskra_ws *w, int b) {
  hipStream_t st = w->ctx->stream;
  const int L = w->L;
  WsTensor *all[] = {w->Ea, w->E, w->P, w->U, w->Da, w->D};
  for (WsTensor *arr : all)
    for (int l = 0; l < L; ++l)
      if (arr[l].g) WRC(ws_zero(w, arr[l].g, ws_npix(w, arr[l].level, b) * arr[l].C));
  /* output layer (linear): its masked gradient is Gout itself */
  const WsTensor &last = L > 1 ? w->D[0] : w->E[0];
  {
    const WsConv &c = w->conv[ws_out(w)];
    const long long npix = ws_npix(w, 0, b);
    /* Gm <- Gout (16 channels) */
    ISK_HIP(hipMemcpyAsync(w->gm, w->gout, npix * WS_OUT_CH * 2, hipMemcpyDeviceToDevice, st));
    WRC(ws_wgrad(w, ws_out(w), b, last.a));
    WRC(ws_launch_conv(w, c, b, w->gm, c.wtt, c.cout, c.cin, c.kpadt, nullptr, nullptr, last.g,
                       0, 1));
  }
  for (int l = 0; l <= L - 2; ++l) {
    WRC(ws_mask(w, w->D[l], b, nullptr));
    WRC(ws_wgrad(w, ws_dec(w, l, 1), b, w->Da[l].a));
    WRC(ws_dgrad(w, ws_dec(w, l, 1), b, w->Da[l].g));
    WRC(ws_mask(w, w->Da[l], b, w->E[l].g));
    WRC(ws_wgrad(w, ws_dec(w, l, 0), b, w->U[l].a));
    WRC(ws_dgrad(w, ws_dec(w, l, 0), b, w->U[l].g));
    const WsTensor &dst = ws_dec_out(w, l + 1);
    const long long n = ws_npix(w, l + 1, b) * dst.C;
    const int H = ws_side(w, l);
    hipLaunchKernelGGL(k_ws_up_bwd, dim3(ws_grid(n >> 2)), dim3(EW_BLOCK), 0, st,
                       (const float *)w->U[l].g, dst.g, H, H, dst.C, n);
    ISK_CHECK_LAUNCH();
  }
  for (int l = L - 1; l >= 0; --l) {
    WRC(ws_mask(w, w->E[l], b, nullptr));
    WRC(ws_wgrad(w, ws_enc(l, 1), b, w->Ea[l].a));
    WRC(ws_dgrad(w, ws_enc(l, 1), b, w->Ea[l].g));
    WRC(ws_mask(w, w->Ea[l], b, nullptr));
    WRC(ws_wgrad(w, ws_enc(l, 0), b, l == 0 ? w->A0.a : w->P[l].a));
    if (l > 0) {
      WRC(ws_dgrad(w, ws_enc(l, 0), b, w->P[l].g));
      const WsTensor &dst = w->E[l - 1];
      const long long n = ws_npix(w, l - 1, b) * dst.C;
      const int H = ws_side(w, l - 1);
      hipLaunchKernelGGL(k_ws_pool_bwd, dim3(ws_grid(n >> 2)), dim3(EW_BLOCK), 0, st,
                         (const float *)w->P[l].g, dst.g, H, H, dst.C, n);
      ISK_CHECK_LAUNCH();
    }
  }
  return ISKRA_OK;
}

static int ws_prep_weights(iskra_ws *w) {
  for (size_t i = 0; i < w->conv.size(); ++i) {
    WsConv &c = w->conv[i];
    const long long n = (long long)c.cout * c.kpad + (c.wtt ? (long long)c.cin * c.kpadt : 0);
    hipLaunchKernelGGL(k_ws_wprep, dim3(ws_grid(n)), dim3(EW_BLOCK), 0, w->ctx->stream,
                       (const float *)(w->params + c.woff), c.wt, c.wtt, c.cin, c.cout, c.kpad,
                       c.kpadt);
    ISK_CHECK_LAUNCH();
  }
  return ISKRA_OK;
}

static void ws_release(iskra_ws *w) {
  for (WsConv &c : w->conv) {
    isk_free(c.wt);
    isk_free(c.wtt);
  }
  ws_free_tensor(&w->A0);
  for (int l = 0; l < WS_MAX_LEVELS; ++l) {
    ws_free_tensor(&w->Ea[l]);
    ws_free_tensor(&w->E[l]);
    ws_free_tensor(&w->P[l]);
    ws_free_tensor(&w->U[l]);
    ws_free_tensor(&w->Da[l]);
    ws_free_tensor(&w->D[l]);
  }
  void *ptrs[] = {w->params, w->grads, w->m,     w->v,     w->out,    w->tgt,   w->gout,
                  w->gm,     w->part,  w->dpart, w->scal,  w->idx,    w->xin,   w->ds_in,
                  w->ds_out, w->eg_e,  w->eg_ae, w->lx,    w->lr,     w->lz,    w->lzs,
                  w->lp[0],  w->lp[1], w->lq[0], w->lq[1], w->lpart,  w->ls,    w->lpre[0],
                  w->lpre[1], w->lrp};
  for (void *p : ptrs) isk_free(p);
}

static int ws_check(iskra_ws *w) {
  if (!w || !w->ctx) {
    isk_set_error("warm start: null network");
    return ISKRA_E_ARG;
  }
  isk_clear_error();
  ISK_HIP(hipSetDevice(w->ctx->device));
  return ISKRA_OK;
}

/* the context's grid must be the network's: square, 2-D, (S+1)^2 nodes */
static int ws_check_grid(iskra_ws *w) {
  const iskra_ctx *c = w->ctx;
  if (!c->have_grid || c->g.dim != 2 || c->g.nx != w->S + 1 || c->g.ny != w->S + 1) {
    isk_set_error("warm start: the context grid must be 2-D with %d x %d nodes", w->S + 1,
                  w->S + 1);
    return ISKRA_E_STATE;
  }
  if (c->g.axial) {
    isk_set_error("warm start: the network maps Cartesian grids, not an axial one");
    return ISKRA_E_UNSUPPORTED;
  }
  return ISKRA_OK;
}

/* f = rho * dx^2 / eps0 of the context -> scale in w->scal */
static int ws_rhs_scale(iskra_ws *w, double *fscale) {
  iskra_ctx *c = w->ctx;
  *fscale = c->g.dx * c->g.dx / c->ps.eps0;
  const long long cnt = (long long)(w->S - 1) * (w->S - 1);
  const unsigned nb = std::min(ws_grid(cnt), (unsigned)w->dpart_n);
  hipLaunchKernelGGL(k_ws_rhs_sumsq, dim3(nb), dim3(EW_BLOCK), 0, c->stream,
                     (const double *)c->rho, c->g.nx, *fscale, w->dpart);
  ISK_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_ws_rhs_scale, dim3(1), dim3(EW_BLOCK), 0, c->stream, (const double *)w->dpart,
                     (int)nb, cnt, w->scal);
  ISK_CHECK_LAUNCH();
  return ISKRA_OK;
}

extern "C" {

int iskra_ws_create(iskra_ctx *ctx, int levels, int base, int side, int max_batch,
                    iskra_ws **out) {
  if (!ctx || !out) {
    isk_set_error("ws_create: null argument");
    return ISKRA_E_ARG;
  }
  ISK_HIP(hipSetDevice(ctx->device));
  isk_clear_error();
  *out = nullptr;
  if (levels < 1 || levels > WS_MAX_LEVELS || base < 16 || base % 16 || side < 16 ||
      (side & (side - 1)) || (side >> (levels - 1)) < 8 || max_batch < 1 || max_batch > 1024) {
    isk_set_error("ws_create: levels in [1,%d], base a multiple of 16, side a power of two with "
                  "side/2^(levels-1) >= 8, max_batch in [1,1024]", WS_MAX_LEVELS);
    return ISKRA_E_ARG;
  }
  iskra_ws *w = new iskra_ws();
  w->ctx = ctx;
  w->L = levels;
  w->base = base;
  w->S = side;
  w->maxb = max_batch;
  for (int l = 0; l < levels; ++l) {
    ws_add_conv(w, l == 0 ? WS_IN_CH : ws_chan(w, l - 1), ws_chan(w, l), l, 1);
    ws_add_conv(w, ws_chan(w, l), ws_chan(w, l), l, 1);
  }
  for (int l = levels - 2; l >= 0; --l) {
    ws_add_conv(w, ws_chan(w, l + 1), ws_chan(w, l), l, 1);
    ws_add_conv(w, ws_chan(w, l), ws_chan(w, l), l, 1);
  }
  ws_add_conv(w, ws_chan(w, 0), WS_OUT_CH, 0, 0);
  int rc = ISKRA_OK;
#define WTRY(x)                  \
  do {                           \
    rc = (x);                    \
    if (rc) goto fail;           \
  } while (0)
  for (size_t i = 0; i < w->conv.size(); ++i) {
    WsConv &c = w->conv[i];
    WTRY(isk_alloc((void **)&c.wt, (size_t)c.cout * c.kpad * 2));
    if (i != 0) WTRY(isk_alloc((void **)&c.wtt, (size_t)c.cin * c.kpadt * 2));
  }
  WTRY(isk_alloc((void **)&w->params, w->nparam * 4));
  WTRY(isk_alloc((void **)&w->grads, w->nparam * 4));
  WTRY(isk_alloc((void **)&w->m, w->nparam * 4));
  WTRY(isk_alloc((void **)&w->v, w->nparam * 4));
  WTRY(ws_tensor(w, &w->A0, 0, WS_IN_CH, false));
  for (int l = 0; l < levels; ++l) {
    WTRY(ws_tensor(w, &w->Ea[l], l, ws_chan(w, l), true));
    WTRY(ws_tensor(w, &w->E[l], l, ws_chan(w, l), true));
    if (l > 0) WTRY(ws_tensor(w, &w->P[l], l, ws_chan(w, l - 1), true));
    if (l <= levels - 2) {
      WTRY(ws_tensor(w, &w->U[l], l, ws_chan(w, l + 1), true));
      WTRY(ws_tensor(w, &w->Da[l], l, ws_chan(w, l), true));
      WTRY(ws_tensor(w, &w->D[l], l, ws_chan(w, l), true));
    }
  }
  {
    const long long npix = ws_npix(w, 0, max_batch);
    w->gm_elems = std::max(w->gm_elems, npix * WS_OUT_CH);
    WTRY(isk_alloc((void **)&w->out, npix * 4));
    WTRY(isk_alloc((void **)&w->tgt, npix * 4));
    WTRY(isk_alloc((void **)&w->xin, npix * 4));
    WTRY(isk_alloc((void **)&w->gout, npix * WS_OUT_CH * 2));
    WTRY(isk_alloc((void **)&w->eg_e, npix * 4));
    WTRY(isk_alloc((void **)&w->eg_ae, npix * 4));
    WTRY(isk_alloc((void **)&w->gm, w->gm_elems * 2));
    WTRY(isk_alloc((void **)&w->part, WS_PART_BYTES));
    w->dpart_n = 8192;
    WTRY(isk_alloc((void **)&w->dpart, w->dpart_n * 8));
    WTRY(isk_alloc((void **)&w->scal, 64));
    WTRY(isk_alloc((void **)&w->idx, (size_t)max_batch * 8));
  }
  WTRY(ws_zero(w, w->params, w->nparam));
  WTRY(ws_zero(w, w->m, w->nparam));
  WTRY(ws_zero(w, w->v, w->nparam));
  WTRY(ws_zero(w, w->grads, w->nparam));
  WTRY(ws_prep_weights(w));
  {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
      rc = isk_hip_fail(e, "hipStreamSynchronize", __FILE__, __LINE__);
      goto fail;
    }
  }
#undef WTRY
  *out = w;
  return ISKRA_OK;
fail:
  ws_release(w);
  delete w;
  return rc;
}

int iskra_ws_destroy(iskra_ws *w) {
  if (!w) return ISKRA_OK;
  if (w->ctx) (void)hipSetDevice(w->ctx->device);
  if (w->ctx) (void)hipStreamSynchronize(w->ctx->stream);
  ws_release(w);
  delete w;
  return ISKRA_OK;
}

int iskra_ws_param_count(iskra_ws *w, int64_t *n) {
  WRC(ws_check(w));
  if (!n) return ISKRA_E_ARG;
  *n = w->nparam;
  return ISKRA_OK;
}

int iskra_ws_set_params(iskra_ws *w, const float *p) {
  WRC(ws_check(w));
  if (!p) return ISKRA_E_ARG;
  ISK_HIP(hipMemcpyAsync(w->params, p, w->nparam * 4, hipMemcpyHostToDevice, w->ctx->stream));
  WRC(ws_zero(w, w->m, w->nparam));
  WRC(ws_zero(w, w->v, w->nparam));
  w->adam_t = 0;
  WRC(ws_prep_weights(w));
  ISK_HIP(hipStreamSynchronize(w->ctx->stream));
  return ISKRA_OK;
}

int iskra_ws_get_params(iskra_ws *w, float *p) {
  WRC(ws_check(w));
  if (!p) return ISKRA_E_ARG;
  ISK_HIP(hipMemcpyAsync(p, w->params, w->nparam * 4, hipMemcpyDeviceToHost, w->ctx->stream));
  ISK_HIP(hipStreamSynchronize(w->ctx->stream));
  return ISKRA_OK;
}

int iskra_ws_set_gain(iskra_ws *w, double gain) {
  WRC(ws_check(w));
  if (!(gain > 0.0)) {
    isk_set_error("ws_set_gain: gain must be positive");
    return ISKRA_E_ARG;
  }
  w->gain = gain;
  return ISKRA_OK;
}

int iskra_ws_set_objective(iskra_ws *w, int objective, double omega, int sweeps) {
  WRC(ws_check(w));
  if ((objective != ISKRA_WS_PHI_MSE && objective != ISKRA_WS_ENERGY &&
       objective != ISKRA_WS_ENERGY_NET) ||
      !(omega >= 0.0) ||
      omega >= 2.0 || sweeps < 0 || sweeps > WS_MAX_SWEEPS) {
    isk_set_error("ws_set_objective: objective ISKRA_WS_PHI_MSE, _ENERGY or _ENERGY_NET, omega in "
                  "[0, 2), sweeps in [0, %d]", WS_MAX_SWEEPS);
    return ISKRA_E_ARG;
  }
  w->obj = objective;
  w->om4 = 0.25 * omega;
  w->sweeps = sweeps;
  return ISKRA_OK;
}

static int ws_host_input(iskra_ws *w, int batch, const float *in) {
  const long long npix = ws_npix(w, 0, batch);
  ISK_HIP(hipMemcpyAsync(w->xin, in, npix * 4, hipMemcpyHostToDevice, w->ctx->stream));
  hipLaunchKernelGGL(k_ws_input, dim3(ws_grid(npix)), dim3(EW_BLOCK), 0, w->ctx->stream,
                     (const float *)w->xin, w->A0.a, npix);
  ISK_CHECK_LAUNCH();
  return ISKRA_OK;
}

int iskra_ws_forward(iskra_ws *w, int batch, const float *in, float *out) {
  WRC(ws_check(w));
  if (batch < 1 || batch > w->maxb || !in || !out) return ISKRA_E_ARG;
  WRC(ws_host_input(w, batch, in));
  WRC(ws_forward(w, batch));
  ISK_HIP(hipMemcpyAsync(out, w->out, ws_npix(w, 0, batch) * 4, hipMemcpyDeviceToHost,
                         w->ctx->stream));
  ISK_HIP(hipStreamSynchronize(w->ctx->stream));
  return ISKRA_OK;
}

int iskra_ws_backward(iskra_ws *w, int batch, const float *in, const float *target,
                      float *grads, double *loss) {
  WRC(ws_check(w));
  if (batch < 1 || batch > w->maxb || !in || !target) return ISKRA_E_ARG;
  WRC(ws_host_input(w, batch, in));
  ISK_HIP(hipMemcpyAsync(w->tgt, target, ws_npix(w, 0, batch) * 4, hipMemcpyHostToDevice,
                         w->ctx->stream));
  WRC(ws_pre(w, batch));
  WRC(ws_forward(w, batch));
  WRC(ws_loss(w, batch, true, loss));
  WRC(ws_backward(w, batch));
  if (grads)
    ISK_HIP(hipMemcpyAsync(grads, w->grads, w->nparam * 4, hipMemcpyDeviceToHost,
                           w->ctx->stream));
  ISK_HIP(hipStreamSynchronize(w->ctx->stream));
  return ISKRA_OK;
}

int iskra_ws_dataset_reserve(iskra_ws *w, int64_t n) {
  WRC(ws_check(w));
  if (n < 1) return ISKRA_E_ARG;
  ISK_HIP(hipStreamSynchronize(w->ctx->stream));
  isk_free(w->ds_in);
  isk_free(w->ds_out);
  w->ds_in = nullptr;
  w->ds_out = nullptr;
  w->ds_cap = w->ds_n = 0;
  const long long S2 = (long long)w->S * w->S;
  WRC(isk_alloc((void **)&w->ds_in, (size_t)n * S2 * 2));
  WRC(isk_alloc((void **)&w->ds_out, (size_t)n * S2 * 4));
  w->ds_cap = n;
  return ISKRA_OK;
}

int iskra_ws_capture(iskra_ws *w, int64_t *index) {
  WRC(ws_check(w));
  WRC(ws_check_grid(w));
  if (w->ds_n >= w->ds_cap) {
    isk_set_error("ws_capture: dataset full (%lld samples)", (long long)w->ds_cap);
    return ISKRA_E_CAPACITY;
  }
  double fscale;
  WRC(ws_rhs_scale(w, &fscale));
  const long long S2 = (long long)w->S * w->S;
  iskra_ctx *c = w->ctx;
  hipLaunchKernelGGL(k_ws_rhs_store, dim3(ws_grid(S2)), dim3(EW_BLOCK), 0, c->stream,
                     (const double *)c->rho, (const double *)c->phi, c->g.nx, fscale,
                     (const double *)w->scal, w->ds_in + w->ds_n * S2, (uint16_t *)nullptr,
                     w->ds_out + w->ds_n * S2);
  ISK_CHECK_LAUNCH();
  if (index) *index = w->ds_n;
  ++w->ds_n;
  return ISKRA_OK;
}

int iskra_ws_dataset_size(iskra_ws *w, int64_t *n) {
  WRC(ws_check(w));
  if (!n) return ISKRA_E_ARG;
  *n = w->ds_n;
  return ISKRA_OK;
}

/* RMS of the stored (phi / s) targets of samples [first, first+count) */
int iskra_ws_dataset_rms(iskra_ws *w, int64_t first, int64_t count, double *rms) {
  WRC(ws_check(w));
  if (first < 0 || count < 1 || first + count > w->ds_n || !rms) return ISKRA_E_ARG;
  const long long S2 = (long long)w->S * w->S;
  const long long n = count * S2;
  const unsigned nb = std::min(ws_grid(n), (unsigned)w->dpart_n);
  hipLaunchKernelGGL(k_ws_sumsq, dim3(nb), dim3(EW_BLOCK), 0, w->ctx->stream,
                     (const float *)(w->ds_out + first * S2), n, w->dpart);
  ISK_CHECK_LAUNCH();
  std::vector<double> h(nb);
  ISK_HIP(hipMemcpyAsync(h.data(), w->dpart, nb * 8, hipMemcpyDeviceToHost, w->ctx->stream));
  ISK_HIP(hipStreamSynchronize(w->ctx->stream));
  double s = 0.0;
  for (unsigned i = 0; i < nb; ++i) s += h[i];
  *rms = sqrt(s / (double)(count * (long long)(w->S - 1) * (w->S - 1)));
  return ISKRA_OK;
}

/* energy objective with k > 0 sweeps: the batch pair (x in A0, t in tgt)
 * becomes (P x, P t), P = (I - om4 A)^k, the residual and correction the
 * network meets after k pre-sweeps from zero (A^-1 commutes with P) */
static int ws_pre(iskra_ws *w, int b) {
  if (w->obj == ISKRA_WS_PHI_MSE || w->sweeps == 0) return ISKRA_OK;
  hipStream_t st = w->ctx->stream;
  const long long npix = ws_npix(w, 0, b);
  const unsigned g = ws_grid(npix);
  float *buf[2] = {w->eg_e, w->eg_ae};
  const float om4 = (float)w->om4;
  for (int pass = 0; pass < 2; ++pass) {
    const float *src = w->tgt;
    if (pass == 0) {
      hipLaunchKernelGGL(k_ws_a0_get, dim3(g), dim3(EW_BLOCK), 0, st, (const uint16_t *)w->A0.a,
                         w->eg_ae, npix);
      ISK_CHECK_LAUNCH();
      src = w->eg_ae;
    }
    int cur = 0;
    for (int k = 0; k < w->sweeps; ++k) {
      hipLaunchKernelGGL(k_ws_eg_poly, dim3(g), dim3(EW_BLOCK), 0, st, src, buf[cur], w->S, npix,
                         om4);
      ISK_CHECK_LAUNCH();
      src = buf[cur];
      cur ^= 1;
    }
    if (pass == 0) {
      hipLaunchKernelGGL(k_ws_input, dim3(g), dim3(EW_BLOCK), 0, st, src, w->A0.a, npix);
      ISK_CHECK_LAUNCH();
    } else {
      ISK_HIP(hipMemcpyAsync(w->tgt, src, npix * 4, hipMemcpyDeviceToDevice, st));
    }
  }
  return ISKRA_OK;
}

static int ws_load_batch(iskra_ws *w, const long long *idx, int b) {
  ISK_HIP(hipMemcpyAsync(w->idx, idx, (size_t)b * 8, hipMemcpyHostToDevice, w->ctx->stream));
  const long long S2 = (long long)w->S * w->S;
  hipLaunchKernelGGL(k_ws_batch, dim3(ws_grid(b * S2)), dim3(EW_BLOCK), 0, w->ctx->stream,
                     (const uint16_t *)w->ds_in, (const float *)w->ds_out,
                     (const long long *)w->idx, w->A0.a, w->tgt, (int)S2, b,
                     (float)(1.0 / w->gain));
  ISK_CHECK_LAUNCH();
  WRC(ws_pre(w, b));
  /* the index upload reads host memory of this call: finish before returning */
  ISK_HIP(hipStreamSynchronize(w->ctx->stream));
  return ISKRA_OK;
}

/* steps of Adam on random batches of samples [first, first+count), drawn with
 * Philox4x32-10 keyed by seed; losses[s] (nullable) = batch loss of step s */
int iskra_ws_train(iskra_ws *w, int64_t first, int64_t count, int batch, int steps, float lr,
                   uint64_t seed, double *losses) {
  WRC(ws_check(w));
  if (first < 0 || count < 1 || first + count > w->ds_n || batch < 1 || batch > w->maxb ||
      steps < 0 || !(lr > 0.f)) {
    isk_set_error("ws_train: bad range, batch or learning rate");
    return ISKRA_E_ARG;
  }
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  std::vector<long long> idx(batch);
  for (int s = 0; s < steps; ++s) {
    for (int k = 0; k < batch; ++k) {
      IskU4 r = isk_philox((IskU4){(uint32_t)(w->adam_t + s), (uint32_t)k, 0x57534200u, 0u},
                           (uint32_t)seed, (uint32_t)(seed >> 32));
      idx[k] = first + (long long)(((uint64_t)r.x << 32 | r.y) % (uint64_t)count);
    }
    WRC(ws_load_batch(w, idx.data(), batch));
    WRC(ws_forward(w, batch));
    double loss = 0.0;
    WRC(ws_loss(w, batch, true, &loss));
    if (losses) losses[s] = loss;
    WRC(ws_backward(w, batch));
    const long long t = ++w->adam_t;
    const float c1 = (float)(1.0 / (1.0 - pow(b1, (double)t)));
    const float c2 = (float)(1.0 / (1.0 - pow(b2, (double)t)));
    hipLaunchKernelGGL(k_ws_adam, dim3(ws_grid(w->nparam)), dim3(EW_BLOCK), 0, w->ctx->stream,
                       w->params, (const float *)w->grads, w->m, w->v, w->nparam, lr, b1, b2, eps,
                       c1, c2);
    ISK_CHECK_LAUNCH();
    WRC(ws_prep_weights(w));
  }
  ISK_HIP(hipStreamSynchronize(w->ctx->stream));
  return ISKRA_OK;
}

/* mean loss over samples [first, first+count), batches of the maximum size */
int iskra_ws_eval(iskra_ws *w, int64_t first, int64_t count, double *mse) {
  WRC(ws_check(w));
  if (first < 0 || count < 1 || first + count > w->ds_n || !mse) return ISKRA_E_ARG;
  double tot = 0.0;
  std::vector<long long> idx(w->maxb);
  for (long long s = 0; s < count; s += w->maxb) {
    const int b = (int)std::min((long long)w->maxb, count - s);
    for (int k = 0; k < b; ++k) idx[k] = first + s + k;
    WRC(ws_load_batch(w, idx.data(), b));
    WRC(ws_forward(w, b));
    double l = 0.0;
    WRC(ws_loss(w, b, false, &l));
    tot += l * b;
  }
  *mse = tot / (double)count;
  return ISKRA_OK;
}

/* z = the two-level cycle applied to the residual r: `sweeps` damped Jacobi
 * sweeps from z = 0, the network's estimate of A^-1 of what they leave
 * (r - A z, scaled by its RMS and the gain) added to z, then `sweeps`
 * sweeps again.  om4 = 0 or sweeps = 0: the network alone. */
static int ls_precond(iskra_ws *w, int nx, double om4, int sweeps, const double **z) {
  iskra_ctx *c = w->ctx;
  hipStream_t st = c->stream;
  const long long n = (long long)(nx - 2) * (nx - 2);
  const unsigned g = (unsigned)std::min((n + LS_BLOCK - 1) / LS_BLOCK, 1024ll);
  const int k = om4 > 0.0 ? sweeps : 0;
  /* pre-sweeps: ping-pong between lpre[0] and lpre[1] */
  const double *zpre = nullptr;
  for (int i = 0; i < k; ++i) {
    hipLaunchKernelGGL(k_ls_smooth, dim3(g), dim3(LS_BLOCK), 0, st, (const double *)w->lr, zpre,
                       w->lpre[i & 1], nx, om4);
    ISK_CHECK_LAUNCH();
    zpre = w->lpre[i & 1];
  }
  const double *rin = w->lr;
  if (zpre) {
    hipLaunchKernelGGL(k_ls_resid, dim3(g), dim3(LS_BLOCK), 0, st, (const double *)w->lr, zpre,
                       w->lrp, nx);
    ISK_CHECK_LAUNCH();
    rin = w->lrp;
  }
  const long long cnt = (long long)(w->S - 1) * (w->S - 1);
  const unsigned nb = std::min(ws_grid(cnt), (unsigned)w->dpart_n);
  hipLaunchKernelGGL(k_ws_rhs_sumsq, dim3(nb), dim3(EW_BLOCK), 0, st, rin, nx, 1.0, w->dpart);
  ISK_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_ws_rhs_scale, dim3(1), dim3(EW_BLOCK), 0, st, (const double *)w->dpart, (int)nb,
                     cnt, w->scal);
  ISK_CHECK_LAUNCH();
  const long long S2 = (long long)w->S * w->S;
  hipLaunchKernelGGL(k_ws_rhs_store, dim3(ws_grid(S2)), dim3(EW_BLOCK), 0, st, rin,
                     (const double *)nullptr, nx, 1.0, (const double *)w->scal,
                     (uint16_t *)nullptr, w->A0.a, (float *)nullptr);
  ISK_CHECK_LAUNCH();
  WRC(ws_forward(w, 1));
  hipLaunchKernelGGL(k_ws_phi_out, dim3(ws_grid(S2)), dim3(EW_BLOCK), 0, st,
                     (const float *)w->out, w->lz, nx, w->gain, (const double *)w->scal);
  ISK_CHECK_LAUNCH();
  if (zpre) {
    hipLaunchKernelGGL(k_ls_add, dim3(g), dim3(LS_BLOCK), 0, st, w->lz, zpre, nx);
    ISK_CHECK_LAUNCH();
  }
  *z = w->lz;
  double *buf[2] = {w->lzs, w->lz};
  for (int i = 0; i < k; ++i) {
    hipLaunchKernelGGL(k_ls_smooth, dim3(g), dim3(LS_BLOCK), 0, st, (const double *)w->lr, *z,
                       buf[i & 1], nx, om4);
    ISK_CHECK_LAUNCH();
    *z = buf[i & 1];
  }
  return ISKRA_OK;
}

/* the context must hold the network's system: a configured solver with
 * phi = 0 on the four walls and no other Dirichlet node or periodic link;
 * allocates the iteration vectors on first use */
static int ls_setup(iskra_ws *w, const char *who, double omega, int sweeps) {
  WRC(ws_check_grid(w));
  iskra_ctx *c = w->ctx;
  IskPoisson &p = c->ps;
  if (p.kind < 0) {
    isk_set_error("%s: no Poisson solver configured", who);
    return ISKRA_E_STATE;
  }
  if (!(omega >= 0.0) || omega >= 2.0 || sweeps < 0 || sweeps > WS_MAX_SWEEPS) {
    isk_set_error("%s: the Jacobi weight must lie in [0, 2), sweeps in [0, %d]", who,
                  WS_MAX_SWEEPS);
    return ISKRA_E_ARG;
  }
  const int nx = c->g.nx;
  bool ok = !p.per_x && !p.per_y && (int)p.tags.size() == nx * nx;
  for (int j = 0; ok && j < nx; ++j)
    for (int i = 0; i < nx; ++i) {
      const int t = p.tags[(size_t)j * nx + i];
      const bool wall = i == 0 || j == 0 || i == nx - 1 || j == nx - 1;
      if (wall ? (t == 0 || p.tag_val[t] != 0.0) : t != 0) {
        ok = false;
        break;
      }
    }
  if (!ok) {
    isk_set_error("%s: the network solves phi = 0 on all four walls with no other Dirichlet "
                  "node or periodic link", who);
    return ISKRA_E_UNSUPPORTED;
  }
  const long long nn = (long long)nx * nx;
  if (!w->lx) {
    double **bufs[] = {&w->lx,    &w->lr,    &w->lz,      &w->lzs,     &w->lp[0], &w->lp[1],
                       &w->lq[0], &w->lq[1], &w->lpre[0], &w->lpre[1], &w->lrp};
    for (double **b : bufs) {
      WRC(isk_alloc((void **)b, nn * 8));
      ISK_HIP(hipMemsetAsync(*b, 0, nn * 8, c->stream));
    }
    WRC(isk_alloc((void **)&w->lpart, 4096 * 8));
    WRC(isk_alloc((void **)&w->ls, LS_NS * 8));
    ISK_HIP(hipMemsetAsync(w->ls, 0, LS_NS * 8, c->stream));
  }
  return ISKRA_OK;
}

/* training pair (r, phi - x) of the current iterate appended to the dataset */
static int ls_capture_pair(iskra_ws *w) {
  iskra_ctx *c = w->ctx;
  const int nx = c->g.nx;
  const long long cnt = (long long)(w->S - 1) * (w->S - 1);
  const unsigned nb = std::min(ws_grid(cnt), (unsigned)w->dpart_n);
  hipLaunchKernelGGL(k_ws_rhs_sumsq, dim3(nb), dim3(EW_BLOCK), 0, c->stream,
                     (const double *)w->lr, nx, 1.0, w->dpart);
  ISK_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_ws_rhs_scale, dim3(1), dim3(EW_BLOCK), 0, c->stream, (const double *)w->dpart,
                     (int)nb, cnt, w->scal);
  ISK_CHECK_LAUNCH();
  const long long S2 = (long long)w->S * w->S;
  hipLaunchKernelGGL(k_ws_pair_store, dim3(ws_grid(S2)), dim3(EW_BLOCK), 0, c->stream,
                     (const double *)w->lr, (const double *)c->phi, (const double *)w->lx, nx,
                     (const double *)w->scal, w->ds_in + w->ds_n * S2, w->ds_out + w->ds_n * S2);
  ISK_CHECK_LAUNCH();
  ++w->ds_n;
  return ISKRA_OK;
}

/* Flexible CG from phi (warm) or 0 on f = dx^2 rho / eps0 until the relative
 * residual meets rtol or maxit iterations; with caps (ascending), the
 * iteration also stops after the last cap and the pair of every capped
 * iterate is captured.  Iterate in lx, residual in lr. */
static int ls_iterate(iskra_ws *w, int warm, double om4, int sweeps, int maxit, double rtol,
                      const int *caps, int ncaps, int *it_out, double *rr_out) {
  iskra_ctx *c = w->ctx;
  const int nx = c->g.nx;
  const long long n = (long long)(nx - 2) * (nx - 2);
  const int nb = (int)std::min((n + LS_BLOCK - 1) / LS_BLOCK, 1024ll);
  const double fscale = c->g.dx * c->g.dx / c->ps.eps0;
  hipLaunchKernelGGL(k_ls_init, dim3(nb), dim3(LS_BLOCK), 0, c->stream, (const double *)c->rho,
                     warm == ISKRA_WARM_PREV ? (const double *)c->phi : (const double *)nullptr,
                     nx, fscale, w->lx, w->lr, w->lpart);
  ISK_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_ls_reduce, dim3(1), dim3(EW_BLOCK), 0, c->stream, (const double *)w->lpart,
                     nb, 2, w->ls + LS_FF);
  ISK_CHECK_LAUNCH();
  double s[2];
  ISK_HIP(hipMemcpyAsync(s, w->ls, 16, hipMemcpyDeviceToHost, c->stream));
  ISK_HIP(hipStreamSynchronize(c->stream));
  double rr = s[0] > 0.0 ? sqrt(s[1] / s[0]) : 0.0;
  int it = 0, cur = 0, next_cap = 0;
  if (caps) maxit = std::min(maxit, caps[ncaps - 1]);
  while (rr > rtol && it < maxit) {
    ++it;
    const double *z = nullptr;
    WRC(ls_precond(w, nx, om4, sweeps, &z));
    const int prev = cur ^ 1;
    if (it > 1) {
      hipLaunchKernelGGL(k_ls_dot, dim3(nb), dim3(LS_BLOCK), 0, c->stream, z,
                         (const double *)w->lq[prev], nx, w->lpart);
      ISK_CHECK_LAUNCH();
      hipLaunchKernelGGL(k_ls_reduce, dim3(1), dim3(EW_BLOCK), 0, c->stream,
                         (const double *)w->lpart, nb, 1, w->ls + LS_ZQ);
      ISK_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_ls_dir, dim3(nb), dim3(LS_BLOCK), 0, c->stream, z,
                       (const double *)w->lp[prev], w->lp[cur], nx, (const double *)w->ls,
                       it == 1 ? 1 : 0);
    ISK_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_ls_Ap, dim3(nb), dim3(LS_BLOCK), 0, c->stream,
                       (const double *)w->lp[cur], (const double *)w->lr, w->lq[cur], nx,
                       w->lpart);
    ISK_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_ls_reduce, dim3(1), dim3(EW_BLOCK), 0, c->stream, (const double *)w->lpart,
                       nb, 2, w->ls + LS_PQ);
    ISK_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_ls_xr, dim3(nb), dim3(LS_BLOCK), 0, c->stream, w->lx, w->lr,
                       (const double *)w->lp[cur], (const double *)w->lq[cur], nx,
                       (const double *)w->ls, w->lpart);
    ISK_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_ls_reduce, dim3(1), dim3(EW_BLOCK), 0, c->stream, (const double *)w->lpart,
                       nb, 1, w->ls + LS_RR);
    ISK_CHECK_LAUNCH();
    if (caps && next_cap < ncaps && it == caps[next_cap]) {
      WRC(ls_capture_pair(w));
      ++next_cap;
    }
    ISK_HIP(hipMemcpyAsync(s, w->ls, 16, hipMemcpyDeviceToHost, c->stream));
    ISK_HIP(hipStreamSynchronize(c->stream));
    rr = s[0] > 0.0 ? sqrt(s[1] / s[0]) : 0.0;
    cur ^= 1;
  }
  *it_out = it;
  *rr_out = rr;
  return ISKRA_OK;
}

int iskra_ws_solve(iskra_ws *w, int warm, double omega, int sweeps, int max_iter, int *iters,
                   double *relres) {
  WRC(ws_check(w));
  if (warm != ISKRA_WARM_ZERO && warm != ISKRA_WARM_PREV) return ISKRA_E_ARG;
  WRC(ls_setup(w, "ws_solve", omega, sweeps));
  iskra_ctx *c = w->ctx;
  IskPoisson &p = c->ps;
  const int nx = c->g.nx;
  hipEvent_t ev = nullptr;
  WRC(isk_prof_begin(c, &ev));
  int it = 0;
  double rr = 0.0;
  WRC(ls_iterate(w, warm, 0.25 * omega, sweeps, max_iter > 0 ? max_iter : p.max_iter, p.rtol,
                 nullptr, 0, &it, &rr));
  const long long n = (long long)(nx - 2) * (nx - 2);
  const int nb = (int)std::min((n + LS_BLOCK - 1) / LS_BLOCK, 1024ll);
  hipLaunchKernelGGL(k_ls_out, dim3(nb), dim3(LS_BLOCK), 0, c->stream, (const double *)w->lx,
                     c->phi, nx);
  ISK_CHECK_LAUNCH();
  WRC(isk_prof_end(c, ev, ISKRA_K_POISSON, 0));
  ISK_HIP(hipStreamSynchronize(c->stream));
  p.last_iters = it;
  p.last_relres = rr;
  c->stats.poisson_solves++;
  c->stats.poisson_iters += it;
  c->stats.last_iters = it;
  c->stats.last_relres = rr;
  if (iters) *iters = it;
  if (relres) *relres = rr;
  if (rr > p.rtol) {
    isk_set_error("ws_solve: relative residual %.3e above %.3e after %d iterations", rr, p.rtol,
                  it);
    return ISKRA_E_CONVERGE;
  }
  return ISKRA_OK;
}

int iskra_ws_capture_iterates(iskra_ws *w, double omega, int sweeps, const int *iters,
                              int n_iters, int64_t *captured) {
  WRC(ws_check(w));
  if (!iters || n_iters < 1 || iters[0] < 1) {
    isk_set_error("ws_capture_iterates: n_iters >= 1 iteration counts, the first >= 1");
    return ISKRA_E_ARG;
  }
  for (int k = 1; k < n_iters; ++k)
    if (iters[k] <= iters[k - 1]) {
      isk_set_error("ws_capture_iterates: iteration counts must increase");
      return ISKRA_E_ARG;
    }
  if (w->ds_n + n_iters > w->ds_cap) {
    isk_set_error("ws_capture_iterates: dataset full (%lld of %lld samples)", (long long)w->ds_n,
                  (long long)w->ds_cap);
    return ISKRA_E_CAPACITY;
  }
  WRC(ls_setup(w, "ws_capture_iterates", omega, sweeps));
  const long long n0 = w->ds_n;
  int it = 0;
  double rr = 0.0;
  WRC(ls_iterate(w, ISKRA_WARM_ZERO, 0.25 * omega, sweeps, iters[n_iters - 1], 0.0, iters,
                 n_iters, &it, &rr));
  ISK_HIP(hipStreamSynchronize(w->ctx->stream));
  if (captured) *captured = w->ds_n - n0;
  return ISKRA_OK;
}

int iskra_ws_dataset_get(iskra_ws *w, int64_t index, float *in, float *out) {
  WRC(ws_check(w));
  if (index < 0 || index >= w->ds_n || !in || !out) {
    isk_set_error("ws_dataset_get: index outside the dataset or null buffer");
    return ISKRA_E_ARG;
  }
  const long long S2 = (long long)w->S * w->S;
  std::vector<uint16_t> h(S2);
  ISK_HIP(hipMemcpyAsync(h.data(), w->ds_in + index * S2, S2 * 2, hipMemcpyDeviceToHost,
                         w->ctx->stream));
  ISK_HIP(hipMemcpyAsync(out, w->ds_out + index * S2, S2 * 4, hipMemcpyDeviceToHost,
                         w->ctx->stream));
  ISK_HIP(hipStreamSynchronize(w->ctx->stream));
  for (long long i = 0; i < S2; ++i) in[i] = ws_bf2f(h[i]);
  return ISKRA_OK;
}

/* phi at the context's interior nodes <- network prediction from its rho */
int iskra_ws_predict(iskra_ws *w) {
  WRC(ws_check(w));
  WRC(ws_check_grid(w));
  iskra_ctx *c = w->ctx;
  hipEvent_t ev = nullptr;
  WRC(isk_prof_begin(c, &ev));
  double fscale;
  WRC(ws_rhs_scale(w, &fscale));
  const long long S2 = (long long)w->S * w->S;
  hipLaunchKernelGGL(k_ws_rhs_store, dim3(ws_grid(S2)), dim3(EW_BLOCK), 0, c->stream,
                     (const double *)c->rho, (const double *)nullptr, c->g.nx, fscale,
                     (const double *)w->scal, (uint16_t *)nullptr, w->A0.a, (float *)nullptr);
  ISK_CHECK_LAUNCH();
  WRC(ws_forward(w, 1));
  hipLaunchKernelGGL(k_ws_phi_out, dim3(ws_grid(S2)), dim3(EW_BLOCK), 0, c->stream,
                     (const float *)w->out, c->phi, c->g.nx, w->gain, (const double *)w->scal);
  ISK_CHECK_LAUNCH();
  WRC(isk_prof_end(c, ev, ISKRA_K_WARMSTART, 0));
  return ISKRA_OK;
}

} /* extern "C" */
