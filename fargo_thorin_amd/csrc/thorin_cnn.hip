// thorin_cnn.hip -- learned warm start of the energy equation's implicit
// solve (EnergySources.c:200-335 solves A.T = RHS by red-black SOR from the
// current temperature).  No counterpart in the reference: a small
// convolutional network predicts the correction T* - T0 of the converged
// solve from the assembled five-point system and T0, so that SOR (or the
// FP32-refinement solve) starts closer to the solution; the solver still
// iterates to the same residual target and rejects a guess that raises the
// residual.
//
// Network: a 4-level U-Net with 32 channels (levels at 1, 1/2, 1/4, 1/8 of
// the NRAD x NSEC grid; average pooling down, nearest upsampling plus skip
// addition up), nine 3x3 convolutions:
//   a1 = relu(conv(x, 16->32))   a2 = relu(conv(a1))          level 0
//   a3 = relu(conv(pool(a2)))                                  level 1
//   a4 = relu(conv(pool(a3)))                                  level 2
//   a5 = relu(conv(pool(a4)))                                  level 3
//   a6 = relu(conv(up(a5) + a4))                               level 2
//   a7 = relu(conv(up(a6) + a3))                               level 1
//   a8 = relu(conv(up(a7) + a2))                               level 0
//   y  = conv(a8, 32->16), channel 0                           level 0
// Radial padding is zero, azimuthal padding periodic (the polar grid).
//
// Coarse network (thorin_cnn_set_coarsen, f = 2 or 4): the features are
// formed on the NRAD x NSEC grid and averaged over f x f cells, the U-Net
// runs on the NRAD/f x NSEC/f grid, and its output is interpolated back
// bilinearly (cell centres; clamped across the rings, periodic along the
// sectors).  The residual loss is taken on the full grid and its gradient
// reaches the network through the transpose of that interpolation.  The
// smooth part of the correction is what the FP32 sweeps remove slowest, and
// the network then costs about 1/f^2 as much.
//
// Arithmetic: activations and weights in BF16 (NHWC, 16 or 32 channels per
// pixel), convolutions as implicit GEMMs on v_mfma_f32_16x16x16_bf16 with
// FP32 accumulation, FP32 master weights and Adam.  Every activation (and
// activation gradient) is rounded to BF16 where it is stored; the CPU
// restatement in oracle/cnn_oracle.py rounds at the same points.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "thorin_internal.h"

#define HCK(call)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess) {                                                              \
      char m_[256];                                                                      \
      snprintf(m_, sizeof m_, "HIP error %s at %s:%d", hipGetErrorString(e_), __FILE__, \
               __LINE__);                                                                \
      thorin_set_error(m_);                                                              \
      return -1;                                                                         \
    }                                                                                    \
  } while (0)

typedef unsigned short bf16_t;
typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

static const int NL = 9;            // convolutions
static const int WID = 32;          // channels
static const int CIN0 = 16;         // input channels (6 used)
static const int COUTL = 16;        // output channels of the last convolution (1 used)
static const int NFEAT = 6;

__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float((unsigned)h << 16); }
__device__ __forceinline__ bf16_t f2bf(float f) {   // round to nearest even
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}
static float h_bf2f(bf16_t h) {
  unsigned u = (unsigned)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}
static bf16_t h_f2bf(float f) {
  unsigned u;
  memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}

// ---------------------------------------------------------------------------
// 3x3 convolution as an implicit GEMM, computed transposed so that a lane
// ends with four consecutive channels of one pixel: D^T[co][px] =
// sum_k W^T[co][k] X^T[k][px] with K running over the 9 taps x CIN channels,
// 16 at a time.  MFMA operand layouts (16x16x16): lane l holds A[l%16][4(l/16)+i],
// B[4(l/16)+i][l%16] and D[4(l/16)+i][l%16], i = 0..3; so the A fragment is
// W[t][cb*16+4(l/16)+i][nb*16+l%16] (packed by k_pack as wpk[((t*CB+cb)*NB+nb)*64+l])
// and the B fragment is 4 consecutive input channels of pixel l%16 (one
// 8-byte load), D gives output channels nb*16+4(l/16)+i of pixel l%16 (one
// 8-byte BF16 or 16-byte FP32 store).  A wave computes MT tiles of 16
// consecutive pixels of a row and reuses each weight fragment for all of them.
// mode: 1 relu; 2 FP32 output; 4 zero the output where mask <= 0 (the relu
// of the layer whose input gradient this is).
template <int CIN, int COUT, int MT>
__global__ void __launch_bounds__(256) k_conv(const bf16_t *__restrict__ x, const bf16_t *__restrict__ wpk,
                                              const float *__restrict__ bias, void *__restrict__ out, int H, int W,
                                              int mode, const bf16_t *__restrict__ mask) {
  constexpr int NB = COUT / 16, CB = CIN / 16;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int tilesW = (W + 16 * MT - 1) / (16 * MT);
  const long tile = (long)blockIdx.x * 4 + wv;
  if (tile >= (long)H * tilesW) return;
  const int h = (int)(tile / tilesW), w0 = (int)(tile % tilesW) * 16 * MT;
  const int m = lane & 15, kq = lane >> 4;
  f4 acc[MT][NB];
#pragma unroll
  for (int u = 0; u < MT; u++)
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc[u][nb] = (f4){0.f, 0.f, 0.f, 0.f};
  const s4 *__restrict__ wp = (const s4 *)wpk;
#pragma unroll
  for (int t = 0; t < 9; t++) {
    const int hh = h + t / 3 - 1;
    const bool rok = hh >= 0 && hh < H;
#pragma unroll
    for (int cb = 0; cb < CB; cb++) {
      s4 wf[NB];
#pragma unroll
      for (int nb = 0; nb < NB; nb++) wf[nb] = wp[((t * CB + cb) * NB + nb) * 64 + lane];
#pragma unroll
      for (int u = 0; u < MT; u++) {
        const int px = w0 + 16 * u + m;
        int ww = px + (t % 3) - 1;
        ww = ww < 0 ? ww + W : (ww >= W ? ww - W : ww);
        s4 xf = (s4){0, 0, 0, 0};
        if (rok && px < W) xf = *(const s4 *)(x + ((size_t)hh * W + ww) * CIN + cb * 16 + kq * 4);
#pragma unroll
        for (int nb = 0; nb < NB; nb++) acc[u][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wf[nb], xf, acc[u][nb], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int u = 0; u < MT; u++) {
    const int px = w0 + 16 * u + m;
    if (px >= W) continue;
#pragma unroll
    for (int nb = 0; nb < NB; nb++) {
      const int co = nb * 16 + 4 * kq;
      const size_t o = ((size_t)h * W + px) * COUT + co;
      float v[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        v[i] = acc[u][nb][i] + (bias ? bias[co + i] : 0.f);
        if (mode & 1) v[i] = v[i] > 0.f ? v[i] : 0.f;
      }
      if (mode & 4) {
        const s4 mk = *(const s4 *)(mask + o);
#pragma unroll
        for (int i = 0; i < 4; i++)
          if (!(bf2f(mk[i]) > 0.f)) v[i] = 0.f;
      }
      if (mode & 2) {
        *(f4 *)((float *)out + o) = (f4){v[0], v[1], v[2], v[3]};
      } else {
        s4 r;
#pragma unroll
        for (int i = 0; i < 4; i++) r[i] = (short)f2bf(v[i]);
        *(s4 *)((bf16_t *)out + o) = r;
      }
    }
  }
}

// The same convolution with the input staged in LDS, four output rows per
// workgroup (one per wave): rows h0-1 .. h0+4 of a CV_SEG4-pixel segment plus
// one halo pixel on each side (periodic) are loaded once (8-byte vectors, all
// in flight before their LDS stores), so an input row is read 1.5 times, and
// every tap reads them from LDS.  The pixel stride is padded (CIN + 4 BF16) so
// that the 16 pixels of a fragment hit distinct banks.
#define CV_SEG4 128
template <int CIN, int COUT>
__global__ void __launch_bounds__(256) k_conv_lds4(const bf16_t *__restrict__ x, const bf16_t *__restrict__ wpk,
                                                   const float *__restrict__ bias, void *__restrict__ out, int H,
                                                   int W, int mode, const bf16_t *__restrict__ mask) {
  constexpr int NB = COUT / 16, CB = CIN / 16, PS = CIN + 4, MT = CV_SEG4 / 16;
  constexpr int NP = CV_SEG4 + 2, NR = 6;
  __shared__ bf16_t sx[NR * NP * PS];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int segs = (W + CV_SEG4 - 1) / CV_SEG4;
  const int h0 = (blockIdx.x / segs) * 4, w0 = (blockIdx.x % segs) * CV_SEG4;
  if (h0 >= H) return;
  {
    constexpr int NE = NR * NP * (CIN / 4), NL = (NE + 255) / 256;
    s4 v[NL];
#pragma unroll
    for (int k = 0; k < NL; k++) {
      const int e = threadIdx.x + k * 256;
      v[k] = (s4){0, 0, 0, 0};
      if (e < NE) {
        const int c4 = e % (CIN / 4), rest = e / (CIN / 4);
        const int p = rest % NP, r = rest / NP;
        const int hh = h0 + r - 1;
        int ww = w0 + p - 1;
        while (ww < 0) ww += W;
        while (ww >= W) ww -= W;
        if (hh >= 0 && hh < H) v[k] = *(const s4 *)(x + ((size_t)hh * W + ww) * CIN + c4 * 4);
      }
    }
#pragma unroll
    for (int k = 0; k < NL; k++) {
      const int e = threadIdx.x + k * 256;
      if (e < NE) {
        const int c4 = e % (CIN / 4), rest = e / (CIN / 4);
        const int p = rest % NP, r = rest / NP;
        *(s4 *)&sx[(r * NP + p) * PS + c4 * 4] = v[k];
      }
    }
  }
  __syncthreads();
  const int h = h0 + wv;
  if (h >= H) return;
  const int m = lane & 15, kq = lane >> 4;
  f4 acc[MT][NB];
#pragma unroll
  for (int u = 0; u < MT; u++)
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc[u][nb] = (f4){0.f, 0.f, 0.f, 0.f};
  const s4 *__restrict__ wp = (const s4 *)wpk;
#pragma unroll
  for (int t = 0; t < 9; t++) {
    const int r = wv + t / 3, dw = t % 3;   // staged row of the wave's tap row; sx pixel p + dw
#pragma unroll
    for (int cb = 0; cb < CB; cb++) {
      s4 wf[NB];
#pragma unroll
      for (int nb = 0; nb < NB; nb++) wf[nb] = wp[((t * CB + cb) * NB + nb) * 64 + lane];
#pragma unroll
      for (int u = 0; u < MT; u++) {
        const s4 xf = *(const s4 *)&sx[(r * NP + m + 16 * u + dw) * PS + cb * 16 + kq * 4];
#pragma unroll
        for (int nb = 0; nb < NB; nb++)
          acc[u][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wf[nb], xf, acc[u][nb], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int u = 0; u < MT; u++) {
    const int px = w0 + m + 16 * u;
    if (px >= W) continue;
#pragma unroll
    for (int nb = 0; nb < NB; nb++) {
      const int co = nb * 16 + 4 * kq;
      const size_t o = ((size_t)h * W + px) * COUT + co;
      float vv[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        vv[i] = acc[u][nb][i] + (bias ? bias[co + i] : 0.f);
        if (mode & 1) vv[i] = vv[i] > 0.f ? vv[i] : 0.f;
      }
      if (mode & 4) {
        const s4 mk = *(const s4 *)(mask + o);
#pragma unroll
        for (int i = 0; i < 4; i++)
          if (!(bf2f(mk[i]) > 0.f)) vv[i] = 0.f;
      }
      if (mode & 2) {
        *(f4 *)((float *)out + o) = (f4){vv[0], vv[1], vv[2], vv[3]};
      } else {
        s4 rr;
#pragma unroll
        for (int i = 0; i < 4; i++) rr[i] = (short)f2bf(vv[i]);
        *(s4 *)((bf16_t *)out + o) = rr;
      }
    }
  }
}

// weights W[t][ci][co] (FP32 master) -> MFMA-fragment order BF16 for the forward
// convolution, or for the input-gradient convolution (taps mirrored, ci and
// co swapped: dX[p][ci] = sum_t sum_co dZ[p - off_t][co] W[t][ci][co])
__global__ void k_pack(const float *__restrict__ w, int cin, int cout, int transposed, bf16_t *__restrict__ wpk) {
  const int CINp = transposed ? cout : cin, COUTp = transposed ? cin : cout;
  const int CB = CINp / 16, NB = COUTp / 16;
  const int n = 9 * CB * NB * 64;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const int lane = q & 63, r = q >> 6;
    const int nb = r % NB, cb = (r / NB) % CB, t = r / (NB * CB);
    for (int i = 0; i < 4; i++) {
      const int k = cb * 16 + 4 * (lane >> 4) + i, nn = nb * 16 + (lane & 15);
      float v;
      if (!transposed) v = w[((size_t)t * cin + k) * cout + nn];
      else v = w[((size_t)(8 - t) * cin + nn) * cout + k];
      wpk[(size_t)q * 4 + i] = f2bf(v);
    }
  }
}

// Weight gradient dW[t][ci][co] = sum_p x[p + off_t][ci] dz[p][co] on MFMA
// (M = ci, N = co, K = pixels).  A workgroup walks row segments of 64
// pixels, staging x (3 rows x 66 columns, periodic) and dz (64 pixels) in LDS;
// wave w owns the 9*CB*NB/4 output tiles w, w+4, ...  Partial tiles (16x16,
// row-major ci x co) go to part[block][tile]; k_wgrad_reduce sums them in
// block order.
#define WG_SEG 64
template <int CIN, int COUT>
__global__ void __launch_bounds__(256) k_wgrad(const bf16_t *__restrict__ x, const bf16_t *__restrict__ dz, int H,
                                               int W, float *__restrict__ part) {
  constexpr int CB = CIN / 16, NB = COUT / 16, NT = 9 * CB * NB, TPW = (NT + 3) / 4;
  __shared__ bf16_t sx[3][WG_SEG + 2][CIN];
  __shared__ bf16_t sdz[WG_SEG][COUT];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int segW = (W + WG_SEG - 1) / WG_SEG;
  const int nseg = H * segW;
  f4 acc[TPW];
#pragma unroll
  for (int q = 0; q < TPW; q++) acc[q] = (f4){0.f, 0.f, 0.f, 0.f};
  for (int sg = blockIdx.x; sg < nseg; sg += gridDim.x) {
    const int h = sg / segW, w0 = (sg % segW) * WG_SEG;
    __syncthreads();
    for (int e = threadIdx.x; e < 3 * (WG_SEG + 2) * CIN / 4; e += 256) {
      const int c4 = e % (CIN / 4), rest = e / (CIN / 4);
      const int col = rest % (WG_SEG + 2), r = rest / (WG_SEG + 2);
      const int hh = h + r - 1;
      const int ww = ((w0 + col - 1) % W + W) % W;    // periodic; columns past the row meet dz = 0
      s4 v = (s4){0, 0, 0, 0};
      if (hh >= 0 && hh < H) v = *(const s4 *)(x + ((size_t)hh * W + ww) * CIN + c4 * 4);
      *(s4 *)&sx[r][col][c4 * 4] = v;
    }
    for (int e = threadIdx.x; e < WG_SEG * COUT / 4; e += 256) {
      const int c4 = e % (COUT / 4), px = e / (COUT / 4);
      s4 v = (s4){0, 0, 0, 0};
      if (w0 + px < W) v = *(const s4 *)(dz + ((size_t)h * W + w0 + px) * COUT + c4 * 4);
      *(s4 *)&sdz[px][c4 * 4] = v;
    }
    __syncthreads();
    // (no early exit from the unrolled loops: acc must stay in registers)
#pragma unroll
    for (int q = 0; q < TPW; q++) {
      const int tile = wv + 4 * q;
      if (tile < NT) {
        const int nb = tile % NB, cb = (tile / NB) % CB, t = tile / (NB * CB);
        const int r = t / 3, dw = t % 3;          // sx column of pixel p for tap dw: p + dw
#pragma unroll
        for (int kc = 0; kc < WG_SEG / 16; kc++) {
          s4 a, b;
#pragma unroll
          for (int i = 0; i < 4; i++) {
            const int p = kc * 16 + 4 * (lane >> 4) + i;
            a[i] = (short)sx[r][p + dw][cb * 16 + (lane & 15)];
            b[i] = (short)sdz[p][nb * 16 + (lane & 15)];
          }
          acc[q] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[q], 0, 0, 0);
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < TPW; q++) {
    const int tile = wv + 4 * q;
    if (tile < NT) {
      float *dst = part + ((size_t)blockIdx.x * NT + tile) * 256;
#pragma unroll
      for (int i = 0; i < 4; i++) dst[(4 * (lane >> 4) + i) * 16 + (lane & 15)] = acc[q][i];
    }
  }
}

// dW[t][ci][co] = sum over blocks (in order) of the partial tiles
__global__ void k_wgrad_reduce(const float *__restrict__ part, int nblk, int cin, int cout, float *__restrict__ dw) {
  const int CB = cin / 16, NB = cout / 16, NT = 9 * CB * NB;
  const int n = NT * 256;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const int tile = e >> 8, mn = e & 255, mi = mn >> 4, ni = mn & 15;
    const int nb = tile % NB, cb = (tile / NB) % CB, t = tile / (NB * CB);
    // blocks added in order; eight loads in flight before their adds
    const float *__restrict__ src = part + (size_t)tile * 256 + mn;
    const size_t stride = (size_t)NT * 256;
    float s = 0.f;
    int b = 0;
    for (; b + 8 <= nblk; b += 8) {
      float v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = src[(size_t)(b + u) * stride];
#pragma unroll
      for (int u = 0; u < 8; u++) s += v[u];
    }
    for (; b < nblk; b++) s += src[(size_t)b * stride];
    dw[((size_t)t * cin + cb * 16 + mi) * cout + nb * 16 + ni] = s;
  }
}

// bias gradient db[co] = sum_p dz[p][co] (block partials, then in order)
__global__ void __launch_bounds__(256) k_bgrad(const bf16_t *__restrict__ dz, long npx, int C, float *__restrict__ part) {
  __shared__ float red[256];
  const int c = threadIdx.x % C, lanes = 256 / C;
  float s = 0.f;
  for (long p = (long)blockIdx.x * lanes + threadIdx.x / C; p < npx; p += (long)gridDim.x * lanes)
    s += bf2f(dz[p * C + c]);
  red[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < C) {
    float t = 0.f;
    for (int k = 0; k < lanes; k++) t += red[k * C + threadIdx.x];
    part[blockIdx.x * C + threadIdx.x] = t;
  }
}
// one workgroup of 256: thread (c, k) sums blocks k, k + 256/C, ... of channel c
__global__ void __launch_bounds__(256) k_bgrad_reduce(const float *__restrict__ part, int nblk, int C,
                                                      float *__restrict__ db) {
  __shared__ float red[256];
  const int c = threadIdx.x % C, k0 = threadIdx.x / C, lanes = 256 / C;
  float s = 0.f;
  int b = k0;
  for (; b + 7 * lanes < nblk; b += 8 * lanes) {   // eight loads in flight, added in order
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = part[(b + u * lanes) * C + c];
#pragma unroll
    for (int u = 0; u < 8; u++) s += v[u];
  }
  for (; b < nblk; b += lanes) s += part[b * C + c];
  red[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < C) {
    float t = 0.f;
    for (int k = 0; k < lanes; k++) t += red[k * C + threadIdx.x];
    db[threadIdx.x] = t;
  }
}

// ---------------------------------------------------------------------------
// level changes (32 channels, 4 at a time)
// pool: out[h][w] = 0.25 (x[2h][2w] + x[2h][2w+1] + x[2h+1][2w] + x[2h+1][2w+1])
__global__ void k_pool(const bf16_t *__restrict__ x, bf16_t *__restrict__ out, int Ho, int Wo) {
  const long n = (long)Ho * Wo * (WID / 4);
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const int c4 = e % (WID / 4);
    const long p = e / (WID / 4);
    const int h = (int)(p / Wo), w = (int)(p % Wo), Wi = 2 * Wo;
    const s4 a = *(const s4 *)(x + ((size_t)(2 * h) * Wi + 2 * w) * WID + c4 * 4);
    const s4 b = *(const s4 *)(x + ((size_t)(2 * h) * Wi + 2 * w + 1) * WID + c4 * 4);
    const s4 c = *(const s4 *)(x + ((size_t)(2 * h + 1) * Wi + 2 * w) * WID + c4 * 4);
    const s4 d = *(const s4 *)(x + ((size_t)(2 * h + 1) * Wi + 2 * w + 1) * WID + c4 * 4);
    s4 o;
#pragma unroll
    for (int i = 0; i < 4; i++)
      o[i] = (short)f2bf(0.25f * (((bf2f(a[i]) + bf2f(b[i])) + bf2f(c[i])) + bf2f(d[i])));
    *(s4 *)(out + (size_t)p * WID + c4 * 4) = o;
  }
}
// up + skip: out[h][w] = x[h/2][w/2] + skip[h][w]
__global__ void k_up_add(const bf16_t *__restrict__ x, const bf16_t *__restrict__ skip, bf16_t *__restrict__ out,
                         int H, int W) {
  const long n = (long)H * W * (WID / 4);
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const int c4 = e % (WID / 4);
    const long p = e / (WID / 4);
    const int h = (int)(p / W), w = (int)(p % W);
    const s4 a = *(const s4 *)(x + ((size_t)(h / 2) * (W / 2) + w / 2) * WID + c4 * 4);
    const s4 b = *(const s4 *)(skip + (size_t)p * WID + c4 * 4);
    s4 o;
#pragma unroll
    for (int i = 0; i < 4; i++) o[i] = (short)f2bf(bf2f(a[i]) + bf2f(b[i]));
    *(s4 *)(out + (size_t)p * WID + c4 * 4) = o;
  }
}
// gradient through up: dz[h][w] = (sum of du over the 2x2 children) where a > 0
__global__ void k_up_bwd_mask(const bf16_t *__restrict__ du, const bf16_t *__restrict__ a, bf16_t *__restrict__ dz,
                              int Ho, int Wo) {
  const long n = (long)Ho * Wo * (WID / 4);
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const int c4 = e % (WID / 4);
    const long p = e / (WID / 4);
    const int h = (int)(p / Wo), w = (int)(p % Wo), Wi = 2 * Wo;
    const s4 q0 = *(const s4 *)(du + ((size_t)(2 * h) * Wi + 2 * w) * WID + c4 * 4);
    const s4 q1 = *(const s4 *)(du + ((size_t)(2 * h) * Wi + 2 * w + 1) * WID + c4 * 4);
    const s4 q2 = *(const s4 *)(du + ((size_t)(2 * h + 1) * Wi + 2 * w) * WID + c4 * 4);
    const s4 q3 = *(const s4 *)(du + ((size_t)(2 * h + 1) * Wi + 2 * w + 1) * WID + c4 * 4);
    const s4 m = *(const s4 *)(a + (size_t)p * WID + c4 * 4);
    s4 o;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const float v = ((bf2f(q0[i]) + bf2f(q1[i])) + bf2f(q2[i])) + bf2f(q3[i]);
      o[i] = (short)f2bf(bf2f(m[i]) > 0.f ? v : 0.f);
    }
    *(s4 *)(dz + (size_t)p * WID + c4 * 4) = o;
  }
}
// gradient through pool plus the skip gradient: dz[h][w] = (0.25 dp[h/2][w/2] + ds[h][w]) where a > 0
__global__ void k_pool_bwd_add_mask(const bf16_t *__restrict__ dp, const bf16_t *__restrict__ ds,
                                    const bf16_t *__restrict__ a, bf16_t *__restrict__ dz, int H, int W) {
  const long n = (long)H * W * (WID / 4);
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const int c4 = e % (WID / 4);
    const long p = e / (WID / 4);
    const int h = (int)(p / W), w = (int)(p % W);
    const s4 q = *(const s4 *)(dp + ((size_t)(h / 2) * (W / 2) + w / 2) * WID + c4 * 4);
    const s4 s = *(const s4 *)(ds + (size_t)p * WID + c4 * 4);
    const s4 m = *(const s4 *)(a + (size_t)p * WID + c4 * 4);
    s4 o;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const float v = 0.25f * bf2f(q[i]) + bf2f(s[i]);
      o[i] = (short)f2bf(bf2f(m[i]) > 0.f ? v : 0.f);
    }
    *(s4 *)(dz + (size_t)p * WID + c4 * 4) = o;
  }
}

// loss = sum over rings 1..H-2 of (y - t tinv)^2 / n, dy = 2 (y - t tinv) / n
// (channel 0); tinv standardises the targets
__global__ void __launch_bounds__(256) k_loss(const float *__restrict__ y, const float *__restrict__ t, float tinv,
                                              int H, int W, bf16_t *__restrict__ dy, double *__restrict__ part) {
  __shared__ double red[256];
  const long n = (long)H * W;
  const double inv = 1.0 / ((double)(H - 2) * W);
  double s = 0.0;
  for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < n; p += (long)gridDim.x * 256) {
    const int h = (int)(p / W);
    float g = 0.f;
    if (h >= 1 && h <= H - 2) {
      const float d = y[p * COUTL] - t[p] * tinv;
      s += (double)d * d;
      g = (float)(2.0 * d * inv);
    }
    s4 z = (s4){(short)f2bf(g), 0, 0, 0};
    *(s4 *)(dy + p * COUTL) = z;
    *(s4 *)(dy + p * COUTL + 4) = (s4){0, 0, 0, 0};
    *(s4 *)(dy + p * COUTL + 8) = (s4){0, 0, 0, 0};
    *(s4 *)(dy + p * COUTL + 12) = (s4){0, 0, 0, 0};
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] * inv;
}

// Adam (FP32): m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2; w -= lr mhat / (sqrt(vhat) + eps)
__global__ void k_adam(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v, const float *__restrict__ g,
                       int n, float lr, float b1, float b2, float c1, float c2, float eps) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const float gg = g[e];
    const float mm = b1 * m[e] + (1.f - b1) * gg;
    const float vv = b2 * v[e] + (1.f - b2) * gg * gg;
    m[e] = mm;
    v[e] = vv;
    w[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
  }
}

// ---------------------------------------------------------------------------
// Features of the assembled system (rings 1..NRAD-2; zero on rings 0 and
// NRAD-1): the Jacobi correction -(A.T0 - RHS)/(Ml T0) scaled by its RMS s,
// the off-diagonals relative to the diagonal (x4) and log Ml.  The system
// carries T, Sigma, kappa, Q+, Q- and dt (EnergySources.c:256-297).
__global__ void __launch_bounds__(256) k_feat1(const double *__restrict__ T, const double *__restrict__ Ml,
                                               const double *__restrict__ Mlip, const double *__restrict__ Mlim,
                                               const double *__restrict__ Mljp, const double *__restrict__ Mljm,
                                               const double *__restrict__ RHS, int H, int W,
                                               float *__restrict__ jac, double *__restrict__ part) {
  __shared__ double red[256];
  const long n = (long)H * W;
  double s = 0.0;
  for (long l = (long)blockIdx.x * 256 + threadIdx.x; l < n; l += (long)gridDim.x * 256) {
    const int i = (int)(l / W), j = (int)(l % W);
    float v = 0.f;
    if (i >= 1 && i <= H - 2) {
      const long ljp = (j == W - 1) ? l - j : l + 1, ljm = (j == 0) ? l + W - 1 : l - 1;
      const double r = T[l] * Ml[l] + T[l + W] * Mlip[l] + T[l - W] * Mlim[l] + T[ljp] * Mljp[l] +
                       T[ljm] * Mljm[l] - RHS[l];
      const double jr = -r / Ml[l] / T[l];
      v = (float)jr;
      s += jr * jr;
    }
    jac[l] = v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ void __launch_bounds__(256) k_scale(const double *__restrict__ part, int nb, long nvalid,
                                               double *__restrict__ sc) {
  __shared__ double red[256];
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) s += part[b];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double v = sqrt(red[0] / (double)nvalid);
    sc[0] = v > 1e-300 ? v : 1e-300;
  }
}
__global__ void k_feat2(const float *__restrict__ jac, const double *__restrict__ Ml, const double *__restrict__ Mlip,
                        const double *__restrict__ Mlim, const double *__restrict__ Mljp,
                        const double *__restrict__ Mljm, int H, int W, const double *__restrict__ sc,
                        bf16_t *__restrict__ x) {
  const long n = (long)H * W;
  const float is = (float)(1.0 / sc[0]);
  for (long l = (long)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (long)gridDim.x * blockDim.x) {
    const int i = (int)(l / W);
    float f[CIN0];
    for (int c = 0; c < CIN0; c++) f[c] = 0.f;
    if (i >= 1 && i <= H - 2) {
      const double ml = Ml[l];
      f[0] = jac[l] * is;
      f[1] = (float)(4.0 * Mlip[l] / ml);
      f[2] = (float)(4.0 * Mlim[l] / ml);
      f[3] = (float)(4.0 * Mljp[l] / ml);
      f[4] = (float)(4.0 * Mljm[l] / ml);
      f[5] = (float)log(ml);
    }
    for (int c4 = 0; c4 < CIN0 / 4; c4++) {
      s4 o;
      for (int k = 0; k < 4; k++) o[k] = (short)f2bf(f[c4 * 4 + k]);
      *(s4 *)(x + l * CIN0 + c4 * 4) = o;
    }
  }
}
// training target (T* - T0)/(T0 s) on rings 1..NRAD-2
__global__ void k_target(const double *__restrict__ Tsol, const double *__restrict__ T0, int H, int W,
                         const double *__restrict__ sc, float *__restrict__ t) {
  const long n = (long)H * W;
  const double s = sc[0];
  for (long l = (long)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (long)gridDim.x * blockDim.x) {
    const int i = (int)(l / W);
    t[l] = (i >= 1 && i <= H - 2) ? (float)((Tsol[l] - T0[l]) / (T0[l] * s)) : 0.f;
  }
}
// Residual loss: with the guess T0 + e, e = y tau T0 s on rings 1..H-2 (0 on
// rings 0 and H-1), the normalised residual of cell l is
//   rho_l = -f0_l + tau (y_l + wip_l y_{l+W} + wim_l y_{l-W} + wjp_l y_{l+1} + wjm_l y_{l-1})
// with f0 = -(A.T0 - RHS)/(Ml T0 s) and w_nb = (M_nb/Ml)(T0_nb/T0_l) (aux
// planes 0..4); loss = sum rho^2 / n over rings 1..H-2.
__global__ void __launch_bounds__(256) k_rloss(const float *__restrict__ y, int ys, const float *__restrict__ aux,
                                               float tau, int H, int W, float *__restrict__ rho,
                                               double *__restrict__ part) {
  __shared__ double red[256];
  const long n = (long)H * W;
  const double inv = 1.0 / ((double)(H - 2) * W);
  double s = 0.0;
  for (long l = (long)blockIdx.x * 256 + threadIdx.x; l < n; l += (long)gridDim.x * 256) {
    const int i = (int)(l / W), j = (int)(l % W);
    float r = 0.f;
    if (i >= 1 && i <= H - 2) {
      const long ljp = (j == W - 1) ? l - j : l + 1, ljm = (j == 0) ? l + W - 1 : l - 1;
      const float yip = (i + 1 <= H - 2) ? y[(l + W) * ys] : 0.f;
      const float yim = (i - 1 >= 1) ? y[(l - W) * ys] : 0.f;
      const float a = y[l * ys] + aux[n + l] * yip + aux[2 * n + l] * yim + aux[3 * n + l] * y[ljp * ys] +
                      aux[4 * n + l] * y[ljm * ys];
      r = -aux[l] + tau * a;
      s += (double)r * r;
    }
    rho[l] = r;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] * inv;
}
// its gradient: dy_m = 2 tau / n (rho_m + wip_{m-W} rho_{m-W} + wim_{m+W} rho_{m+W}
//                                 + wjp_{m-1} rho_{m-1} + wjm_{m+1} rho_{m+1}) on rings 1..H-2
// (into dy as BF16 channel 0 of COUTL, or with gf != NULL into the float plane gf)
__global__ void k_rgrad(const float *__restrict__ rho, const float *__restrict__ aux, float tau, int H, int W,
                        bf16_t *__restrict__ dy, float *__restrict__ gf) {
  const long n = (long)H * W;
  const float c = (float)(2.0 / ((double)(H - 2) * W)) * tau;
  for (long m = (long)blockIdx.x * blockDim.x + threadIdx.x; m < n; m += (long)gridDim.x * blockDim.x) {
    const int i = (int)(m / W), j = (int)(m % W);
    float g = 0.f;
    if (i >= 1 && i <= H - 2) {
      const long mjm = (j == 0) ? m + W - 1 : m - 1, mjp = (j == W - 1) ? m - j : m + 1;
      g = rho[m] + aux[n + m - W] * rho[m - W] + aux[2 * n + m + W] * rho[m + W] + aux[3 * n + mjm] * rho[mjm] +
          aux[4 * n + mjp] * rho[mjp];
      g *= c;
    }
    if (gf) {
      gf[m] = g;
      continue;
    }
    *(s4 *)(dy + m * COUTL) = (s4){(short)f2bf(g), 0, 0, 0};
    *(s4 *)(dy + m * COUTL + 4) = (s4){0, 0, 0, 0};
    *(s4 *)(dy + m * COUTL + 8) = (s4){0, 0, 0, 0};
    *(s4 *)(dy + m * COUTL + 12) = (s4){0, 0, 0, 0};
  }
}
// aux planes of a training pair from the system, T0, the Jacobi feature and s
__global__ void k_aux(const float *__restrict__ jac, const double *__restrict__ T0, const double *__restrict__ Ml,
                      const double *__restrict__ Mlip, const double *__restrict__ Mlim, const double *__restrict__ Mljp,
                      const double *__restrict__ Mljm, int H, int W, const double *__restrict__ sc,
                      float *__restrict__ aux) {
  const long n = (long)H * W;
  const double is = 1.0 / sc[0];
  for (long l = (long)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (long)gridDim.x * blockDim.x) {
    const int i = (int)(l / W), j = (int)(l % W);
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f, a4 = 0.f;
    if (i >= 1 && i <= H - 2) {
      const long ljp = (j == W - 1) ? l - j : l + 1, ljm = (j == 0) ? l + W - 1 : l - 1;
      const double ml = Ml[l], t0 = T0[l];
      a0 = (float)((double)jac[l] * is);
      a1 = (float)(Mlip[l] / ml * (T0[l + W] / t0));
      a2 = (float)(Mlim[l] / ml * (T0[l - W] / t0));
      a3 = (float)(Mljp[l] / ml * (T0[ljp] / t0));
      a4 = (float)(Mljm[l] / ml * (T0[ljm] / t0));
    }
    aux[l] = a0; aux[n + l] = a1; aux[2 * n + l] = a2; aux[3 * n + l] = a3; aux[4 * n + l] = a4;
  }
}

// sum of squares (block partials)
__global__ void __launch_bounds__(256) k_sumsq(const float *__restrict__ t, long n, double *__restrict__ part) {
  __shared__ double red[256];
  double s = 0.0;
  for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < n; p += (long)gridDim.x * 256) s += (double)t[p] * t[p];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// warm start: Tg = T0 + y tscale T0 s on rings 1..NRAD-2, T0 elsewhere
__global__ void k_apply(const float *__restrict__ y, float tscale, const double *__restrict__ T0, int H, int W,
                        const double *__restrict__ sc, double *__restrict__ Tg) {
  const long n = (long)H * W;
  const double s = sc[0] * (double)tscale;
  for (long l = (long)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (long)gridDim.x * blockDim.x) {
    const int i = (int)(l / W);
    const double t0 = T0[l];
    Tg[l] = (i >= 1 && i <= H - 2) ? t0 + (double)y[l * COUTL] * t0 * s : t0;
  }
}

// ---------------------------------------------------------------------------
// Coarse network: restriction of the features and of the target (f x f
// means), bilinear interpolation of the output and its transpose.

// fine index i of an axis -> the two coarse cells whose centres bracket the
// fine cell's centre, u = (i + 0.5)/f - 0.5 in coarse units, and the weight
// w of the second (exact in FP32 for f = 1, 2, 4); across the rings the cells
// are clamped to 0..nc-1, along the sectors taken modulo nc
__device__ __forceinline__ void bil_axis(int i, int f, int nc, bool periodic, int &c0, int &c1, float &w) {
  const float u = ((float)i + 0.5f) / (float)f - 0.5f;
  const float fl = floorf(u);
  w = u - fl;
  int a = (int)fl, b = a + 1;
  if (periodic) {
    a = (a + nc) % nc;
    b = b % nc;
  } else {
    a = min(max(a, 0), nc - 1);
    b = min(max(b, 0), nc - 1);
  }
  c0 = a;
  c1 = b;
}
// bilinear value at fine cell (i, j) of the coarse plane y (stride ys)
__device__ __forceinline__ float bil_at(const float *__restrict__ y, int ys, int i, int j, int f, int Hc, int Wc) {
  int i0, i1, j0, j1;
  float wi, wj;
  bil_axis(i, f, Hc, false, i0, i1, wi);
  bil_axis(j, f, Wc, true, j0, j1, wj);
  const float r0 = (1.f - wj) * y[((size_t)i0 * Wc + j0) * ys] + wj * y[((size_t)i0 * Wc + j1) * ys];
  const float r1 = (1.f - wj) * y[((size_t)i1 * Wc + j0) * ys] + wj * y[((size_t)i1 * Wc + j1) * ys];
  return (1.f - wi) * r0 + wi * r1;
}
// features averaged over the f x f fine cells of each coarse cell (the
// fine features as k_feat2 forms them; 0 on rings 0 and H-1)
__global__ void k_feat2c(const float *__restrict__ jac, const double *__restrict__ Ml, const double *__restrict__ Mlip,
                         const double *__restrict__ Mlim, const double *__restrict__ Mljp,
                         const double *__restrict__ Mljm, int H, int W, int f, const double *__restrict__ sc,
                         bf16_t *__restrict__ x) {
  const int Hc = H / f, Wc = W / f;
  const long n = (long)Hc * Wc;
  const float is = (float)(1.0 / sc[0]);
  const float inv = 1.f / (float)(f * f);
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) {
    const int I = (int)(p / Wc), J = (int)(p % Wc);
    float acc[NFEAT];
    for (int c = 0; c < NFEAT; c++) acc[c] = 0.f;
    for (int di = 0; di < f; di++) {
      const int i = I * f + di;
      if (i < 1 || i > H - 2) continue;
      for (int dj = 0; dj < f; dj++) {
        const long l = (long)i * W + J * f + dj;
        const double ml = Ml[l];
        acc[0] += jac[l] * is;
        acc[1] += (float)(4.0 * Mlip[l] / ml);
        acc[2] += (float)(4.0 * Mlim[l] / ml);
        acc[3] += (float)(4.0 * Mljp[l] / ml);
        acc[4] += (float)(4.0 * Mljm[l] / ml);
        acc[5] += (float)log(ml);
      }
    }
    s4 o0, o1;
    for (int k = 0; k < 4; k++) o0[k] = (short)f2bf(acc[k] * inv);
    o1[0] = (short)f2bf(acc[4] * inv);
    o1[1] = (short)f2bf(acc[5] * inv);
    o1[2] = 0;
    o1[3] = 0;
    *(s4 *)(x + p * CIN0) = o0;
    *(s4 *)(x + p * CIN0 + 4) = o1;
    *(s4 *)(x + p * CIN0 + 8) = (s4){0, 0, 0, 0};
    *(s4 *)(x + p * CIN0 + 12) = (s4){0, 0, 0, 0};
  }
}
// target (T* - T0)/(T0 s) averaged over the f x f fine cells (0 on rings 0, H-1)
__global__ void k_target_c(const double *__restrict__ Tsol, const double *__restrict__ T0, int H, int W, int f,
                           const double *__restrict__ sc, float *__restrict__ t) {
  const int Hc = H / f, Wc = W / f;
  const long n = (long)Hc * Wc;
  const double s = sc[0];
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) {
    const int I = (int)(p / Wc), J = (int)(p % Wc);
    double acc = 0.0;
    for (int di = 0; di < f; di++) {
      const int i = I * f + di;
      if (i < 1 || i > H - 2) continue;
      for (int dj = 0; dj < f; dj++) {
        const long l = (long)i * W + J * f + dj;
        acc += (Tsol[l] - T0[l]) / (T0[l] * s);
      }
    }
    t[p] = (float)(acc / (double)(f * f));
  }
}
// yf = bilinear interpolation of channel 0 of y (Hc x Wc, COUTL channels) on the H x W grid
__global__ void k_upy(const float *__restrict__ y, int f, int H, int W, float *__restrict__ yf) {
  const int Hc = H / f, Wc = W / f;
  const long n = (long)H * W;
  for (long l = (long)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (long)gridDim.x * blockDim.x)
    yf[l] = bil_at(y, COUTL, (int)(l / W), (int)(l % W), f, Hc, Wc);
}
// transpose of k_upy: dy[p] (channel 0, BF16) = sum over the fine cells of
// their interpolation weight on coarse cell p times g; the other channels 0
__global__ void k_upy_adj(const float *__restrict__ g, int f, int H, int W, bf16_t *__restrict__ dy) {
  const int Hc = H / f, Wc = W / f;
  const long n = (long)Hc * Wc;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) {
    const int I = (int)(p / Wc), J = (int)(p % Wc);
    float acc = 0.f;
    for (int di = -f; di < 2 * f; di++) {
      const int i = I * f + di;
      if (i < 0 || i >= H) continue;
      int a0, a1;
      float wi;
      bil_axis(i, f, Hc, false, a0, a1, wi);
      const float wr = (a0 == I ? 1.f - wi : 0.f) + (a1 == I ? wi : 0.f);
      if (wr == 0.f) continue;
      for (int dj = -f; dj < 2 * f; dj++) {
        const int j = (J * f + dj + W) % W;
        int b0, b1;
        float wj;
        bil_axis(j, f, Wc, true, b0, b1, wj);
        const float wc = (b0 == J ? 1.f - wj : 0.f) + (b1 == J ? wj : 0.f);
        if (wc == 0.f) continue;
        acc += (wr * wc) * g[(size_t)i * W + j];
      }
    }
    *(s4 *)(dy + p * COUTL) = (s4){(short)f2bf(acc), 0, 0, 0};
    *(s4 *)(dy + p * COUTL + 4) = (s4){0, 0, 0, 0};
    *(s4 *)(dy + p * COUTL + 8) = (s4){0, 0, 0, 0};
    *(s4 *)(dy + p * COUTL + 12) = (s4){0, 0, 0, 0};
  }
}
// warm start of the coarse network: Tg = T0 + up(y) tscale T0 s on rings 1..H-2
__global__ void k_apply_up(const float *__restrict__ y, float tscale, const double *__restrict__ T0, int H, int W,
                           int f, const double *__restrict__ sc, double *__restrict__ Tg) {
  const int Hc = H / f, Wc = W / f;
  const long n = (long)H * W;
  const double s = sc[0] * (double)tscale;
  for (long l = (long)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (long)gridDim.x * blockDim.x) {
    const int i = (int)(l / W), j = (int)(l % W);
    const double t0 = T0[l];
    Tg[l] = (i >= 1 && i <= H - 2) ? t0 + (double)bil_at(y, COUTL, i, j, f, Hc, Wc) * t0 * s : t0;
  }
}

// ---------------------------------------------------------------------------
// host side
struct Layer {
  int cin, cout, lvl;
  size_t off;                 // offset of W (then b) in the parameter vector
};

struct Sample {
  int H, W;                   // the system's grid (the network's is H/f x W/f)
  bf16_t *x;                  // features on the network's grid
  float *t;                   // correction target on the network's grid
  float *aux;                 // residual-loss planes (f0, wip, wim, wjp, wjm) on the system's grid
};

struct thorin_cnn {
  hipStream_t s;
  Layer L[NL];
  size_t nparam;
  float *w, *g, *m, *v;       // parameters, gradient, Adam moments (device, FP32)
  bf16_t *wpk[NL], *wpkT[NL];
  int adam_t;
  float tscale;               // RMS of the training targets (standardisation), 1 before any pair
  double tss;                 // running sum of the pairs' mean squared targets
  float tinv;                 // 1/tscale during training, 1 for thorin_cnn_grad
  int f;                      // coarsening of the network's grid (1, 2 or 4)
  int H, W;                   // network grid of the current pass (0: none)
  int HF, WF;                 // system grid of the current pass (H = HF/f, W = WF/f)
  size_t cap;                 // pixels the activation buffers hold (grids up to that size)
  size_t capf;                // pixels the system-grid planes hold
  float *yf, *gf;             // f > 1: interpolated output and loss gradient on the system grid
  bf16_t *x, *a[9], *p1, *p2, *p3, *u0, *u1, *u2, *dy, *dz[9], *du0, *du1, *du2, *dp1, *dp2, *dp3;
  float *y, *t, *jac, *wpart, *bpart, *aux, *rho;
  int lossmode;               // 0: mean square of the correction error; 1: of the guess' residual
  double *lpart, *fpart, *sc;
  int nwblk;
  std::vector<Sample> samples;
};

static const int LVL[NL] = {0, 0, 1, 2, 3, 2, 1, 0, 0};

static int grid_for(long n, int bs = 256) {
  long g = (n + bs - 1) / bs;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  return (int)g;
}

static void free_acts(thorin_cnn *n) {
  bf16_t **bs[] = {&n->x, &n->p1, &n->p2, &n->p3, &n->u0, &n->u1, &n->u2, &n->dy, &n->du0, &n->du1,
                   &n->du2, &n->dp1, &n->dp2, &n->dp3};
  for (auto p : bs)
    if (*p) { (void)hipFree(*p); *p = nullptr; }
  for (int k = 0; k < 9; k++) {
    if (n->a[k]) { (void)hipFree(n->a[k]); n->a[k] = nullptr; }
    if (n->dz[k]) { (void)hipFree(n->dz[k]); n->dz[k] = nullptr; }
  }
  float **fs[] = {&n->y, &n->t, &n->jac, &n->wpart, &n->bpart, &n->aux, &n->rho, &n->yf, &n->gf};
  for (auto p : fs)
    if (*p) { (void)hipFree(*p); *p = nullptr; }
  double **ds[] = {&n->lpart, &n->fpart};
  for (auto p : ds)
    if (*p) { (void)hipFree(*p); *p = nullptr; }
  n->H = n->W = n->HF = n->WF = 0;
  n->cap = n->capf = 0;
}

// buffers for a system grid HF x WF: activations on the network's grid
// HF/f x WF/f, the features' and the loss's planes on the system grid; they
// grow to the largest grid met (the pairs of an ensemble have several sizes)
// and are reused for smaller ones
static int alloc_acts(thorin_cnn *n, int HF, int WF) {
  const int f = n->f;
  if (HF % f || WF % f || (HF / f) % 8 || (WF / f) % 8 || HF / f < 16 || WF / f < 16) {
    thorin_set_errorf("the warm-start network needs NRAD/%d and NSEC/%d divisible by 8 (three 2x2 poolings) "
                      "and >= 16; the grid is %dx%d", f, f, HF, WF);
    return -1;
  }
  const int H = HF / f, W = WF / f;
  const size_t px0 = (size_t)H * W, pxf = (size_t)HF * WF;
  if (px0 <= n->cap && pxf <= n->capf) {
    n->H = H;
    n->W = W;
    n->HF = HF;
    n->WF = WF;
    return 0;
  }
  free_acts(n);
  auto px = [&](int l) { return px0 >> (2 * l); };
  auto A = [&](bf16_t **p, size_t cnt) -> int { HCK(hipMalloc(p, cnt * sizeof(bf16_t))); return 0; };
  if (A(&n->x, px0 * CIN0)) return -1;
  for (int k = 1; k <= 8; k++)
    if (A(&n->a[k], px(LVL[k - 1]) * WID) || A(&n->dz[k], px(LVL[k - 1]) * WID)) return -1;
  if (A(&n->p1, px(1) * WID) || A(&n->p2, px(2) * WID) || A(&n->p3, px(3) * WID)) return -1;
  if (A(&n->u2, px(2) * WID) || A(&n->u1, px(1) * WID) || A(&n->u0, px(0) * WID)) return -1;
  if (A(&n->du2, px(2) * WID) || A(&n->du1, px(1) * WID) || A(&n->du0, px(0) * WID)) return -1;
  if (A(&n->dp3, px(3) * WID) || A(&n->dp2, px(2) * WID) || A(&n->dp1, px(1) * WID)) return -1;
  if (A(&n->dy, px0 * COUTL)) return -1;
  HCK(hipMalloc(&n->y, px0 * COUTL * sizeof(float)));
  HCK(hipMalloc(&n->t, px0 * sizeof(float)));
  HCK(hipMalloc(&n->jac, pxf * sizeof(float)));
  HCK(hipMalloc(&n->aux, 5 * pxf * sizeof(float)));
  HCK(hipMalloc(&n->rho, pxf * sizeof(float)));
  if (f > 1) {
    HCK(hipMalloc(&n->yf, pxf * sizeof(float)));
    HCK(hipMalloc(&n->gf, pxf * sizeof(float)));
  }
  n->nwblk = 512;
  HCK(hipMalloc(&n->wpart, (size_t)n->nwblk * 9 * 4 * 256 * sizeof(float)));
  HCK(hipMalloc(&n->bpart, (size_t)8192 * WID * sizeof(float)));
  HCK(hipMalloc(&n->lpart, 8192 * sizeof(double)));
  HCK(hipMalloc(&n->fpart, 8192 * sizeof(double)));
  n->H = H;
  n->W = W;
  n->HF = HF;
  n->WF = WF;
  n->cap = px0;
  n->capf = pxf;
  return 0;
}

static int pack_all(thorin_cnn *n) {
  for (int k = 0; k < NL; k++) {
    hipLaunchKernelGGL(k_pack, dim3(64), dim3(256), 0, n->s, (const float *)(n->w + n->L[k].off), n->L[k].cin,
                       n->L[k].cout, 0, n->wpk[k]);
    hipLaunchKernelGGL(k_pack, dim3(64), dim3(256), 0, n->s, (const float *)(n->w + n->L[k].off), n->L[k].cin,
                       n->L[k].cout, 1, n->wpkT[k]);
  }
  HCK(hipGetLastError());
  return 0;
}

// one convolution launch of the right instance
static const int CONV_MT = 4;      // 16-pixel tiles per wave
static int conv(thorin_cnn *n, int cin, int cout, const bf16_t *x, const bf16_t *wpk, const float *bias, void *out,
                int H, int W, int mode, const bf16_t *mask) {
  if (W >= CV_SEG4) {
    // rows of at least one 128-pixel segment: LDS-staged input, four rows per workgroup
    const dim3 g((unsigned)(((H + 3) / 4) * ((W + CV_SEG4 - 1) / CV_SEG4))), b(256);
    if (cin == 16 && cout == 32)
      hipLaunchKernelGGL((k_conv_lds4<16, 32>), g, b, 0, n->s, x, wpk, bias, out, H, W, mode, mask);
    else if (cin == 32 && cout == 32)
      hipLaunchKernelGGL((k_conv_lds4<32, 32>), g, b, 0, n->s, x, wpk, bias, out, H, W, mode, mask);
    else if (cin == 32 && cout == 16)
      hipLaunchKernelGGL((k_conv_lds4<32, 16>), g, b, 0, n->s, x, wpk, bias, out, H, W, mode, mask);
    else {
      thorin_set_error("no convolution instance for these channels");
      return -1;
    }
    HCK(hipGetLastError());
    return 0;
  }
  const long tiles = (long)H * ((W + 16 * CONV_MT - 1) / (16 * CONV_MT));
  const dim3 g((unsigned)((tiles + 3) / 4)), b(256);
  if (cin == 16 && cout == 32)
    hipLaunchKernelGGL((k_conv<16, 32, CONV_MT>), g, b, 0, n->s, x, wpk, bias, out, H, W, mode, mask);
  else if (cin == 32 && cout == 32)
    hipLaunchKernelGGL((k_conv<32, 32, CONV_MT>), g, b, 0, n->s, x, wpk, bias, out, H, W, mode, mask);
  else if (cin == 32 && cout == 16)
    hipLaunchKernelGGL((k_conv<32, 16, CONV_MT>), g, b, 0, n->s, x, wpk, bias, out, H, W, mode, mask);
  else {
    thorin_set_error("no convolution instance for these channels");
    return -1;
  }
  HCK(hipGetLastError());
  return 0;
}

static int wgrad(thorin_cnn *n, int k, const bf16_t *x, const bf16_t *dz, int H, int W) {
  const Layer &Lk = n->L[k];
  const int nseg = H * ((W + WG_SEG - 1) / WG_SEG);
  const int nblk = nseg < n->nwblk ? nseg : n->nwblk;
  if (Lk.cin == 16 && Lk.cout == 32)
    hipLaunchKernelGGL((k_wgrad<16, 32>), dim3(nblk), dim3(256), 0, n->s, x, dz, H, W, n->wpart);
  else if (Lk.cin == 32 && Lk.cout == 32)
    hipLaunchKernelGGL((k_wgrad<32, 32>), dim3(nblk), dim3(256), 0, n->s, x, dz, H, W, n->wpart);
  else
    hipLaunchKernelGGL((k_wgrad<32, 16>), dim3(nblk), dim3(256), 0, n->s, x, dz, H, W, n->wpart);
  float *dw = n->g + Lk.off;
  hipLaunchKernelGGL(k_wgrad_reduce, dim3(grid_for((long)9 * Lk.cin * Lk.cout)), dim3(256), 0, n->s,
                     (const float *)n->wpart, nblk, Lk.cin, Lk.cout, dw);
  const long npx = (long)H * W;
  const int lanes = 256 / Lk.cout;
  const int bb = grid_for(npx, lanes) < 4096 ? grid_for(npx, lanes) : 4096;
  hipLaunchKernelGGL(k_bgrad, dim3(bb), dim3(256), 0, n->s, dz, npx, Lk.cout, n->bpart);
  hipLaunchKernelGGL(k_bgrad_reduce, dim3(1), dim3(256), 0, n->s, (const float *)n->bpart, bb, Lk.cout,
                     dw + (size_t)9 * Lk.cin * Lk.cout);
  HCK(hipGetLastError());
  return 0;
}

static int forward(thorin_cnn *n) {
  const int H = n->H, W = n->W;
  auto Hl = [&](int l) { return H >> l; };
  auto Wl = [&](int l) { return W >> l; };
  auto bias = [&](int k) { return (const float *)(n->w + n->L[k].off + (size_t)9 * n->L[k].cin * n->L[k].cout); };
  if (conv(n, 16, 32, n->x, n->wpk[0], bias(0), n->a[1], H, W, 1, nullptr)) return -1;
  if (conv(n, 32, 32, n->a[1], n->wpk[1], bias(1), n->a[2], H, W, 1, nullptr)) return -1;
  hipLaunchKernelGGL(k_pool, dim3(grid_for((long)Hl(1) * Wl(1) * 8)), dim3(256), 0, n->s, (const bf16_t *)n->a[2],
                     n->p1, Hl(1), Wl(1));
  if (conv(n, 32, 32, n->p1, n->wpk[2], bias(2), n->a[3], Hl(1), Wl(1), 1, nullptr)) return -1;
  hipLaunchKernelGGL(k_pool, dim3(grid_for((long)Hl(2) * Wl(2) * 8)), dim3(256), 0, n->s, (const bf16_t *)n->a[3],
                     n->p2, Hl(2), Wl(2));
  if (conv(n, 32, 32, n->p2, n->wpk[3], bias(3), n->a[4], Hl(2), Wl(2), 1, nullptr)) return -1;
  hipLaunchKernelGGL(k_pool, dim3(grid_for((long)Hl(3) * Wl(3) * 8)), dim3(256), 0, n->s, (const bf16_t *)n->a[4],
                     n->p3, Hl(3), Wl(3));
  if (conv(n, 32, 32, n->p3, n->wpk[4], bias(4), n->a[5], Hl(3), Wl(3), 1, nullptr)) return -1;
  hipLaunchKernelGGL(k_up_add, dim3(grid_for((long)Hl(2) * Wl(2) * 8)), dim3(256), 0, n->s, (const bf16_t *)n->a[5],
                     (const bf16_t *)n->a[4], n->u2, Hl(2), Wl(2));
  if (conv(n, 32, 32, n->u2, n->wpk[5], bias(5), n->a[6], Hl(2), Wl(2), 1, nullptr)) return -1;
  hipLaunchKernelGGL(k_up_add, dim3(grid_for((long)Hl(1) * Wl(1) * 8)), dim3(256), 0, n->s, (const bf16_t *)n->a[6],
                     (const bf16_t *)n->a[3], n->u1, Hl(1), Wl(1));
  if (conv(n, 32, 32, n->u1, n->wpk[6], bias(6), n->a[7], Hl(1), Wl(1), 1, nullptr)) return -1;
  hipLaunchKernelGGL(k_up_add, dim3(grid_for((long)H * W * 8)), dim3(256), 0, n->s, (const bf16_t *)n->a[7],
                     (const bf16_t *)n->a[2], n->u0, H, W);
  if (conv(n, 32, 32, n->u0, n->wpk[7], bias(7), n->a[8], H, W, 1, nullptr)) return -1;
  if (conv(n, 32, 16, n->a[8], n->wpk[8], bias(8), n->y, H, W, 2, nullptr)) return -1;
  HCK(hipGetLastError());
  return 0;
}

// loss of the current y against n->t, gradient of every parameter into n->g
static int backward(thorin_cnn *n, double *loss) {
  const int H = n->H, W = n->W;
  auto Hl = [&](int l) { return H >> l; };
  auto Wl = [&](int l) { return W >> l; };
  // the residual loss of a coarse network is taken on the system grid
  const long nl = (n->lossmode == 1 && n->f > 1) ? (long)n->HF * n->WF : (long)H * W;
  const int lb = grid_for(nl) < 8192 ? grid_for(nl) : 8192;
  if (n->lossmode == 0) {
    hipLaunchKernelGGL(k_loss, dim3(lb), dim3(256), 0, n->s, (const float *)n->y, (const float *)n->t, n->tinv, H, W,
                       n->dy, n->lpart);
  } else if (n->f == 1) {
    const float tau = 1.0f / n->tinv;
    hipLaunchKernelGGL(k_rloss, dim3(lb), dim3(256), 0, n->s, (const float *)n->y, COUTL, (const float *)n->aux, tau,
                       H, W, n->rho, n->lpart);
    hipLaunchKernelGGL(k_rgrad, dim3(grid_for((long)H * W)), dim3(256), 0, n->s, (const float *)n->rho,
                       (const float *)n->aux, tau, H, W, n->dy, (float *)nullptr);
  } else {
    // y -> yf (bilinear), residual loss and its gradient gf on the system
    // grid, then dy = transpose of the interpolation applied to gf
    const float tau = 1.0f / n->tinv;
    const int HF = n->HF, WF = n->WF;
    hipLaunchKernelGGL(k_upy, dim3(grid_for(nl)), dim3(256), 0, n->s, (const float *)n->y, n->f, HF, WF, n->yf);
    hipLaunchKernelGGL(k_rloss, dim3(lb), dim3(256), 0, n->s, (const float *)n->yf, 1, (const float *)n->aux, tau,
                       HF, WF, n->rho, n->lpart);
    hipLaunchKernelGGL(k_rgrad, dim3(grid_for(nl)), dim3(256), 0, n->s, (const float *)n->rho,
                       (const float *)n->aux, tau, HF, WF, (bf16_t *)nullptr, n->gf);
    hipLaunchKernelGGL(k_upy_adj, dim3(grid_for((long)H * W)), dim3(256), 0, n->s, (const float *)n->gf, n->f, HF,
                       WF, n->dy);
  }
  // layer 9 (32 -> 16): weights from (a8, dy); input gradient masked by a8 > 0
  if (wgrad(n, 8, n->a[8], n->dy, H, W)) return -1;
  if (conv(n, 16, 32, n->dy, n->wpkT[8], nullptr, n->dz[8], H, W, 4, n->a[8])) return -1;
  // layer 8 (level 0): input u0 = up(a7) + a2
  if (wgrad(n, 7, n->u0, n->dz[8], H, W)) return -1;
  if (conv(n, 32, 32, n->dz[8], n->wpkT[7], nullptr, n->du0, H, W, 0, nullptr)) return -1;
  hipLaunchKernelGGL(k_up_bwd_mask, dim3(grid_for((long)Hl(1) * Wl(1) * 8)), dim3(256), 0, n->s,
                     (const bf16_t *)n->du0, (const bf16_t *)n->a[7], n->dz[7], Hl(1), Wl(1));
  // layer 7 (level 1): input u1 = up(a6) + a3
  if (wgrad(n, 6, n->u1, n->dz[7], Hl(1), Wl(1))) return -1;
  if (conv(n, 32, 32, n->dz[7], n->wpkT[6], nullptr, n->du1, Hl(1), Wl(1), 0, nullptr)) return -1;
  hipLaunchKernelGGL(k_up_bwd_mask, dim3(grid_for((long)Hl(2) * Wl(2) * 8)), dim3(256), 0, n->s,
                     (const bf16_t *)n->du1, (const bf16_t *)n->a[6], n->dz[6], Hl(2), Wl(2));
  // layer 6 (level 2): input u2 = up(a5) + a4
  if (wgrad(n, 5, n->u2, n->dz[6], Hl(2), Wl(2))) return -1;
  if (conv(n, 32, 32, n->dz[6], n->wpkT[5], nullptr, n->du2, Hl(2), Wl(2), 0, nullptr)) return -1;
  hipLaunchKernelGGL(k_up_bwd_mask, dim3(grid_for((long)Hl(3) * Wl(3) * 8)), dim3(256), 0, n->s,
                     (const bf16_t *)n->du2, (const bf16_t *)n->a[5], n->dz[5], Hl(3), Wl(3));
  // layer 5 (level 3): input p3 = pool(a4)
  if (wgrad(n, 4, n->p3, n->dz[5], Hl(3), Wl(3))) return -1;
  if (conv(n, 32, 32, n->dz[5], n->wpkT[4], nullptr, n->dp3, Hl(3), Wl(3), 0, nullptr)) return -1;
  hipLaunchKernelGGL(k_pool_bwd_add_mask, dim3(grid_for((long)Hl(2) * Wl(2) * 8)), dim3(256), 0, n->s,
                     (const bf16_t *)n->dp3, (const bf16_t *)n->du2, (const bf16_t *)n->a[4], n->dz[4], Hl(2), Wl(2));
  // layer 4 (level 2): input p2 = pool(a3)
  if (wgrad(n, 3, n->p2, n->dz[4], Hl(2), Wl(2))) return -1;
  if (conv(n, 32, 32, n->dz[4], n->wpkT[3], nullptr, n->dp2, Hl(2), Wl(2), 0, nullptr)) return -1;
  hipLaunchKernelGGL(k_pool_bwd_add_mask, dim3(grid_for((long)Hl(1) * Wl(1) * 8)), dim3(256), 0, n->s,
                     (const bf16_t *)n->dp2, (const bf16_t *)n->du1, (const bf16_t *)n->a[3], n->dz[3], Hl(1), Wl(1));
  // layer 3 (level 1): input p1 = pool(a2)
  if (wgrad(n, 2, n->p1, n->dz[3], Hl(1), Wl(1))) return -1;
  if (conv(n, 32, 32, n->dz[3], n->wpkT[2], nullptr, n->dp1, Hl(1), Wl(1), 0, nullptr)) return -1;
  hipLaunchKernelGGL(k_pool_bwd_add_mask, dim3(grid_for((long)H * W * 8)), dim3(256), 0, n->s, (const bf16_t *)n->dp1,
                     (const bf16_t *)n->du0, (const bf16_t *)n->a[2], n->dz[2], H, W);
  // layer 2 (level 0): input a1; input gradient masked by a1 > 0
  if (wgrad(n, 1, n->a[1], n->dz[2], H, W)) return -1;
  if (conv(n, 32, 32, n->dz[2], n->wpkT[1], nullptr, n->dz[1], H, W, 4, n->a[1])) return -1;
  // layer 1: input x
  if (wgrad(n, 0, n->x, n->dz[1], H, W)) return -1;
  HCK(hipGetLastError());
  if (loss) {
    std::vector<double> lp(lb);
    HCK(hipMemcpyAsync(lp.data(), n->lpart, sizeof(double) * lb, hipMemcpyDeviceToHost, n->s));
    HCK(hipStreamSynchronize(n->s));
    double s = 0.0;
    for (int b = 0; b < lb; b++) s += lp[b];
    *loss = s;
  }
  return 0;
}

// features of the context's assembled system and T (into n->x); scale in n->sc
static int features(thorin_cnn *n, const thorin_sysptrs &sp) {
  const int H = sp.nr, W = sp.ns;
  if (alloc_acts(n, H, W)) return -1;
  const long npx = (long)H * W;
  const int nb = grid_for(npx) < 8192 ? grid_for(npx) : 8192;
  hipLaunchKernelGGL(k_feat1, dim3(nb), dim3(256), 0, n->s, sp.T, sp.ml, sp.mlip, sp.mlim, sp.mljp, sp.mljm, sp.rhs,
                     H, W, n->jac, n->fpart);
  hipLaunchKernelGGL(k_scale, dim3(1), dim3(256), 0, n->s, (const double *)n->fpart, nb, (long)(H - 2) * W, n->sc);
  if (n->f == 1)
    hipLaunchKernelGGL(k_feat2, dim3(grid_for(npx)), dim3(256), 0, n->s, (const float *)n->jac, sp.ml, sp.mlip,
                       sp.mlim, sp.mljp, sp.mljm, H, W, (const double *)n->sc, n->x);
  else
    hipLaunchKernelGGL(k_feat2c, dim3(grid_for((long)n->H * n->W)), dim3(256), 0, n->s, (const float *)n->jac, sp.ml,
                       sp.mlip, sp.mlim, sp.mljp, sp.mljm, H, W, n->f, (const double *)n->sc, n->x);
  HCK(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// C ABI
extern "C" int thorin_cnn_create(unsigned long long seed, thorin_cnn **out) {
  if (!out) { thorin_set_error("null argument"); return -1; }
  thorin_cnn *n = new thorin_cnn();
  memset(n->wpk, 0, sizeof n->wpk);
  memset(n->wpkT, 0, sizeof n->wpkT);
  n->x = n->p1 = n->p2 = n->p3 = n->u0 = n->u1 = n->u2 = n->dy = nullptr;
  n->du0 = n->du1 = n->du2 = n->dp1 = n->dp2 = n->dp3 = nullptr;
  for (int k = 0; k < 9; k++) { n->a[k] = nullptr; n->dz[k] = nullptr; }
  n->y = n->t = n->jac = n->wpart = n->bpart = n->aux = n->rho = nullptr;
  n->lossmode = 0;
  n->lpart = n->fpart = nullptr;
  n->f = 1;
  n->yf = n->gf = nullptr;
  n->H = n->W = n->HF = n->WF = 0;
  n->cap = n->capf = 0;
  n->adam_t = 0;
  n->tscale = 1.0f;
  n->tss = 0.0;
  n->tinv = 1.0f;
  const int cin[NL] = {CIN0, WID, WID, WID, WID, WID, WID, WID, WID};
  const int cout[NL] = {WID, WID, WID, WID, WID, WID, WID, WID, COUTL};
  size_t off = 0;
  for (int k = 0; k < NL; k++) {
    n->L[k].cin = cin[k];
    n->L[k].cout = cout[k];
    n->L[k].lvl = LVL[k];
    n->L[k].off = off;
    off += (size_t)9 * cin[k] * cout[k] + cout[k];
  }
  n->nparam = off;
  // He-normal weights from a SplitMix64 stream (Box-Muller), zero biases;
  // only the used input channels and output channel 0 are nonzero
  std::vector<float> h(off, 0.f);
  unsigned long long st = seed;
  auto next = [&]() {
    unsigned long long z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  };
  for (int k = 0; k < NL; k++) {
    const int ci_used = k == 0 ? NFEAT : cin[k], co_used = k == NL - 1 ? 1 : cout[k];
    const double sd = sqrt(2.0 / (9.0 * ci_used)) * (k == NL - 1 ? 0.1 : 1.0);
    for (int t = 0; t < 9; t++)
      for (int ci = 0; ci < ci_used; ci++)
        for (int co = 0; co < co_used; co++) {
          const double u1 = next(), u2 = next();
          h[n->L[k].off + ((size_t)t * cin[k] + ci) * cout[k] + co] =
              (float)(sd * sqrt(-2.0 * log(u1)) * cos(2.0 * 3.14159265358979323846 * u2));
        }
  }
  if (hipStreamCreateWithFlags(&n->s, hipStreamNonBlocking) != hipSuccess) {
    thorin_set_error("hipStreamCreate failed");
    delete n;
    return -1;
  }
  HCK(hipMalloc(&n->w, off * sizeof(float)));
  HCK(hipMalloc(&n->g, off * sizeof(float)));
  HCK(hipMalloc(&n->m, off * sizeof(float)));
  HCK(hipMalloc(&n->v, off * sizeof(float)));
  HCK(hipMalloc(&n->sc, 8 * sizeof(double)));
  HCK(hipMemcpyAsync(n->w, h.data(), off * sizeof(float), hipMemcpyHostToDevice, n->s));
  HCK(hipMemsetAsync(n->m, 0, off * sizeof(float), n->s));
  HCK(hipMemsetAsync(n->v, 0, off * sizeof(float), n->s));
  HCK(hipMemsetAsync(n->g, 0, off * sizeof(float), n->s));
  for (int k = 0; k < NL; k++) {
    const size_t nw = (size_t)9 * cin[k] * cout[k];
    HCK(hipMalloc(&n->wpk[k], nw * sizeof(bf16_t)));
    HCK(hipMalloc(&n->wpkT[k], nw * sizeof(bf16_t)));
  }
  if (pack_all(n)) return -1;
  HCK(hipStreamSynchronize(n->s));
  *out = n;
  return 0;
}

extern "C" int thorin_cnn_destroy(thorin_cnn *n) {
  if (!n) return 0;
  (void)hipStreamSynchronize(n->s);
  free_acts(n);
  for (auto &s : n->samples) { (void)hipFree(s.x); (void)hipFree(s.t); (void)hipFree(s.aux); }
  float *fs[] = {n->w, n->g, n->m, n->v};
  for (auto p : fs)
    if (p) (void)hipFree(p);
  if (n->sc) (void)hipFree(n->sc);
  for (int k = 0; k < NL; k++) {
    if (n->wpk[k]) (void)hipFree(n->wpk[k]);
    if (n->wpkT[k]) (void)hipFree(n->wpkT[k]);
  }
  (void)hipStreamDestroy(n->s);
  delete n;
  return 0;
}

extern "C" size_t thorin_cnn_param_count(const thorin_cnn *n) { return n ? n->nparam : 0; }

extern "C" int thorin_cnn_get_params(thorin_cnn *n, float *out, size_t count) {
  if (!n || !out || count != n->nparam) { thorin_set_error("thorin_cnn_get_params: bad arguments"); return -1; }
  HCK(hipMemcpyAsync(out, n->w, count * sizeof(float), hipMemcpyDeviceToHost, n->s));
  HCK(hipStreamSynchronize(n->s));
  return 0;
}

extern "C" int thorin_cnn_set_params(thorin_cnn *n, const float *in, size_t count) {
  if (!n || !in || count != n->nparam) { thorin_set_error("thorin_cnn_set_params: bad arguments"); return -1; }
  HCK(hipMemcpyAsync(n->w, in, count * sizeof(float), hipMemcpyHostToDevice, n->s));
  if (pack_all(n)) return -1;
  HCK(hipStreamSynchronize(n->s));
  return 0;
}

// Forward and backward on host arrays (x: H*W*16 features, t: H*W target):
// loss, the parameter gradient (nparam floats) and the prediction y (H*W,
// channel 0).  For checks of the kernels against a CPU restatement.
extern "C" int thorin_cnn_grad(thorin_cnn *n, int H, int W, const float *x, const float *t, double *loss, float *grad,
                               float *y) {
  return thorin_cnn_grad2(n, H, W, x, t, nullptr, loss, grad, y);
}

// the same with aux != NULL (5 planes of H*W: f0, wip, wim, wjp, wjm): the
// residual loss with tau = 1 instead of the correction loss (t unused)
extern "C" int thorin_cnn_grad2(thorin_cnn *n, int H, int W, const float *x, const float *t, const float *aux,
                                double *loss, float *grad, float *y) {
  if (!n || !x || (!t && !aux)) { thorin_set_error("thorin_cnn_grad: null argument"); return -1; }
  if (alloc_acts(n, H, W)) return -1;
  const size_t npx = (size_t)n->H * n->W;        // the network's grid
  std::vector<bf16_t> xb(npx * CIN0);
  for (size_t e = 0; e < xb.size(); e++) xb[e] = h_f2bf(x[e]);
  HCK(hipMemcpyAsync(n->x, xb.data(), xb.size() * sizeof(bf16_t), hipMemcpyHostToDevice, n->s));
  if (t) HCK(hipMemcpyAsync(n->t, t, npx * sizeof(float), hipMemcpyHostToDevice, n->s));
  if (aux) HCK(hipMemcpyAsync(n->aux, aux, 5 * (size_t)H * W * sizeof(float), hipMemcpyHostToDevice, n->s));
  n->tinv = 1.0f;
  const int keep = n->lossmode;
  n->lossmode = aux ? 1 : 0;
  int rc = forward(n);
  if (!rc) rc = backward(n, loss);
  n->lossmode = keep;
  if (rc) return -1;
  if (grad) HCK(hipMemcpyAsync(grad, n->g, n->nparam * sizeof(float), hipMemcpyDeviceToHost, n->s));
  if (y) {
    std::vector<float> yy(npx * COUTL);
    HCK(hipMemcpyAsync(yy.data(), n->y, yy.size() * sizeof(float), hipMemcpyDeviceToHost, n->s));
    HCK(hipStreamSynchronize(n->s));
    for (size_t p = 0; p < npx; p++) y[p] = yy[p * COUTL];
  }
  HCK(hipStreamSynchronize(n->s));
  return 0;
}

// a training pair from the context: the features of the system assembled by
// thorin_rad_system (with its initial temperature T0) and the target from
// the temperature the last thorin_rad_solve left
extern "C" int thorin_cnn_add_pair(thorin_cnn *n, thorin_ctx *c) {
  thorin_sysptrs sp;
  if (!n || thorin_dev_system(c, &sp)) { if (!n) thorin_set_error("null network"); return -1; }
  HCK(hipStreamSynchronize((hipStream_t)sp.stream));   // the solve that left T
  thorin_sysptrs s0 = sp;
  s0.T = sp.T0;               // features of the initial temperature
  if (features(n, s0)) return -1;
  Sample sm;
  sm.H = sp.nr;
  sm.W = sp.ns;
  const size_t npf = (size_t)sp.nr * sp.ns;      // the system's grid
  const size_t npx = (size_t)n->H * n->W;        // the network's grid
  HCK(hipMalloc(&sm.x, npx * CIN0 * sizeof(bf16_t)));
  HCK(hipMalloc(&sm.t, npx * sizeof(float)));
  HCK(hipMalloc(&sm.aux, 5 * npf * sizeof(float)));
  hipLaunchKernelGGL(k_aux, dim3(grid_for((long)npf)), dim3(256), 0, n->s, (const float *)n->jac, sp.T0, sp.ml, sp.mlip,
                     sp.mlim, sp.mljp, sp.mljm, sp.nr, sp.ns, (const double *)n->sc, sm.aux);
  HCK(hipMemcpyAsync(sm.x, n->x, npx * CIN0 * sizeof(bf16_t), hipMemcpyDeviceToDevice, n->s));
  if (n->f == 1)
    hipLaunchKernelGGL(k_target, dim3(grid_for((long)npx)), dim3(256), 0, n->s, sp.T, sp.T0, sp.nr, sp.ns,
                       (const double *)n->sc, sm.t);
  else
    hipLaunchKernelGGL(k_target_c, dim3(grid_for((long)npx)), dim3(256), 0, n->s, sp.T, sp.T0, sp.nr, sp.ns, n->f,
                       (const double *)n->sc, sm.t);
  // running RMS of the targets
  const int nb = grid_for((long)npx) < 8192 ? grid_for((long)npx) : 8192;
  hipLaunchKernelGGL(k_sumsq, dim3(nb), dim3(256), 0, n->s, (const float *)sm.t, (long)npx, n->lpart);
  HCK(hipGetLastError());
  std::vector<double> part(nb);
  HCK(hipMemcpyAsync(part.data(), n->lpart, sizeof(double) * nb, hipMemcpyDeviceToHost, n->s));
  HCK(hipStreamSynchronize(n->s));
  double ss = 0.0;
  for (int b = 0; b < nb; b++) ss += part[b];
  n->samples.push_back(sm);
  n->tss += ss / ((double)(n->H - 2) * n->W);
  const double rms = sqrt(n->tss / (double)n->samples.size());
  n->tscale = rms > 0.0 ? (float)rms : 1.0f;
  return 0;
}

extern "C" int thorin_cnn_num_pairs(const thorin_cnn *n) { return n ? (int)n->samples.size() : 0; }

// iters Adam steps over the stored pairs in turn; the mean loss of the first
// and of the last pass over the pairs
extern "C" int thorin_cnn_train(thorin_cnn *n, int iters, float lr, double *loss_first, double *loss_last) {
  if (!n || n->samples.empty() || iters < 1) { thorin_set_error("thorin_cnn_train: no pairs"); return -1; }
  const int ns = (int)n->samples.size();
  {
    int big = 0;     // room for the largest pair first
    for (int k = 1; k < ns; k++)
      if ((size_t)n->samples[k].H * n->samples[k].W > (size_t)n->samples[big].H * n->samples[big].W) big = k;
    if (alloc_acts(n, n->samples[big].H, n->samples[big].W)) return -1;
  }
  double lf = 0.0, ll = 0.0;
  int cf = 0, cl = 0;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  n->tinv = 1.0f / n->tscale;
  for (int it = 0; it < iters; it++) {
    const Sample &sm = n->samples[it % ns];
    if (alloc_acts(n, sm.H, sm.W)) return -1;
    const size_t npx = (size_t)n->H * n->W, npf = (size_t)sm.H * sm.W;
    HCK(hipMemcpyAsync(n->x, sm.x, npx * CIN0 * sizeof(bf16_t), hipMemcpyDeviceToDevice, n->s));
    HCK(hipMemcpyAsync(n->t, sm.t, npx * sizeof(float), hipMemcpyDeviceToDevice, n->s));
    if (n->lossmode) HCK(hipMemcpyAsync(n->aux, sm.aux, 5 * npf * sizeof(float), hipMemcpyDeviceToDevice, n->s));
    if (forward(n)) return -1;
    const bool want = it < ns || it >= iters - ns;
    double L = 0.0;
    if (backward(n, want ? &L : nullptr)) return -1;
    if (it < ns) { lf += L; cf++; }
    if (it >= iters - ns) { ll += L; cl++; }
    n->adam_t++;
    const float c1 = 1.f - powf(b1, (float)n->adam_t), c2 = 1.f - powf(b2, (float)n->adam_t);
    hipLaunchKernelGGL(k_adam, dim3(grid_for((long)n->nparam)), dim3(256), 0, n->s, n->w, n->m, n->v,
                       (const float *)n->g, (int)n->nparam, lr, b1, b2, c1, c2, eps);
    if (pack_all(n)) return -1;
  }
  HCK(hipStreamSynchronize(n->s));
  if (loss_first) *loss_first = cf ? lf / cf : 0.0;
  if (loss_last) *loss_last = cl ? ll / cl : 0.0;
  return 0;
}

// warm start for thorin_rad_solve: Tg = T0 + prediction from the system and
// the context's current T (= T0), written to Tg (device, NRAD x NSEC doubles)
extern "C" int thorin_cnn_predict_dev(thorin_cnn *n, thorin_ctx *c, double *Tg) {
  thorin_sysptrs sp;
  if (!n || thorin_dev_system(c, &sp)) { if (!n) thorin_set_error("null network"); return -1; }
  // the network runs on the context's stream so that no extra sync is needed
  hipStream_t keep = n->s;
  n->s = (hipStream_t)sp.stream;
  int rc = features(n, sp);
  if (!rc) rc = forward(n);
  if (!rc) {
    const long npx = (long)sp.nr * sp.ns;
    if (n->f == 1)
      hipLaunchKernelGGL(k_apply, dim3(grid_for(npx)), dim3(256), 0, n->s, (const float *)n->y, n->tscale, sp.T,
                         sp.nr, sp.ns, (const double *)n->sc, Tg);
    else
      hipLaunchKernelGGL(k_apply_up, dim3(grid_for(npx)), dim3(256), 0, n->s, (const float *)n->y, n->tscale, sp.T,
                         sp.nr, sp.ns, n->f, (const double *)n->sc, Tg);
    if (hipGetLastError() != hipSuccess) { thorin_set_error("k_apply launch failed"); rc = -1; }
  }
  n->s = keep;
  return rc;
}

// training objective: 0 = mean square of (prediction - normalised correction),
// 1 = mean square of the normalised residual of the guess T0 + e
extern "C" int thorin_cnn_set_loss(thorin_cnn *n, int mode) {
  if (!n || mode < 0 || mode > 1) { thorin_set_error("thorin_cnn_set_loss: mode 0 or 1"); return -1; }
  n->lossmode = mode;
  return 0;
}

// the network's grid: the system's grid coarsened f = 1, 2 or 4 times along
// both axes; set before the first pair (pairs keep features of that grid)
extern "C" int thorin_cnn_set_coarsen(thorin_cnn *n, int f) {
  if (!n || (f != 1 && f != 2 && f != 4)) { thorin_set_error("thorin_cnn_set_coarsen: f is 1, 2 or 4"); return -1; }
  if (f == n->f) return 0;
  if (!n->samples.empty()) {
    thorin_set_error("thorin_cnn_set_coarsen: the network already holds pairs of another grid");
    return -1;
  }
  HCK(hipStreamSynchronize(n->s));
  free_acts(n);
  n->f = f;
  return 0;
}

extern "C" int thorin_cnn_coarsen(const thorin_cnn *n) { return n ? n->f : 0; }
