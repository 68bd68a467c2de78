// ams_surrogate.hip -- the placement surrogate's network: 3x3 convolutions
// (padding 1) over [B][C][64][64] maps with bf16 weights and FP32
// accumulation, their input and weight gradients, and the MSE loss.
//
// Two implementations of each convolution pass:
//
// MFMA (the default): implicit GEMMs on v_mfma_f32_16x16x32_bf16.  The
// forward is Y[co][p] = sum_k A[co][k] Xcol[p][k] with k = ci*9 + 3*ky + kx.
// Every FP32 value is split exactly into three bf16 parts, x = hi + mid +
// lo (24 significant bits), so the products carry FP32 precision: the
// weights are bf16 already (A exact), so the forward takes three MFMAs per
// 16 x 16 x 32 tile step (A.hi, A.mid, A.lo).  One workgroup per (image,
// 16 x 16 output tile): the input tile with its one-pixel halo is split once
// and staged in LDS ((hi << 16 | mid) words and lo halves); each wave owns
// four rows of the tile (four pixel tiles of 16) and every output channel,
// and gathers its Xcol fragments from LDS through a table of the offsets of
// k.  The input gradient is the same kernel with flipped, channel-transposed
// weights.  The weight gradient is dW[co][k] = sum_p dY[co][p] Xcol[p][k]
// with both operands split (the six products hi.hi, hi.mid, mid.hi, hi.lo,
// lo.hi, mid.mid; the rest are below FP32 rounding); one workgroup per
// (image, tile) writes its partial sums, a second kernel adds them in a
// fixed order.  Column k = Cin*9 of Xcol is 1, so that column of dW is the
// bias gradient.  (A two-part split, 2^-17 per value, put weight gradients
// 4e-3 off on random data, where the sums over 10^4 pixels cancel.)
//
// FP32 (comparison): one workgroup per (image, 16 x 16 output tile); the
// input tile with a one-pixel halo (Cin x 18 x 18 f32) and all weights (Cout
// x Cin x 9, bf16 -> f32) are staged in LDS; each thread owns one output
// pixel and accumulates every output channel in registers.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/ams.h"
#include "ams_frag.h"
#include "ams_internal.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int kTile = 16;
constexpr int kHalo = kTile + 2;
constexpr int kMaxC = 32;
constexpr int kHaloN = kHalo * kHalo;
constexpr int kMaxKb = (kMaxC * 9 + 31) / 32;      // reduction blocks of the forward (9)
constexpr int kMaxKt = (kMaxC * 9 + 1 + 15) / 16;  // 16-column tiles of dW with the bias column (19)
constexpr int kDwWaveTiles = (2 * kMaxKt + 3) / 4;  // dW tiles per wave (10)

__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float(((unsigned)h) << 16); }

// Y[b][co][y][x] = bias[co] + sum_{ci,ky,kx} W[co][ci][ky][kx] X[b][ci][y+ky-1][x+kx-1]
// (optional ReLU).  H, W multiples of 16; Cin, Cout <= 32.
__global__ void __launch_bounds__(256) k_conv3x3(const float *__restrict__ X, const unsigned short *__restrict__ Wt,
                                                 const float *__restrict__ bias, float *__restrict__ Y, int Cin,
                                                 int Cout, int H, int Wd, int relu) {
  __shared__ float sx[kMaxC][kHalo][kHalo + 1];
  __shared__ unsigned short sw[kMaxC * kMaxC * 9];  // bf16 weights (62 KB of LDS in all)
  const int tiles_x = Wd / kTile;
  const int b = blockIdx.y;
  const int ty0 = (blockIdx.x / tiles_x) * kTile, tx0 = (blockIdx.x % tiles_x) * kTile;
  const int tid = threadIdx.x;
  for (int e = tid; e < Cout * Cin * 9; e += 256) sw[e] = Wt[e];
  for (int e = tid; e < Cin * kHalo * kHalo; e += 256) {
    int ci = e / (kHalo * kHalo), r = e % (kHalo * kHalo);
    int yy = ty0 + r / kHalo - 1, xx = tx0 + r % kHalo - 1;
    float v = 0.0f;
    if (yy >= 0 && yy < H && xx >= 0 && xx < Wd) v = X[(((size_t)b * Cin + ci) * H + yy) * Wd + xx];
    sx[ci][r / kHalo][r % kHalo] = v;
  }
  __syncthreads();
  const int ly = tid / kTile, lx = tid % kTile;
  float acc[kMaxC];
#pragma unroll
  for (int co = 0; co < kMaxC; co++) acc[co] = 0.0f;
  for (int ci = 0; ci < Cin; ci++) {
    float v[9];
#pragma unroll
    for (int k = 0; k < 9; k++) v[k] = sx[ci][ly + k / 3][lx + k % 3];
#pragma unroll
    for (int co = 0; co < kMaxC; co++) {
      if (co < Cout) {
        const unsigned short *w = sw + (co * Cin + ci) * 9;
        float s = acc[co];
#pragma unroll
        for (int k = 0; k < 9; k++) s = fmaf(bf2f(w[k]), v[k], s);
        acc[co] = s;
      }
    }
  }
  const int y = ty0 + ly, x = tx0 + lx;
#pragma unroll
  for (int co = 0; co < kMaxC; co++) {
    if (co < Cout) {
      float o = acc[co] + (bias ? bias[co] : 0.0f);
      if (relu) o = fmaxf(o, 0.0f);
      Y[(((size_t)b * Cout + co) * H + y) * Wd + x] = o;
    }
  }
}

// dW[co][ci][k] = sum_{b,y,x} dY[b][co][y][x] X[b][ci][y+ky-1][x+kx-1];
// one workgroup per (co, ci); thread 0 of ci == 0 also sums db[co].
__global__ void __launch_bounds__(256) k_conv3x3_dw(const float *__restrict__ X, const float *__restrict__ dY,
                                                    float *__restrict__ dW, float *__restrict__ db, int B, int Cin,
                                                    int Cout, int H, int Wd) {
  __shared__ float red[10][8];
  const int co = blockIdx.x / Cin, ci = blockIdx.x % Cin;
  float acc[10];
#pragma unroll
  for (int k = 0; k < 10; k++) acc[k] = 0.0f;
  const long long n = (long long)B * H * Wd;
  for (long long p = threadIdx.x; p < n; p += blockDim.x) {
    int b = (int)(p / (H * Wd)), r = (int)(p % (H * Wd));
    int y = r / Wd, x = r % Wd;
    float g = dY[(((size_t)b * Cout + co) * H + y) * Wd + x];
    const float *xs = X + ((size_t)b * Cin + ci) * H * Wd;
#pragma unroll
    for (int k = 0; k < 9; k++) {
      int yy = y + k / 3 - 1, xx = x + k % 3 - 1;
      float v = (yy >= 0 && yy < H && xx >= 0 && xx < Wd) ? xs[yy * Wd + xx] : 0.0f;
      acc[k] = fmaf(g, v, acc[k]);
    }
    acc[9] += g;
  }
#pragma unroll
  for (int k = 0; k < 10; k++) {
    float v = acc[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 10) {
    float s = 0.0f;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += red[threadIdx.x][w];
    if (threadIdx.x < 9) dW[(co * Cin + ci) * 9 + threadIdx.x] = s;
    else if (ci == 0 && db) db[co] = s;
  }
}

__device__ __forceinline__ bf16x8 as_bf(u32x4 v) { return __builtin_bit_cast(bf16x8, v); }
#define MFMA(a, b, c) __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(a), as_bf(b), (c), 0, 0, 0)

// x = hi + mid + lo exactly (three bf16, rounded to nearest in turn): the
// word (hi << 16 | mid) and lo
__device__ __forceinline__ void split3(float x, unsigned &hm, unsigned short &lo) {
  const unsigned short h = ams::bf16_rne(x);
  const float r = x - ams::bf16_to_f(h);
  const unsigned short m = ams::bf16_rne(r);
  lo = ams::bf16_rne(r - ams::bf16_to_f(m));
  hm = ((unsigned)h << 16) | m;
}

// the hi, mid and lo fragments of 8 split values (element e in the half
// (e & 1) of word e / 2)
__device__ __forceinline__ void split_frags(const unsigned (&hm)[8], const unsigned short (&l)[8], u32x4 &fh,
                                            u32x4 &fm, u32x4 &fl) {
#pragma unroll
  for (int q = 0; q < 4; q++) {
    fh[q] = (hm[2 * q] >> 16) | (hm[2 * q + 1] & 0xffff0000u);
    fm[q] = (hm[2 * q] & 0xffffu) | (hm[2 * q + 1] << 16);
    fl[q] = (unsigned)l[2 * q] | ((unsigned)l[2 * q + 1] << 16);
  }
}

// Stage the input tile with its halo, split (zero outside the map).
__device__ __forceinline__ void stage_split_tile(const float *__restrict__ X, unsigned *shm, unsigned short *sl,
                                                 int b, int Cin, int H, int Wd, int ty0, int tx0) {
  for (int e = threadIdx.x; e < Cin * kHaloN; e += blockDim.x) {
    const int ci = e / kHaloN, r = e % kHaloN;
    const int yy = ty0 + r / kHalo - 1, xx = tx0 + r % kHalo - 1;
    float v = 0.0f;
    if (yy >= 0 && yy < H && xx >= 0 && xx < Wd) v = X[(((size_t)b * Cin + ci) * H + yy) * Wd + xx];
    split3(v, shm[e], sl[e]);
  }
}

// A fragments of the weights: A[co][k] = W[co][k] (bf16, [Cout][Cin*9]),
// rows padded to 16 and k to 32; item t = (ct * nkb + kb) * 64 + lane.
__global__ void __launch_bounds__(256) k_conv_wfrag(const unsigned short *__restrict__ W, int Cout, int Cin, int nkb,
                                                    int n_items, u32x4 *__restrict__ frag) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_items) return;
  const int lane = t & 63, blk = t >> 6, kb = blk % nkb, ct = blk / nkb;
  const int co = ct * 16 + (lane & 15), K = Cin * 9;
  unsigned short h[8];
#pragma unroll
  for (int e = 0; e < 8; e++) {
    const int k = kb * 32 + 8 * (lane >> 4) + e;
    h[e] = (co < Cout && k < K) ? W[(size_t)co * K + k] : (unsigned short)0;
  }
  u32x4 w;
  w.x = h[0] | ((unsigned)h[1] << 16);
  w.y = h[2] | ((unsigned)h[3] << 16);
  w.z = h[4] | ((unsigned)h[5] << 16);
  w.w = h[6] | ((unsigned)h[7] << 16);
  frag[t] = w;
}

// Y = conv(X) + bias (optional ReLU) on MFMA; NCT = output-channel tiles of 16.
template <int NCT>
__global__ void __launch_bounds__(256) k_conv3x3_mfma(const float *__restrict__ X, const u32x4 *__restrict__ Wf,
                                                      const float *__restrict__ bias, float *__restrict__ Y, int Cin,
                                                      int Cout, int H, int Wd, int relu) {
  __shared__ unsigned shm[kMaxC * kHaloN];
  __shared__ unsigned short sl[kMaxC * kHaloN];
  __shared__ int koff[kMaxKb * 32];  // LDS offset of k = (ci, ky, kx) in the tile, -1 past Cin*9
  const int tiles_x = Wd / kTile, b = blockIdx.y;
  const int ty0 = (blockIdx.x / tiles_x) * kTile, tx0 = (blockIdx.x % tiles_x) * kTile;
  const int nkb = (Cin * 9 + 31) / 32;
  stage_split_tile(X, shm, sl, b, Cin, H, Wd, ty0, tx0);
  for (int k = threadIdx.x; k < nkb * 32; k += blockDim.x)
    koff[k] = k < Cin * 9 ? (k / 9) * kHaloN + ((k % 9) / 3) * kHalo + (k % 3) : -1;
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, px = lane & 15, g = lane >> 4;
  f32x4 acc[NCT][4];
#pragma unroll
  for (int c = 0; c < NCT; c++)
#pragma unroll
    for (int t = 0; t < 4; t++) acc[c][t] = (f32x4){0, 0, 0, 0};
  for (int kb = 0; kb < nkb; kb++) {
    u32x4 a[NCT];
#pragma unroll
    for (int c = 0; c < NCT; c++) a[c] = Wf[((size_t)c * nkb + kb) * 64 + lane];
    int ko[8];
#pragma unroll
    for (int e = 0; e < 8; e++) ko[e] = koff[kb * 32 + 8 * g + e];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const int pix = (4 * w + t) * kHalo + px;
      unsigned hm[8];
      unsigned short l[8];
#pragma unroll
      for (int e = 0; e < 8; e++) {
        hm[e] = ko[e] >= 0 ? shm[ko[e] + pix] : 0u;
        l[e] = ko[e] >= 0 ? sl[ko[e] + pix] : (unsigned short)0;
      }
      u32x4 bh, bm, bl;
      split_frags(hm, l, bh, bm, bl);
#pragma unroll
      for (int c = 0; c < NCT; c++) {
        acc[c][t] = MFMA(a[c], bh, acc[c][t]);
        acc[c][t] = MFMA(a[c], bm, acc[c][t]);
        acc[c][t] = MFMA(a[c], bl, acc[c][t]);
      }
    }
  }
  // D layout: lane l holds output channels 16 c + 4 (l >> 4) + v of pixel l & 15
#pragma unroll
  for (int c = 0; c < NCT; c++)
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int co = c * 16 + 4 * g + v;
        if (co < Cout) {
          float o = acc[c][t][v] + (bias ? bias[co] : 0.0f);
          if (relu) o = fmaxf(o, 0.0f);
          Y[(((size_t)b * Cout + co) * H + ty0 + 4 * w + t) * Wd + tx0 + px] = o;
        }
      }
}

// Partial dW of one (image, tile): part[pid][co][k], k < 16 n_kt, column
// Cin*9 the bias gradient.  NCT = output-channel tiles; wave w owns the
// tiles j = w, w + 4, ... of the NCT x n_kt output tiles.
template <int NCT>
__global__ void __launch_bounds__(256) k_conv3x3_dw_mfma(const float *__restrict__ X, const float *__restrict__ dY,
                                                         float *__restrict__ part, int Cin, int Cout, int H, int Wd,
                                                         int n_kt) {
  __shared__ unsigned shm[kMaxC * kHaloN];
  __shared__ unsigned short sl[kMaxC * kHaloN];
  __shared__ int koff[kMaxKt * 16];  // as the forward's; -2 at the bias column (Xcol = 1)
  const int n_tiles = (H / kTile) * (Wd / kTile), tiles_x = Wd / kTile, b = blockIdx.y;
  const int ty0 = (blockIdx.x / tiles_x) * kTile, tx0 = (blockIdx.x % tiles_x) * kTile;
  const int K = Cin * 9, Kp = 16 * n_kt;
  stage_split_tile(X, shm, sl, b, Cin, H, Wd, ty0, tx0);
  for (int k = threadIdx.x; k < Kp; k += blockDim.x)
    koff[k] = k < K ? (k / 9) * kHaloN + ((k % 9) / 3) * kHalo + (k % 3) : (k == K ? -2 : -1);
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4;
  const int n_out = NCT * n_kt;
  f32x4 acc[kDwWaveTiles];
#pragma unroll
  for (int j = 0; j < kDwWaveTiles; j++) acc[j] = (f32x4){0, 0, 0, 0};
  for (int pb = 0; pb < (kTile * kTile) / 32; pb++) {
    // pixels q = 32 pb + 8 g + e of the tile: row q >> 4, columns px0 .. px0 + 7
    const int q0 = 32 * pb + 8 * g, py = q0 >> 4, px0 = q0 & 15;
    u32x4 ah[NCT], am[NCT], al[NCT];
#pragma unroll
    for (int c = 0; c < NCT; c++) {
      const int co = c * 16 + (lane & 15);
      float d[8];
      if (co < Cout) {
        const float4 *src = reinterpret_cast<const float4 *>(
            dY + (((size_t)b * Cout + co) * H + ty0 + py) * Wd + tx0 + px0);
        const float4 d0 = src[0], d1 = src[1];
        d[0] = d0.x;
        d[1] = d0.y;
        d[2] = d0.z;
        d[3] = d0.w;
        d[4] = d1.x;
        d[5] = d1.y;
        d[6] = d1.z;
        d[7] = d1.w;
      } else {
#pragma unroll
        for (int e = 0; e < 8; e++) d[e] = 0.0f;
      }
      unsigned hm[8];
      unsigned short l[8];
#pragma unroll
      for (int e = 0; e < 8; e++) split3(d[e], hm[e], l[e]);
      split_frags(hm, l, ah[c], am[c], al[c]);
    }
#pragma unroll
    for (int jj = 0; jj < kDwWaveTiles; jj++) {
      const int j = w + 4 * jj;
      if (j < n_out) {
        const int c = j / n_kt, kt = j % n_kt;
        const int ko = koff[kt * 16 + (lane & 15)];
        const int base = py * kHalo + px0;
        unsigned hm[8];
        unsigned short l[8];
#pragma unroll
        for (int e = 0; e < 8; e++) {
          hm[e] = ko >= 0 ? shm[ko + base + e] : (ko == -2 ? 0x3f800000u : 0u);  // 1.0 = (0x3f80, 0, 0)
          l[e] = ko >= 0 ? sl[ko + base + e] : (unsigned short)0;
        }
        u32x4 bh, bm, bl;
        split_frags(hm, l, bh, bm, bl);
        f32x4 r = acc[jj];
        r = MFMA(ah[c], bh, r);
        r = MFMA(ah[c], bm, r);
        r = MFMA(am[c], bh, r);
        r = MFMA(ah[c], bl, r);
        r = MFMA(al[c], bh, r);
        r = MFMA(am[c], bm, r);
        acc[jj] = r;
      }
    }
  }
  const long long pid = (long long)b * n_tiles + blockIdx.x;
#pragma unroll
  for (int jj = 0; jj < kDwWaveTiles; jj++) {
    const int j = w + 4 * jj;
    if (j < n_out) {
      const int c = j / n_kt, kt = j % n_kt;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int co = c * 16 + 4 * g + v;
        if (co < Cout) part[((size_t)pid * Cout + co) * Kp + kt * 16 + (lane & 15)] = acc[jj][v];
      }
    }
  }
}

// dW[co][k] = sum_p part[p][co][k] (k < Cin*9), db[co] = the bias column;
// 16 columns x 16 partial groups per workgroup, groups added in order.
__global__ void __launch_bounds__(256) k_conv_dw_reduce(const float *__restrict__ part, long long n_part, int Cout,
                                                        int Cin, int Kp, float *__restrict__ dW,
                                                        float *__restrict__ db) {
  __shared__ float red[16][17];
  const int kl = threadIdx.x & 15, grp = threadIdx.x >> 4;
  const int co = blockIdx.y, k = blockIdx.x * 16 + kl, K = Cin * 9;
  float s = 0.0f;
  if (k <= K)
    for (long long p = grp; p < n_part; p += 16) s += part[((size_t)p * Cout + co) * Kp + k];
  red[grp][kl] = s;
  __syncthreads();
  if (grp != 0 || k > K) return;
  float t = red[0][kl];
  for (int q = 1; q < 16; q++) t += red[q][kl];
  if (k < K) dW[(size_t)co * K + k] = t;
  else if (db) db[co] = t;
}

// dX_pre = dX * (Y > 0): the ReLU's backward on a layer output.
__global__ void k_relu_bwd(const float *__restrict__ Y, float *__restrict__ g, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && !(Y[i] > 0.0f)) g[i] = 0.0f;
}

// loss partials of mean((pred - target)^2 * mask) and its gradient.
__global__ void k_mse(const float *__restrict__ pred, const float *__restrict__ target,
                      const float *__restrict__ mask, long long n, float inv_n, float *__restrict__ grad,
                      float *__restrict__ partial) {
  __shared__ float red[8];
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  float s = 0.0f;
  if (i < n) {
    float m = mask ? mask[i] : 1.0f;
    float d = (pred[i] - target[i]) * m;
    s = d * d;
    grad[i] = 2.0f * d * m * inv_n;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.0f;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += red[w];
    partial[blockIdx.x] = t;
  }
}

// f32 -> bf16 (round to nearest even) for the weight copies.
__global__ void k_to_bf16(const float *__restrict__ x, unsigned short *__restrict__ y, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  unsigned u = __float_as_uint(x[i]);
  if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);
  y[i] = (unsigned short)(u >> 16);
}

int hip_status(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ams::set_error(std::string(what) + ": " + hipGetErrorString(e));
    return AMS_ERR_HIP;
  }
  return AMS_OK;
}

}  // namespace

extern "C" {

int ams_conv3x3_fwd(const float *X, const void *Wbf16, const float *bias, float *Y, int32_t B, int32_t Cin,
                    int32_t Cout, int32_t H, int32_t W, int32_t relu, void *stream) {
  if (!X || !Wbf16 || !Y || B <= 0 || Cin <= 0 || Cin > kMaxC || Cout <= 0 || Cout > kMaxC || H <= 0 || W <= 0 ||
      H % kTile || W % kTile) {
    ams::set_error("ams_conv3x3_fwd: bad arguments (channels <= 32, H and W multiples of 16)");
    return AMS_ERR_ARG;
  }
  dim3 grid((H / kTile) * (W / kTile), B);
  k_conv3x3<<<grid, 256, 0, (hipStream_t)stream>>>(X, (const unsigned short *)Wbf16, bias, Y, Cin, Cout, H, W, relu);
  return hip_status("ams_conv3x3_fwd");
}

int ams_conv3x3_bwd_weight(const float *X, const float *dY, float *dW, float *db, int32_t B, int32_t Cin,
                           int32_t Cout, int32_t H, int32_t W, void *stream) {
  if (!X || !dY || !dW || B <= 0 || Cin <= 0 || Cout <= 0 || H <= 0 || W <= 0) {
    ams::set_error("ams_conv3x3_bwd_weight: bad arguments");
    return AMS_ERR_ARG;
  }
  k_conv3x3_dw<<<Cout * Cin, 256, 0, (hipStream_t)stream>>>(X, dY, dW, db, B, Cin, Cout, H, W);
  return hip_status("ams_conv3x3_bwd_weight");
}

int64_t ams_conv3x3_wfrag_bytes(int32_t Cout, int32_t Cin) {
  if (Cout <= 0 || Cin <= 0) return 0;
  return (int64_t)((Cout + 15) / 16) * ((Cin * 9 + 31) / 32) * 64 * 16;
}

int ams_conv3x3_pack_weights(const void *Wbf16, int32_t Cout, int32_t Cin, void *frag, void *stream) {
  if (!Wbf16 || !frag || Cout <= 0 || Cout > kMaxC || Cin <= 0 || Cin > kMaxC) {
    ams::set_error("ams_conv3x3_pack_weights: bad arguments (channels <= 32)");
    return AMS_ERR_ARG;
  }
  const int nkb = (Cin * 9 + 31) / 32, n_items = ((Cout + 15) / 16) * nkb * 64;
  k_conv_wfrag<<<(n_items + 255) / 256, 256, 0, (hipStream_t)stream>>>((const unsigned short *)Wbf16, Cout, Cin, nkb,
                                                                       n_items, (u32x4 *)frag);
  return hip_status("ams_conv3x3_pack_weights");
}

int ams_conv3x3_fwd_mfma(const float *X, const void *Wfrag, const float *bias, float *Y, int32_t B, int32_t Cin,
                         int32_t Cout, int32_t H, int32_t W, int32_t relu, void *stream) {
  if (!X || !Wfrag || !Y || B <= 0 || B > 65535 || Cin <= 0 || Cin > kMaxC || Cout <= 0 || Cout > kMaxC || H <= 0 ||
      W <= 0 || H % kTile || W % kTile) {
    ams::set_error("ams_conv3x3_fwd_mfma: bad arguments (channels <= 32, H and W multiples of 16)");
    return AMS_ERR_ARG;
  }
  dim3 grid((H / kTile) * (W / kTile), B);
  hipStream_t st = (hipStream_t)stream;
  if (Cout > 16)
    k_conv3x3_mfma<2><<<grid, 256, 0, st>>>(X, (const u32x4 *)Wfrag, bias, Y, Cin, Cout, H, W, relu);
  else
    k_conv3x3_mfma<1><<<grid, 256, 0, st>>>(X, (const u32x4 *)Wfrag, bias, Y, Cin, Cout, H, W, relu);
  return hip_status("ams_conv3x3_fwd_mfma");
}

int64_t ams_conv3x3_dw_workspace_bytes(int32_t B, int32_t Cin, int32_t Cout, int32_t H, int32_t W) {
  if (B <= 0 || Cin <= 0 || Cout <= 0 || H <= 0 || W <= 0) return 0;
  const int64_t n_kt = (Cin * 9 + 1 + 15) / 16;
  return (int64_t)B * (H / kTile) * (W / kTile) * Cout * 16 * n_kt * 4;
}

int ams_conv3x3_bwd_weight_mfma(const float *X, const float *dY, float *dW, float *db, int32_t B, int32_t Cin,
                                int32_t Cout, int32_t H, int32_t W, void *workspace, void *stream) {
  if (!X || !dY || !dW || !workspace || B <= 0 || B > 65535 || Cin <= 0 || Cin > kMaxC || Cout <= 0 ||
      Cout > kMaxC || H <= 0 || W <= 0 || H % kTile || W % kTile) {
    ams::set_error("ams_conv3x3_bwd_weight_mfma: bad arguments (channels <= 32, H and W multiples of 16)");
    return AMS_ERR_ARG;
  }
  const int n_kt = (Cin * 9 + 1 + 15) / 16, n_tiles = (H / kTile) * (W / kTile);
  dim3 grid(n_tiles, B);
  hipStream_t st = (hipStream_t)stream;
  float *part = (float *)workspace;
  if (Cout > 16)
    k_conv3x3_dw_mfma<2><<<grid, 256, 0, st>>>(X, dY, part, Cin, Cout, H, W, n_kt);
  else
    k_conv3x3_dw_mfma<1><<<grid, 256, 0, st>>>(X, dY, part, Cin, Cout, H, W, n_kt);
  k_conv_dw_reduce<<<dim3((Cin * 9 + 1 + 15) / 16, Cout), 256, 0, st>>>(part, (long long)B * n_tiles, Cout, Cin,
                                                                        16 * n_kt, dW, db);
  return hip_status("ams_conv3x3_bwd_weight_mfma");
}

int ams_relu_bwd(const float *Y, float *grad, int64_t n, void *stream) {
  if (!Y || !grad || n <= 0) {
    ams::set_error("ams_relu_bwd: bad arguments");
    return AMS_ERR_ARG;
  }
  k_relu_bwd<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(Y, grad, n);
  return hip_status("ams_relu_bwd");
}

int64_t ams_mse_partials(int64_t n) { return (n + 255) / 256; }

int ams_mse(const float *pred, const float *target, const float *mask, int64_t n, float *grad, float *partial,
            void *stream) {
  if (!pred || !target || !grad || !partial || n <= 0) {
    ams::set_error("ams_mse: bad arguments");
    return AMS_ERR_ARG;
  }
  k_mse<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(pred, target, mask, n, 1.0f / (float)n, grad,
                                                                       partial);
  return hip_status("ams_mse");
}

int ams_to_bf16(const float *x, void *y, int64_t n, void *stream) {
  if (!x || !y || n <= 0) {
    ams::set_error("ams_to_bf16: bad arguments");
    return AMS_ERR_ARG;
  }
  k_to_bf16<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(x, (unsigned short *)y, n);
  return hip_status("ams_to_bf16");
}

}  // extern "C"
