// ams_surrogate.hip -- the placement surrogate's network: 3x3 convolutions
// (padding 1) over [B][C][64][64] maps with bf16 weights and FP32
// accumulation, their input and weight gradients, and the MSE loss.
//
// Forward: one workgroup per (image, 16 x 16 output tile); the input tile
// with a one-pixel halo (Cin x 18 x 18 f32) and all weights (Cout x Cin x 9,
// bf16 -> f32) are staged in LDS; each thread owns one output pixel and
// accumulates every output channel in registers.  The input gradient is
// the same kernel run with flipped, channel-transposed weights.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/ams.h"
#include "ams_internal.h"

namespace {

constexpr int kTile = 16;
constexpr int kHalo = kTile + 2;
constexpr int kMaxC = 32;

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
