// Fused phase-map operand of one metasurface with single-bounce chains, and
// its backward (go.py:107-163 for one metasurface, max_ms_reflex = 1).
//
// Rows r = m*Cn + c (map m, code c) of the stacked operand A [nm*Cn][K+M]:
//   A[r][k]     = WG[c][k] * u[m][k]          k < K,   WG = W . G
//   A[r][K + j] = W[c][j]
//   W[c][j] = amp[c][j] e^{-i th_ap[c][j]}          (go.py:163)
//   u[m][k] = ms_amp[k] e^{i rem(th_ms[m][k], P)}   (go.py:131, P = max phase)
// so that y = A . [F ; H_d] is the received signal of every (map, code).
//
// The backward takes gA in torch's convention for complex tensors (the
// gradient with respect to the conjugate, as torch's autograd hands it to a
// complex input) and returns the gradients of the real phases:
//   gu[m][k]  = sum_c gA[r][k] conj(WG[c][k])        g_ms = Im(conj(u) gu)
//   gWG[c][k] = sum_m gA[r][k] conj(u[m][k])
//   gW[c][j]  = sum_m gA[r][K+j] + sum_k gWG[c][k] conj(G[j][k])
//   g_ap      = -Im(conj(W) gW)
// (d u / d th = i u and d W / d th = -i W; the remainder passes the gradient
// through unchanged, as torch.remainder's backward does).
//
// Chains over several metasurfaces (go.py:107-157, max_ms_reflex >= 1): the
// surfaces' elements are concatenated (K = sum K_b, period[k] the max phase of
// element k's surface) and T is the [K][K] block matrix of the ms-to-ms
// channels (block (a, b) = ms{a}->ms{b}, zero on the diagonal; symmetric,
// since ms{b}->ms{a} is the transpose of ms{a}->ms{b}).  Then
//   P_1 = WG o u,   P_r = (P_{r-1} . T) o u,   A[:, :K] = sum_r P_r
// (o: per row of map m, u[m]), the products P_{r-1} . T on the MFMA
// contraction (ams_cgemm_frag), the elementwise steps and their backward here:
//   gP_R = gA,  gQ_r = gP_r o conj(u),  gP_{r-1} = gA + gQ_r . conj(T)^T,
//   gu[m][k] = sum_r sum_c gP_r[r][k] conj(Q_r[r][k])  (Q_1 = WG),
//   gWG[c][k] = sum_m gP_1[r][k] conj(u[m][k]).

#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "../../include/ams.h"
#include "ams_frag.h"
#include "ams_internal.h"

namespace {

constexpr int kPmBlock = 256;
constexpr int kPmMaxM = 4096;      // AP elements (LDS row of W: 32 KB)
constexpr int kPmFragMaxM = 128;   // AP elements of the fragment writer (W rows and G columns: 48 KB)
constexpr int kPmApBlock = 256;
constexpr int kPmMsK = 32;  // metasurface elements per workgroup of k_pm_bwd_ms
constexpr int kPmMsC = 8;   // code groups per workgroup of k_pm_bwd_ms

// a * b and acc + a * b with every rounding fixed (explicit fma): the
// operand kernels below form A the same way, value for value.
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cmac(float2 acc, float2 a, float2 b) {
  acc.x = fmaf(a.x, b.x, acc.x);
  acc.x = fmaf(-a.y, b.y, acc.x);
  acc.y = fmaf(a.x, b.y, acc.y);
  acc.y = fmaf(a.y, b.x, acc.y);
  return acc;
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// torch.remainder for floats: the sign of the divisor
__device__ __forceinline__ float remainder_like_torch(float a, float b) {
  float m = fmodf(a, b);
  if (m != 0.0f && ((b < 0.0f) != (m < 0.0f))) m += b;
  return m;
}

__device__ __forceinline__ float2 ap_weight(float amp, float th) {
  float s, c;
  sincosf(th, &s, &c);
  return make_float2(amp * c, -amp * s);
}

__device__ __forceinline__ float2 ms_coef(float a, float th, float max_phase) {
  float s, c;
  sincosf(remainder_like_torch(th, max_phase), &s, &c);
  return make_float2(a * c, a * s);
}

// grid (ceil(K / 256), Cn): a workgroup holds W[c][:] in LDS and forms
// WG[c][k] for 256 elements k, then A for every map.  Workgroups of column
// block 0 also write the W columns of A.
// period (may be null): the max phase of each element k, else max_phase.
__global__ void __launch_bounds__(kPmBlock) k_pm_operand(const float *__restrict__ th_ap,
                                                         const float *__restrict__ amp,
                                                         const float *__restrict__ th_ms,
                                                         const float *__restrict__ ms_amp,
                                                         const float2 *__restrict__ G, int Cn, int M, int K, int nm,
                                                         float max_phase, const float *__restrict__ period,
                                                         float2 *__restrict__ A,
                                                         float2 *__restrict__ WG, float2 *__restrict__ U) {
  extern __shared__ float2 w_row[];
  const int c = blockIdx.y;
  const long long Kc = (long long)K + M;
  for (int j = threadIdx.x; j < M; j += blockDim.x) {
    float2 w = ap_weight(amp[(size_t)c * M + j], th_ap[(size_t)c * M + j]);
    w_row[j] = w;
    if (blockIdx.x == 0)
      for (int m = 0; m < nm; m++) A[(size_t)(m * Cn + c) * Kc + K + j] = w;
  }
  __syncthreads();
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  float2 acc = make_float2(0.0f, 0.0f);
  for (int j = 0; j < M; j++) acc = cmac(acc, w_row[j], G[(size_t)j * K + k]);
  WG[(size_t)c * K + k] = acc;
  const float a = ms_amp[k];
  const float pk = period ? period[k] : max_phase;
  for (int m = 0; m < nm; m++) {
    float2 u = ms_coef(a, th_ms[(size_t)m * K + k], pk);
    if (c == 0) U[(size_t)m * K + k] = u;
    A[(size_t)(m * Cn + c) * Kc + k] = cmul(acc, u);
  }
}

// The same operand written straight into the split-bf16 fragment layout of
// the contraction (ams_frag.h) instead of a complex matrix.  One workgroup
// per 16 x 32 fragment block: W of its 16 rows in LDS, two elements per lane
// (the same products in the same order as k_pm_operand, so the fragments
// equal ams_frag_pack of that kernel's A), the block's values gathered in
// LDS and packed by one wave.  The block's 32 columns of G are staged in LDS
// first, so the M-term products read no global memory.  Lanes of map 0 also
// write WG, lanes of code 0 write u.  Dynamic LDS: 48 * M float2.
__global__ void __launch_bounds__(kPmBlock) k_pm_operand_frag(const float *__restrict__ th_ap,
                                                              const float *__restrict__ amp,
                                                              const float *__restrict__ th_ms,
                                                              const float *__restrict__ ms_amp,
                                                              const float2 *__restrict__ G, int Cn, int M, int K,
                                                              int nm, float max_phase, int n_kb,
                                                              u32x4 *__restrict__ frag, float2 *__restrict__ WG,
                                                              float2 *__restrict__ U) {
  extern __shared__ float2 pm_lds[];
  float2 *w_rows = pm_lds;            // [16][M]
  float2 *g_cols = pm_lds + 16 * M;  // [M][32]
  __shared__ float2 tile[16][32];
  const int R = nm * Cn, Kc = K + M;
  const int kb = blockIdx.x % n_kb, rb = blockIdx.x / n_kb;
  for (int idx = threadIdx.x; idx < 32 * M; idx += blockDim.x) {
    const int j = idx >> 5, k = kb * 32 + (idx & 31);
    g_cols[idx] = k < K ? G[(size_t)j * K + k] : make_float2(0.0f, 0.0f);
  }
  for (int idx = threadIdx.x; idx < 16 * M; idx += blockDim.x) {
    const int rl = idx / M, j = idx % M;
    const int r = rb * 16 + rl;
    float2 w = make_float2(0.0f, 0.0f);
    if (r < R) {
      const int c = r % Cn;
      w = ap_weight(amp[(size_t)c * M + j], th_ap[(size_t)c * M + j]);
    }
    w_rows[idx] = w;
  }
  __syncthreads();
  const int rl = threadIdx.x / 16, kl0 = 2 * (threadIdx.x % 16);
  const int r = rb * 16 + rl;
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int kl = kl0 + h, k = kb * 32 + kl;
    float2 v = make_float2(0.0f, 0.0f);
    if (r < R) {
      const int c = r % Cn, m = r / Cn;
      const float2 *wr = w_rows + rl * M;
      if (k < K) {
        float2 acc = make_float2(0.0f, 0.0f);
        for (int j = 0; j < M; j++) acc = cmac(acc, wr[j], g_cols[j * 32 + kl]);
        const float2 u = ms_coef(ms_amp[k], th_ms[(size_t)m * K + k], max_phase);
        v = cmul(acc, u);
        if (m == 0) WG[(size_t)c * K + k] = acc;
        if (c == 0) U[(size_t)m * K + k] = u;
      } else if (k < Kc) {
        v = wr[k - K];
      }
    }
    tile[rl][kl] = v;
  }
  __syncthreads();
  if (threadIdx.x >= 64) return;
  const ams::FragItem it = ams::frag_item((long long)blockIdx.x * 64 + threadIdx.x, n_kb);
  float2 v8[8];
#pragma unroll
  for (int e = 0; e < 8; e++) v8[e] = tile[threadIdx.x & 15][8 * (threadIdx.x >> 4) + e];
  ams::frag_store8(frag, it, v8);
}

// Workgroup = kPmMsK metasurface elements k x kPmMsC code groups, group q
// summing codes c = q, q + kPmMsC, ...; the group sums are added in group
// order.  g_ms for every map, then gWG when the AP phases need a gradient.
__global__ void __launch_bounds__(kPmMsK *kPmMsC) k_pm_bwd_ms(const float2 *__restrict__ gA,
                                                              const float2 *__restrict__ WG,
                                                              const float2 *__restrict__ U, int Cn, int M, int K,
                                                              int nm, float *__restrict__ g_ms,
                                                              float2 *__restrict__ gWG) {
  __shared__ float2 red[kPmMsC][kPmMsK];
  const int kl = threadIdx.x % kPmMsK, cg = threadIdx.x / kPmMsK;
  const int k = blockIdx.x * kPmMsK + kl;
  const bool ok = k < K;
  const long long Kc = (long long)K + M;
  for (int m = 0; m < nm; m++) {
    float2 gu = make_float2(0.0f, 0.0f);
    if (ok)
      for (int c = cg; c < Cn; c += kPmMsC) {
        float2 t = cmulc(gA[(size_t)(m * Cn + c) * Kc + k], WG[(size_t)c * K + k]);
        gu.x += t.x;
        gu.y += t.y;
      }
    red[cg][kl] = gu;
    __syncthreads();
    if (cg == 0 && ok) {
      float2 s = red[0][kl];
      for (int q = 1; q < kPmMsC; q++) {
        s.x += red[q][kl].x;
        s.y += red[q][kl].y;
      }
      float2 u = U[(size_t)m * K + k];
      g_ms[(size_t)m * K + k] = u.x * s.y - u.y * s.x;
    }
    __syncthreads();
  }
  if (!gWG || !ok) return;
  for (int c = cg; c < Cn; c += kPmMsC) {
    float2 s = make_float2(0.0f, 0.0f);
    for (int m = 0; m < nm; m++) {
      float2 t = cmulc(gA[(size_t)(m * Cn + c) * Kc + k], U[(size_t)m * K + k]);
      s.x += t.x;
      s.y += t.y;
    }
    gWG[(size_t)c * K + k] = s;
  }
}

// One workgroup of 256 lanes per (AP element j, code c): gW[c][j] =
// sum_k gWG[c][k] conj(G[j][k]) (lanes stride over k; a wave reduction, then
// the four waves added in order), plus the direct columns, then g_ap[c][j].
__global__ void __launch_bounds__(kPmApBlock) k_pm_bwd_ap(const float2 *__restrict__ gA,
                                                          const float2 *__restrict__ gWG,
                                                          const float2 *__restrict__ G,
                                                          const float *__restrict__ th_ap,
                                                          const float *__restrict__ amp, int Cn, int M, int K,
                                                          int nm, float *__restrict__ g_ap) {
  __shared__ float2 red[kPmApBlock / 64];
  const int j = blockIdx.x, c = blockIdx.y;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long long Kc = (long long)K + M;
  const float2 *gw = gWG + (size_t)c * K;
  const float2 *gj = G + (size_t)j * K;
  float x = 0.0f, y = 0.0f;
#pragma unroll 4
  for (int k = threadIdx.x; k < K; k += kPmApBlock) {
    float2 t = cmulc(gw[k], gj[k]);
    x += t.x;
    y += t.y;
  }
  for (int off = 32; off > 0; off >>= 1) {
    x += __shfl_down(x, off);
    y += __shfl_down(y, off);
  }
  if (lane == 0) red[wave] = make_float2(x, y);
  __syncthreads();
  if (threadIdx.x != 0) return;
  float2 s = red[0];
  for (int w = 1; w < kPmApBlock / 64; w++) {
    s.x += red[w].x;
    s.y += red[w].y;
  }
  for (int m = 0; m < nm; m++) {
    float2 d = gA[(size_t)(m * Cn + c) * Kc + K + j];
    s.x += d.x;
    s.y += d.y;
  }
  float2 w = ap_weight(amp[(size_t)c * M + j], th_ap[(size_t)c * M + j]);
  g_ap[(size_t)c * M + j] = -(w.x * s.y - w.y * s.x);
}

// One chain round: P[r][k] = Q[r][k] u[m][k] (r = m Cn + c), A[r][k] += P.
// P may be null (the last round).  Thread per element.
__global__ void __launch_bounds__(kPmBlock) k_chain_step(const float2 *__restrict__ Q, const float2 *__restrict__ U,
                                                         int Cn, int nm, int K, float2 *__restrict__ A,
                                                         long long lda, float2 *__restrict__ P) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)nm * Cn * K;
  if (t >= n) return;
  const long long r = t / K;
  const int k = (int)(t - r * K);
  const int m = (int)(r / Cn);
  const float2 p = cmul(Q[t], U[(size_t)m * K + k]);
  float2 *a = A + r * lda + k;
  *a = make_float2(a->x + p.x, a->y + p.y);
  if (P) P[t] = p;
}

// gP[r][k] = gA[r][k] (+ Gacc[r][k]): the gradient reaching chain round r's
// output, gA at row stride lda, Gacc (the later rounds' part) dense.
__device__ __forceinline__ float2 chain_gp(const float2 *__restrict__ gA, long long lda,
                                           const float2 *__restrict__ Gacc, long long r, int K, int k) {
  float2 g = gA[r * lda + k];
  if (Gacc) {
    const float2 h = Gacc[r * K + k];
    g.x += h.x;
    g.y += h.y;
  }
  return g;
}

// gQ[r][k] = gP[r][k] conj(u[m][k]), dense [R][K].
__global__ void __launch_bounds__(kPmBlock) k_chain_bwd_scale(const float2 *__restrict__ gA, long long lda,
                                                              const float2 *__restrict__ Gacc,
                                                              const float2 *__restrict__ U, int Cn, int nm, int K,
                                                              float2 *__restrict__ gQ) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)nm * Cn * K;
  if (t >= n) return;
  const long long r = t / K;
  const int k = (int)(t - r * K);
  const int m = (int)(r / Cn);
  gQ[t] = cmulc(chain_gp(gA, lda, Gacc, r, K, k), U[(size_t)m * K + k]);
}

// Round r's share of the metasurface gradient: gu[m][k] (+)= sum_c gP[r][k]
// conj(X[.][k]), X = Q_r (row r) or, for the first round, WG (row c).  With
// g_ms: g_ms = Im(conj(u) gu) after the update; with gWG: gWG[c][k] = sum_m
// gP[r][k] conj(u[m][k]).  Workgroups as k_pm_bwd_ms (group sums in order).
__global__ void __launch_bounds__(kPmMsK *kPmMsC) k_chain_bwd_ms(const float2 *__restrict__ gA, long long lda,
                                                                 const float2 *__restrict__ Gacc,
                                                                 const float2 *__restrict__ X, int x_by_code,
                                                                 const float2 *__restrict__ U, int Cn, int nm, int K,
                                                                 float2 *__restrict__ gu, int accumulate,
                                                                 float *__restrict__ g_ms,
                                                                 float2 *__restrict__ gWG) {
  __shared__ float2 red[kPmMsC][kPmMsK];
  const int kl = threadIdx.x % kPmMsK, cg = threadIdx.x / kPmMsK;
  const int k = blockIdx.x * kPmMsK + kl;
  const bool ok = k < K;
  for (int m = 0; m < nm; m++) {
    float2 s = make_float2(0.0f, 0.0f);
    if (ok)
      for (int c = cg; c < Cn; c += kPmMsC) {
        const long long r = (long long)m * Cn + c;
        const float2 t = cmulc(chain_gp(gA, lda, Gacc, r, K, k), X[(x_by_code ? (long long)c : r) * K + k]);
        s.x += t.x;
        s.y += t.y;
      }
    red[cg][kl] = s;
    __syncthreads();
    if (cg == 0 && ok) {
      float2 v = red[0][kl];
      for (int q = 1; q < kPmMsC; q++) {
        v.x += red[q][kl].x;
        v.y += red[q][kl].y;
      }
      const size_t i = (size_t)m * K + k;
      if (accumulate) {
        v.x += gu[i].x;
        v.y += gu[i].y;
      }
      gu[i] = v;
      if (g_ms) {
        const float2 u = U[i];
        g_ms[i] = u.x * v.y - u.y * v.x;
      }
    }
    __syncthreads();
  }
  if (!gWG || !ok) return;
  for (int c = cg; c < Cn; c += kPmMsC) {
    float2 s = make_float2(0.0f, 0.0f);
    for (int m = 0; m < nm; m++) {
      const float2 t = cmulc(chain_gp(gA, lda, Gacc, (long long)m * Cn + c, K, k), U[(size_t)m * K + k]);
      s.x += t.x;
      s.y += t.y;
    }
    gWG[(size_t)c * K + k] = s;
  }
}

bool pm_dims_ok(int32_t Cn, int32_t M, int32_t K, int32_t nm) {
  return Cn > 0 && M > 0 && M <= kPmMaxM && K > 0 && nm > 0 && (long long)nm * Cn <= (1LL << 20) &&
         (long long)K <= (1LL << 24);
}

}  // namespace

extern "C" {

int ams_pm_operand(const float *th_ap, const float *amp, const float *th_ms, const float *ms_amp, const float *G,
                   int32_t Cn, int32_t M, int32_t K, int32_t nm, float max_phase, float *A, float *WG, float *U,
                   void *stream) {
  if (!th_ap || !amp || !th_ms || !ms_amp || !G || !A || !WG || !U || !pm_dims_ok(Cn, M, K, nm) ||
      !(max_phase > 0.0f)) {
    ams::set_error("ams_pm_operand: bad arguments");
    return AMS_ERR_ARG;
  }
  dim3 grid((unsigned)((K + kPmBlock - 1) / kPmBlock), (unsigned)Cn);
  k_pm_operand<<<grid, kPmBlock, (size_t)M * sizeof(float2), (hipStream_t)stream>>>(
      th_ap, amp, th_ms, ms_amp, (const float2 *)G, Cn, M, K, nm, max_phase, nullptr, (float2 *)A, (float2 *)WG,
      (float2 *)U);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ams::set_error(std::string("ams_pm_operand: ") + hipGetErrorString(e));
    return AMS_ERR_HIP;
  }
  return AMS_OK;
}

int ams_pm_operand_frag(const float *th_ap, const float *amp, const float *th_ms, const float *ms_amp,
                        const float *G, int32_t Cn, int32_t M, int32_t K, int32_t nm, float max_phase, void *frag,
                        float *WG, float *U, void *stream) {
  if (!th_ap || !amp || !th_ms || !ms_amp || !G || !frag || !WG || !U || !pm_dims_ok(Cn, M, K, nm) ||
      M > kPmFragMaxM || !(max_phase > 0.0f)) {
    ams::set_error("ams_pm_operand_frag: bad arguments (M <= 128)");
    return AMS_ERR_ARG;
  }
  const long long R = (long long)nm * Cn, Kc = (long long)K + M;
  const int n_kb = (int)((Kc + 31) / 32);
  const long long n_blocks = ((R + 15) / 16) * n_kb;
  if (n_blocks > 0x7fffffffLL) {
    ams::set_error("ams_pm_operand_frag: operand too large for one launch");
    return AMS_ERR_ARG;
  }
  k_pm_operand_frag<<<(unsigned)n_blocks, kPmBlock, (size_t)48 * M * sizeof(float2), (hipStream_t)stream>>>(
      th_ap, amp, th_ms, ms_amp, (const float2 *)G, Cn, M, K, nm, max_phase, n_kb, (u32x4 *)frag, (float2 *)WG,
      (float2 *)U);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ams::set_error(std::string("ams_pm_operand_frag: ") + hipGetErrorString(e));
    return AMS_ERR_HIP;
  }
  return AMS_OK;
}

int ams_pm_operand_bwd(const float *gA, const float *th_ap, const float *amp, const float *G, const float *WG,
                       const float *U, int32_t Cn, int32_t M, int32_t K, int32_t nm, float *g_ms, float *g_ap,
                       float *gWG, void *stream) {
  if (!gA || !WG || !U || !g_ms || !pm_dims_ok(Cn, M, K, nm) || (g_ap && (!th_ap || !amp || !G || !gWG))) {
    ams::set_error("ams_pm_operand_bwd: bad arguments");
    return AMS_ERR_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  k_pm_bwd_ms<<<(unsigned)((K + kPmMsK - 1) / kPmMsK), kPmMsK * kPmMsC, 0, st>>>(
      (const float2 *)gA, (const float2 *)WG, (const float2 *)U, Cn, M, K, nm, g_ms, g_ap ? (float2 *)gWG : nullptr);
  if (g_ap)
    k_pm_bwd_ap<<<dim3((unsigned)M, (unsigned)Cn), kPmApBlock, 0, st>>>(
        (const float2 *)gA, (const float2 *)gWG, (const float2 *)G, th_ap, amp, Cn, M, K, nm, g_ap);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ams::set_error(std::string("ams_pm_operand_bwd: ") + hipGetErrorString(e));
    return AMS_ERR_HIP;
  }
  return AMS_OK;
}

int ams_chain_first(const float *th_ap, const float *amp, const float *th_ms, const float *ms_amp,
                    const float *period, const float *G, int32_t Cn, int32_t M, int32_t K, int32_t nm, float *A,
                    float *WG, float *U, void *stream) {
  if (!th_ap || !amp || !th_ms || !ms_amp || !period || !G || !A || !WG || !U || !pm_dims_ok(Cn, M, K, nm)) {
    ams::set_error("ams_chain_first: bad arguments");
    return AMS_ERR_ARG;
  }
  dim3 grid((unsigned)((K + kPmBlock - 1) / kPmBlock), (unsigned)Cn);
  k_pm_operand<<<grid, kPmBlock, (size_t)M * sizeof(float2), (hipStream_t)stream>>>(
      th_ap, amp, th_ms, ms_amp, (const float2 *)G, Cn, M, K, nm, 1.0f, period, (float2 *)A, (float2 *)WG,
      (float2 *)U);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ams::set_error(std::string("ams_chain_first: ") + hipGetErrorString(e));
    return AMS_ERR_HIP;
  }
  return AMS_OK;
}

int ams_chain_step(const float *Q, const float *U, int32_t Cn, int32_t nm, int32_t K, float *A, int64_t lda,
                   float *P, void *stream) {
  if (!Q || !U || !A || Cn <= 0 || nm <= 0 || K <= 0 || lda < K || (long long)nm * Cn > (1LL << 20) ||
      (long long)K > (1LL << 24)) {
    ams::set_error("ams_chain_step: bad arguments");
    return AMS_ERR_ARG;
  }
  const long long n = (long long)nm * Cn * K;
  if ((n + kPmBlock - 1) / kPmBlock > 0x7fffffffLL) {
    ams::set_error("ams_chain_step: operand too large for one launch");
    return AMS_ERR_ARG;
  }
  k_chain_step<<<(unsigned)((n + kPmBlock - 1) / kPmBlock), kPmBlock, 0, (hipStream_t)stream>>>(
      (const float2 *)Q, (const float2 *)U, Cn, nm, K, (float2 *)A, lda, (float2 *)P);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ams::set_error(std::string("ams_chain_step: ") + hipGetErrorString(e));
    return AMS_ERR_HIP;
  }
  return AMS_OK;
}

int ams_chain_bwd_scale(const float *gA, int64_t lda, const float *Gacc, const float *U, int32_t Cn, int32_t nm,
                        int32_t K, float *gQ, void *stream) {
  if (!gA || !U || !gQ || Cn <= 0 || nm <= 0 || K <= 0 || lda < K || (long long)nm * Cn > (1LL << 20) ||
      (long long)K > (1LL << 24)) {
    ams::set_error("ams_chain_bwd_scale: bad arguments");
    return AMS_ERR_ARG;
  }
  const long long n = (long long)nm * Cn * K;
  if ((n + kPmBlock - 1) / kPmBlock > 0x7fffffffLL) {
    ams::set_error("ams_chain_bwd_scale: operand too large for one launch");
    return AMS_ERR_ARG;
  }
  k_chain_bwd_scale<<<(unsigned)((n + kPmBlock - 1) / kPmBlock), kPmBlock, 0, (hipStream_t)stream>>>(
      (const float2 *)gA, lda, (const float2 *)Gacc, (const float2 *)U, Cn, nm, K, (float2 *)gQ);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ams::set_error(std::string("ams_chain_bwd_scale: ") + hipGetErrorString(e));
    return AMS_ERR_HIP;
  }
  return AMS_OK;
}

int ams_chain_bwd_ms(const float *gA, int64_t lda, const float *Gacc, const float *X, int32_t x_by_code,
                     const float *U, int32_t Cn, int32_t nm, int32_t K, float *gu, int32_t accumulate, float *g_ms,
                     float *gWG, void *stream) {
  if (!gA || !X || !U || !gu || Cn <= 0 || nm <= 0 || K <= 0 || lda < K || (long long)nm * Cn > (1LL << 20) ||
      (long long)K > (1LL << 24)) {
    ams::set_error("ams_chain_bwd_ms: bad arguments");
    return AMS_ERR_ARG;
  }
  k_chain_bwd_ms<<<(unsigned)((K + kPmMsK - 1) / kPmMsK), kPmMsK * kPmMsC, 0, (hipStream_t)stream>>>(
      (const float2 *)gA, lda, (const float2 *)Gacc, (const float2 *)X, x_by_code, (const float2 *)U, Cn, nm, K,
      (float2 *)gu, accumulate, g_ms, (float2 *)gWG);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ams::set_error(std::string("ams_chain_bwd_ms: ") + hipGetErrorString(e));
    return AMS_ERR_HIP;
  }
  return AMS_OK;
}

int ams_chain_bwd_ap(const float *gA, const float *gWG, const float *G, const float *th_ap, const float *amp,
                     int32_t Cn, int32_t M, int32_t K, int32_t nm, float *g_ap, void *stream) {
  if (!gA || !gWG || !G || !th_ap || !amp || !g_ap || !pm_dims_ok(Cn, M, K, nm)) {
    ams::set_error("ams_chain_bwd_ap: bad arguments");
    return AMS_ERR_ARG;
  }
  k_pm_bwd_ap<<<dim3((unsigned)M, (unsigned)Cn), kPmApBlock, 0, (hipStream_t)stream>>>(
      (const float2 *)gA, (const float2 *)gWG, (const float2 *)G, th_ap, amp, Cn, M, K, nm, g_ap);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ams::set_error(std::string("ams_chain_bwd_ap: ") + hipGetErrorString(e));
    return AMS_ERR_HIP;
  }
  return AMS_OK;
}

}  // extern "C"
