// ams_contract.hip -- phase-map contraction on gfx950 MFMA.
//
// The reference forms H = H_d + sum G diag(u) F and y = W H with torch
// complex matmuls (go.py:107-164).  Here both products become ONE complex
// GEMM  Y = [ (W G) diag(u) | W ] . [ F ; H_d ]  whose large operand is
// stored once in a split-precision bf16 layout (x = hi + lo, each bf16) in
// the lane order of v_mfma_f32_16x16x32_bf16; each real product is
// hi*hi + hi*lo + lo*hi with FP32 accumulation (relative error ~1e-5).
//
// Fragment layout of a complex matrix X[R][Kd]: ams_frag.h.
// The same layout serves the A operand (rows = output rows) and the B
// operand (rows = output columns), so C = A . B^T with both operands
// contiguous along the reduction index.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/ams.h"
#include "ams_frag.h"
#include "ams_internal.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Split partials of ams_cgemm_frag_noalloc: a file-scope device array, so
// that call allocates nothing.  Each GPU holds its own copy until the process
// ends; kernels get its device address (scratch_dev), never the symbol.
constexpr size_t kScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// One thread per (rb, kb, lane): 8 complex values -> 4 planes of 8 bf16.
__global__ void __launch_bounds__(256) k_frag_pack(const float *__restrict__ X, int R, int Kd, long long sr,
                                                   long long sk, int n_kb, long long n_items,
                                                   u32x4 *__restrict__ frag) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_items) return;
  const ams::FragItem it = ams::frag_item(t, n_kb);
  float2 v[8];
#pragma unroll
  for (int e = 0; e < 8; e++) {
    v[e] = make_float2(0.0f, 0.0f);
    int k = it.k0 + e;
    if (it.r < R && k < Kd) {
      const float *src = X + 2 * (it.r * sr + (long long)k * sk);
      v[e] = make_float2(src[0], src[1]);
    }
  }
  ams::frag_store8(frag, it, v);
}

__device__ __forceinline__ bf16x8 as_bf(u32x4 v) { return __builtin_bit_cast(bf16x8, v); }
__device__ __forceinline__ bf16x8 neg_bf(u32x4 v) {
  u32x4 s = v ^ (u32x4){0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
  return __builtin_bit_cast(bf16x8, s);
}
#define MFMA(a, b, c) __builtin_amdgcn_mfma_f32_16x16x32_bf16((a), (b), (c), 0, 0, 0)

// One 32-deep reduction block of a wave's MT x NT tiles: the A and B
// fragments of block kb (rows and columns past the matrix are clamped to the
// last block; they are never stored).
template <int MT, int NT>
__device__ __forceinline__ void load_blocks(const u32x4 *__restrict__ A, const u32x4 *__restrict__ B, int mg,
                                            int ng, int ra_blocks, int rb_blocks, int n_kb, int kb, int lane,
                                            u32x4 (&a)[MT][4], u32x4 (&b)[NT][4]) {
#pragma unroll
  for (int i = 0; i < MT; i++) {
    int rb = mg * MT + i;
    rb = rb < ra_blocks ? rb : ra_blocks - 1;
    const u32x4 *p = A + ((long long)rb * n_kb + kb) * 256 + lane;
#pragma unroll
    for (int q = 0; q < 4; q++) a[i][q] = p[q * 64];
  }
#pragma unroll
  for (int j = 0; j < NT; j++) {
    int cb = ng * NT + j;
    cb = cb < rb_blocks ? cb : rb_blocks - 1;
    const u32x4 *p = B + ((long long)cb * n_kb + kb) * 256 + lane;
#pragma unroll
    for (int q = 0; q < 4; q++) b[j][q] = p[q * 64];
  }
}

// The 12 MFMAs per tile pair of one reduction block: 4 real products x
// (hi.hi, hi.lo, lo.hi), negation as a sign-bit XOR.
template <int MT, int NT, bool CONJB>
__device__ __forceinline__ void mac_blocks(const u32x4 (&a)[MT][4], const u32x4 (&b)[NT][4],
                                           f32x4 (&acc_re)[MT][NT], f32x4 (&acc_im)[MT][NT]) {
#pragma unroll
  for (int i = 0; i < MT; i++) {
    bf16x8 arh = as_bf(a[i][0]), arl = as_bf(a[i][1]), aih = as_bf(a[i][2]), ail = as_bf(a[i][3]);
    bf16x8 narh = neg_bf(a[i][0]), narl = neg_bf(a[i][1]);
    bf16x8 naih = neg_bf(a[i][2]), nail = neg_bf(a[i][3]);
#pragma unroll
    for (int j = 0; j < NT; j++) {
      bf16x8 brh = as_bf(b[j][0]), brl = as_bf(b[j][1]), bih = as_bf(b[j][2]), bil = as_bf(b[j][3]);
      f32x4 re = acc_re[i][j], im = acc_im[i][j];
      // real part: ar*br -/+ ai*bi
      re = MFMA(arh, brh, re);
      re = MFMA(arh, brl, re);
      re = MFMA(arl, brh, re);
      if (CONJB) {
        re = MFMA(aih, bih, re);
        re = MFMA(aih, bil, re);
        re = MFMA(ail, bih, re);
      } else {
        re = MFMA(naih, bih, re);
        re = MFMA(naih, bil, re);
        re = MFMA(nail, bih, re);
      }
      // imaginary part: ai*br +/- ar*bi
      im = MFMA(aih, brh, im);
      im = MFMA(aih, brl, im);
      im = MFMA(ail, brh, im);
      if (CONJB) {
        im = MFMA(narh, bih, im);
        im = MFMA(narh, bil, im);
        im = MFMA(narl, bih, im);
      } else {
        im = MFMA(arh, bih, im);
        im = MFMA(arh, bil, im);
        im = MFMA(arl, bih, im);
      }
      acc_re[i][j] = re;
      acc_im[i][j] = im;
    }
  }
}

// C[Ra][Rb] (complex f32) = A . B^T  (CONJB: A . conj(B)^T) over reduction
// blocks [kb0, kb1).  One wave owns MT x NT tiles of 16 x 16; waves of a
// workgroup share the column group (so B blocks are reused through L1/L2).
// The fragments of the next block are loaded while the current one is
// multiplied (two register sets, the loop unrolled by two).  With splits > 1
// (gridDim.z), partial sums go to part[z][Ra][Rb].
template <int MT, int NT, bool CONJB>
__global__ void __launch_bounds__(256) k_cgemm(const u32x4 *__restrict__ A, const u32x4 *__restrict__ B, int Ra,
                                               int Rb, int n_kb, int m_groups, int n_groups,
                                               float *__restrict__ C, long long ldc, float *__restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int gw = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int mg = gw % m_groups, ng = gw / m_groups;
  if (ng >= n_groups) return;
  const int splits = gridDim.z;
  const int kb0 = (int)((long long)n_kb * blockIdx.z / splits);
  const int kb1 = (int)((long long)n_kb * (blockIdx.z + 1) / splits);
  const int ra_blocks = (Ra + 15) >> 4, rb_blocks = (Rb + 15) >> 4;
  f32x4 acc_re[MT][NT], acc_im[MT][NT];
#pragma unroll
  for (int i = 0; i < MT; i++)
#pragma unroll
    for (int j = 0; j < NT; j++) {
      acc_re[i][j] = (f32x4){0, 0, 0, 0};
      acc_im[i][j] = (f32x4){0, 0, 0, 0};
    }
  u32x4 a0[MT][4], b0[NT][4], a1[MT][4], b1[NT][4];
  int kb = kb0;
  if (kb < kb1) load_blocks<MT, NT>(A, B, mg, ng, ra_blocks, rb_blocks, n_kb, kb, lane, a0, b0);
  while (kb < kb1) {
    if (kb + 1 < kb1) load_blocks<MT, NT>(A, B, mg, ng, ra_blocks, rb_blocks, n_kb, kb + 1, lane, a1, b1);
    mac_blocks<MT, NT, CONJB>(a0, b0, acc_re, acc_im);
    if (++kb >= kb1) break;
    if (kb + 1 < kb1) load_blocks<MT, NT>(A, B, mg, ng, ra_blocks, rb_blocks, n_kb, kb + 1, lane, a0, b0);
    mac_blocks<MT, NT, CONJB>(a1, b1, acc_re, acc_im);
    ++kb;
  }
  // D layout of 16x16 MFMA: lane l holds rows 4*(l>>4)+v, column l&15
  float *out = splits > 1 ? part + (long long)blockIdx.z * 2 * (long long)Ra * Rb : C;
  long long ld = splits > 1 ? Rb : ldc;
#pragma unroll
  for (int i = 0; i < MT; i++) {
#pragma unroll
    for (int j = 0; j < NT; j++) {
      int col = (ng * NT + j) * 16 + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; v++) {
        int row = (mg * MT + i) * 16 + 4 * (lane >> 4) + v;
        if (row < Ra && col < Rb) {
          float2 val = make_float2(acc_re[i][j][v], acc_im[i][j][v]);
          *reinterpret_cast<float2 *>(out + 2 * ((long long)row * ld + col)) = val;
        }
      }
    }
  }
}

// Workgroup-tiled variant for Ra >= 128.  Four waves (2 x 2) own a 128 x 64
// output tile, each wave 64 x 32 (4 x 2 MFMA tiles).  Per 32-deep step the
// workgroup stages its 8 A blocks and 4 B blocks (48 KB) in LDS once, so a
// fragment crosses L2 -> CU once per workgroup instead of once per wave (the
// per-wave kernel needs ~62 B/clk/CU at full MFMA rate, the CU's load limit).
// The next step's blocks are loaded into registers while the current step
// computes.  Workgroups are numbered so that the row groups of one column
// group run on the same XCD (blockIdx.x mod 8) and share B through its L2.
constexpr int kWgA = 8, kWgB = 4, kWgBlocks = kWgA + kWgB;

template <bool CONJB>
__global__ void __launch_bounds__(256) k_cgemm_wg(const u32x4 *__restrict__ A, const u32x4 *__restrict__ B, int Ra,
                                                  int Rb, int n_kb, int row_groups, int col_groups,
                                                  float *__restrict__ C, long long ldc, float *__restrict__ part) {
  __shared__ u32x4 lds[kWgBlocks * 256];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int xcd = (int)(blockIdx.x & 7), s = (int)(blockIdx.x >> 3);
  const int rg = s % row_groups;
  const int cg = (s / row_groups) * 8 + xcd;
  if (cg >= col_groups) return;  // uniform over the workgroup, before any barrier
  const int ra_blocks = (Ra + 15) >> 4, rb_blocks = (Rb + 15) >> 4;
  const int splits = gridDim.z;
  const int kb0 = (int)((long long)n_kb * blockIdx.z / splits);
  const int kb1 = (int)((long long)n_kb * (blockIdx.z + 1) / splits);
  long long base[kWgBlocks];  // first fragment (kb = 0) of each staged block; clamped blocks are never stored
#pragma unroll
  for (int q = 0; q < kWgA; q++) {
    int rb = rg * kWgA + q;
    base[q] = (long long)(rb < ra_blocks ? rb : ra_blocks - 1) * n_kb * 256;
  }
#pragma unroll
  for (int q = 0; q < kWgB; q++) {
    int cb = cg * kWgB + q;
    base[kWgA + q] = (long long)(cb < rb_blocks ? cb : rb_blocks - 1) * n_kb * 256;
  }
  u32x4 pf[kWgBlocks];
  if (kb0 < kb1) {
#pragma unroll
    for (int q = 0; q < kWgBlocks; q++) pf[q] = (q < kWgA ? A : B)[base[q] + (long long)kb0 * 256 + tid];
  }
  f32x4 acc_re[4][2], acc_im[4][2];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) {
      acc_re[i][j] = (f32x4){0, 0, 0, 0};
      acc_im[i][j] = (f32x4){0, 0, 0, 0};
    }
  const int wr = w & 1, wc = w >> 1;
  for (int kb = kb0; kb < kb1; kb++) {
    __syncthreads();  // every wave is done reading the previous step
#pragma unroll
    for (int q = 0; q < kWgBlocks; q++) lds[q * 256 + tid] = pf[q];
    if (kb + 1 < kb1) {
#pragma unroll
      for (int q = 0; q < kWgBlocks; q++) pf[q] = (q < kWgA ? A : B)[base[q] + (long long)(kb + 1) * 256 + tid];
    }
    __syncthreads();
    u32x4 b[2][4];
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int q = 0; q < 4; q++) b[j][q] = lds[(kWgA + wc * 2 + j) * 256 + q * 64 + lane];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      u32x4 a[4];
#pragma unroll
      for (int q = 0; q < 4; q++) a[q] = lds[(wr * 4 + i) * 256 + q * 64 + lane];
      bf16x8 arh = as_bf(a[0]), arl = as_bf(a[1]), aih = as_bf(a[2]), ail = as_bf(a[3]);
      bf16x8 narh = neg_bf(a[0]), narl = neg_bf(a[1]), naih = neg_bf(a[2]), nail = neg_bf(a[3]);
#pragma unroll
      for (int j = 0; j < 2; j++) {
        bf16x8 brh = as_bf(b[j][0]), brl = as_bf(b[j][1]), bih = as_bf(b[j][2]), bil = as_bf(b[j][3]);
        f32x4 re = acc_re[i][j], im = acc_im[i][j];
        re = MFMA(arh, brh, re);
        re = MFMA(arh, brl, re);
        re = MFMA(arl, brh, re);
        if (CONJB) {
          re = MFMA(aih, bih, re);
          re = MFMA(aih, bil, re);
          re = MFMA(ail, bih, re);
        } else {
          re = MFMA(naih, bih, re);
          re = MFMA(naih, bil, re);
          re = MFMA(nail, bih, re);
        }
        im = MFMA(aih, brh, im);
        im = MFMA(aih, brl, im);
        im = MFMA(ail, brh, im);
        if (CONJB) {
          im = MFMA(narh, bih, im);
          im = MFMA(narh, bil, im);
          im = MFMA(narl, bih, im);
        } else {
          im = MFMA(arh, bih, im);
          im = MFMA(arh, bil, im);
          im = MFMA(arl, bih, im);
        }
        acc_re[i][j] = re;
        acc_im[i][j] = im;
      }
    }
  }
  float *out = splits > 1 ? part + (long long)blockIdx.z * 2 * (long long)Ra * Rb : C;
  long long ld = splits > 1 ? Rb : ldc;
#pragma unroll
  for (int i = 0; i < 4; i++) {
#pragma unroll
    for (int j = 0; j < 2; j++) {
      int col = (cg * kWgB + wc * 2 + j) * 16 + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; v++) {
        int row = (rg * kWgA + wr * 4 + i) * 16 + 4 * (lane >> 4) + v;
        if (row < Ra && col < Rb) {
          float2 val = make_float2(acc_re[i][j][v], acc_im[i][j][v]);
          *reinterpret_cast<float2 *>(out + 2 * ((long long)row * ld + col)) = val;
        }
      }
    }
  }
}

// Ordered sum of split partials: C = sum_z part[z]  (deterministic), the
// splits' loads issued four at a time.  k_reduce_parts2 takes two complex
// elements per lane (16-byte loads) when rows have even length; both add
// the splits in the same order, so their sums are identical.
__global__ void k_reduce_parts(const float *__restrict__ part, int splits, int Ra, int Rb, float *__restrict__ C,
                               long long ldc) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long n = (long long)Ra * Rb;
  if (t >= n) return;
  int row = (int)(t / Rb), col = (int)(t % Rb);
  float re = 0.0f, im = 0.0f;
#pragma unroll 4
  for (int z = 0; z < splits; z++) {
    const float *p = part + (long long)z * 2 * n + 2 * t;
    re += p[0];
    im += p[1];
  }
  C[2 * ((long long)row * ldc + col)] = re;
  C[2 * ((long long)row * ldc + col) + 1] = im;
}

__global__ void k_reduce_parts2(const float *__restrict__ part, int splits, int Ra, int Rb, float *__restrict__ C,
                                long long ldc) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;  // pair index
  long long n = (long long)Ra * Rb;
  if (2 * t >= n) return;
  long long e = 2 * t;
  int row = (int)(e / Rb), col = (int)(e % Rb);
  float4 s = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll 4
  for (int z = 0; z < splits; z++) {
    float4 p = *(const float4 *)(part + (long long)z * 2 * n + 2 * e);
    s.x += p.x;
    s.y += p.y;
    s.z += p.z;
    s.w += p.w;
  }
  *(float4 *)(C + 2 * ((long long)row * ldc + col)) = s;
}

// Plain FP32 complex GEMM C = A . B^T on the vector ALU (comparison path).
// A [Ra][Kd] and B [Rb][Kd] complex f32 with row strides lda, ldb.
__global__ void __launch_bounds__(256) k_cgemm_f32(const float2 *__restrict__ A, long long lda,
                                                   const float2 *__restrict__ B, long long ldb, int Ra, int Rb,
                                                   int Kd, int conj_b, float2 *__restrict__ C, long long ldc) {
  __shared__ float2 sa[16][33], sb[64][33];
  int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 columns x 4 row groups
  int col = blockIdx.x * 64 + tx;
  int row0 = blockIdx.y * 16;
  float accr[4] = {0, 0, 0, 0}, acci[4] = {0, 0, 0, 0};
  for (int k0 = 0; k0 < Kd; k0 += 32) {
    for (int e = threadIdx.x; e < 16 * 32; e += 256) {
      int r = e / 32, k = e % 32;
      sa[r][k] = (row0 + r < Ra && k0 + k < Kd) ? A[(long long)(row0 + r) * lda + k0 + k] : make_float2(0, 0);
    }
    for (int e = threadIdx.x; e < 64 * 32; e += 256) {
      int r = e / 32, k = e % 32;
      int cc = blockIdx.x * 64 + r;
      float2 v = (cc < Rb && k0 + k < Kd) ? B[(long long)cc * ldb + k0 + k] : make_float2(0, 0);
      if (conj_b) v.y = -v.y;
      sb[r][k] = v;
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < 32; k++) {
      float2 b = sb[tx][k];
#pragma unroll
      for (int v = 0; v < 4; v++) {
        float2 a = sa[ty * 4 + v][k];
        accr[v] += a.x * b.x - a.y * b.y;
        acci[v] += a.x * b.y + a.y * b.x;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int v = 0; v < 4; v++) {
    int row = row0 + ty * 4 + v;
    if (row < Ra && col < Rb) C[(long long)row * ldc + col] = make_float2(accr[v], acci[v]);
  }
}

template <int MT, int NT, bool CJ>
void launch_cgemm(const u32x4 *A, const u32x4 *B, int Ra, int Rb, int n_kb, int splits, float *C, long long ldc,
                  float *part, hipStream_t st) {
  int m_groups = ((Ra + 15) / 16 + MT - 1) / MT;
  int n_groups = ((Rb + 15) / 16 + NT - 1) / NT;
  long long waves = (long long)m_groups * n_groups;
  unsigned blocks = (unsigned)((waves + 3) / 4);
  k_cgemm<MT, NT, CJ><<<dim3(blocks, 1, splits), 256, 0, st>>>(A, B, Ra, Rb, n_kb, m_groups, n_groups, C, ldc,
                                                               part);
}

template <bool CJ>
void launch_cgemm_wg(const u32x4 *A, const u32x4 *B, int Ra, int Rb, int n_kb, int splits, float *C, long long ldc,
                     float *part, hipStream_t st) {
  int row_groups = ((Ra + 15) / 16 + kWgA - 1) / kWgA;
  int col_groups = ((Rb + 15) / 16 + kWgB - 1) / kWgB;
  unsigned blocks = (unsigned)(8LL * row_groups * ((col_groups + 7) / 8));
  k_cgemm_wg<CJ><<<dim3(blocks, 1, splits), 256, 0, st>>>(A, B, Ra, Rb, n_kb, row_groups, col_groups, C, ldc, part);
}

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_dev[kMaxDevices] = {};

// The current device's address of `scratch`, looked up once per device.
int scratch_dev(void **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    ams::set_error(std::string("ams_cgemm_frag_noalloc: no current device: ") + hipGetErrorString(e));
    return AMS_ERR_HIP;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      ams::set_error(std::string("ams_cgemm_frag_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
      return AMS_ERR_HIP;
    }
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return AMS_OK;
}

}  // namespace

extern "C" {

int64_t ams_frag_bytes(int32_t R, int32_t Kd) {
  if (R <= 0 || Kd <= 0) return 0;
  return (int64_t)((R + 15) / 16) * ((Kd + 31) / 32) * 4096;
}

int ams_frag_pack(const float *X, int32_t R, int32_t Kd, int64_t sr, int64_t sk, void *frag, void *stream) {
  if (!X || !frag || R <= 0 || Kd <= 0) {
    ams::set_error("ams_frag_pack: bad arguments");
    return AMS_ERR_ARG;
  }
  int n_kb = (Kd + 31) / 32;
  long long n_items = (long long)((R + 15) / 16) * n_kb * 64;
  k_frag_pack<<<(unsigned)((n_items + 255) / 256), 256, 0, (hipStream_t)stream>>>(X, R, Kd, sr, sk, n_kb, n_items,
                                                                                  (u32x4 *)frag);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ams::set_error(std::string("ams_frag_pack: ") + hipGetErrorString(e));
    return AMS_ERR_HIP;
  }
  return AMS_OK;
}

int64_t ams_cgemm_workspace_bytes(int32_t Ra, int32_t Rb, int32_t splits) {
  return splits > 1 ? (int64_t)splits * 2 * Ra * (int64_t)Rb * 4 : 0;
}

int ams_cgemm_frag(const void *A, const void *B, int32_t Ra, int32_t Rb, int32_t Kd, int32_t conj_b, float *C,
                   int64_t ldc, int32_t splits, void *workspace, void *stream) {
  if (!A || !B || !C || Ra <= 0 || Rb <= 0 || Kd <= 0 || ldc < Rb || splits < 1 ||
      (splits > 1 && !workspace)) {
    ams::set_error("ams_cgemm_frag: bad arguments");
    return AMS_ERR_ARG;
  }
  int n_kb = (Kd + 31) / 32;
  if (splits > n_kb) splits = n_kb;
  const bool per_wave = (conj_b & 2) != 0;
  conj_b &= 1;
  hipStream_t st = (hipStream_t)stream;
  const u32x4 *a = (const u32x4 *)A, *b = (const u32x4 *)B;
  float *part = (float *)workspace;
  int ra_blocks = (Ra + 15) / 16;
  if (ra_blocks == 1) {
    if (conj_b) launch_cgemm<1, 4, true>(a, b, Ra, Rb, n_kb, splits, C, ldc, part, st);
    else launch_cgemm<1, 4, false>(a, b, Ra, Rb, n_kb, splits, C, ldc, part, st);
  } else if (ra_blocks == 2) {
    if (conj_b) launch_cgemm<2, 4, true>(a, b, Ra, Rb, n_kb, splits, C, ldc, part, st);
    else launch_cgemm<2, 4, false>(a, b, Ra, Rb, n_kb, splits, C, ldc, part, st);
  } else if (ra_blocks < kWgA || per_wave) {
    if (conj_b) launch_cgemm<4, 2, true>(a, b, Ra, Rb, n_kb, splits, C, ldc, part, st);
    else launch_cgemm<4, 2, false>(a, b, Ra, Rb, n_kb, splits, C, ldc, part, st);
  } else {
    if (conj_b) launch_cgemm_wg<true>(a, b, Ra, Rb, n_kb, splits, C, ldc, part, st);
    else launch_cgemm_wg<false>(a, b, Ra, Rb, n_kb, splits, C, ldc, part, st);
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && splits > 1) {
    long long n = (long long)Ra * Rb;
    // pairs stay in one row and 16-byte aligned when Rb and ldc are even
    // (the workspace and C start 16-byte aligned)
    bool pairs = (Rb % 2 == 0) && (ldc % 2 == 0) && ((uintptr_t)C % 16 == 0) && ((uintptr_t)part % 16 == 0);
    if (pairs)
      k_reduce_parts2<<<(unsigned)((n / 2 + 255) / 256), 256, 0, st>>>(part, splits, Ra, Rb, C, ldc);
    else
      k_reduce_parts<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(part, splits, Ra, Rb, C, ldc);
    e = hipGetLastError();
  }
  if (e != hipSuccess) {
    ams::set_error(std::string("ams_cgemm_frag: ") + hipGetErrorString(e));
    return AMS_ERR_HIP;
  }
  return AMS_OK;
}

int64_t ams_cgemm_noalloc_max_bytes(void) { return (int64_t)kScratchBytes; }

int ams_cgemm_frag_noalloc(const void *A, const void *B, int32_t Ra, int32_t Rb, int32_t Kd, int32_t conj_b,
                           float *C, int64_t ldc, int32_t splits, void *stream) {
  if (splits < 1 || Kd <= 0) {
    ams::set_error("ams_cgemm_frag_noalloc: bad arguments");
    return AMS_ERR_ARG;
  }
  const int n_kb = (Kd + 31) / 32;
  if (splits > n_kb) splits = n_kb;  // as ams_cgemm_frag does
  void *ws = nullptr;
  if (splits > 1) {
    const int64_t need = ams_cgemm_workspace_bytes(Ra, Rb, splits);
    if (need > (int64_t)kScratchBytes) {
      ams::set_error("ams_cgemm_frag_noalloc: " + std::to_string(splits) + " split partials of " +
                     std::to_string(Ra) + " x " + std::to_string(Rb) + " need " + std::to_string(need) +
                     " bytes, beyond the " + std::to_string(kScratchBytes) + "-byte scratch");
      return AMS_ERR_ARG;
    }
    int rc = scratch_dev(&ws);
    if (rc) return rc;
  }
  return ams_cgemm_frag(A, B, Ra, Rb, Kd, conj_b, C, ldc, splits, ws, stream);
}

int ams_cgemm_f32(const float *A, int64_t lda, const float *B, int64_t ldb, int32_t Ra, int32_t Rb, int32_t Kd,
                  int32_t conj_b, float *C, int64_t ldc, void *stream) {
  if (!A || !B || !C || Ra <= 0 || Rb <= 0 || Kd <= 0) {
    ams::set_error("ams_cgemm_f32: bad arguments");
    return AMS_ERR_ARG;
  }
  dim3 grid((Rb + 63) / 64, (Ra + 15) / 16);
  k_cgemm_f32<<<grid, 256, 0, (hipStream_t)stream>>>((const float2 *)A, lda, (const float2 *)B, ldb, Ra, Rb, Kd,
                                                     conj_b, (float2 *)C, ldc);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ams::set_error(std::string("ams_cgemm_f32: ") + hipGetErrorString(e));
    return AMS_ERR_HIP;
  }
  return AMS_OK;
}

}  // extern "C"
