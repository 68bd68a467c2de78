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
// With split_n > 0 (a reduction split split_n ways, one-dimensional grid)
// each XCD instead takes whole splits, z = xcd, xcd + 8, ..., with every
// (row group, column group) of a split in turn: the split's slice of A
// (e.g. gY in the backward, read by every column group) then crosses HBM
// once, into that XCD's L2.
constexpr int kWgA = 8, kWgB = 4, kWgBlocks = kWgA + kWgB;

template <bool CONJB, bool SCHED>
__global__ void __launch_bounds__(256) k_cgemm_wg(const u32x4 *__restrict__ A, const u32x4 *__restrict__ B, int Ra,
                                                  int Rb, int n_kb, int row_groups, int col_groups, int split_n,
                                                  float *__restrict__ C, long long ldc, float *__restrict__ part) {
  __shared__ u32x4 lds[kWgBlocks * 256];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int xcd = (int)(blockIdx.x & 7), s = (int)(blockIdx.x >> 3);
  int rg, cg, z, splits;
  if (split_n > 0) {
    const int per = row_groups * col_groups;
    const int zz = s / per, rem = s - zz * per;
    z = zz * 8 + xcd;
    rg = rem % row_groups;
    cg = rem / row_groups;
    splits = split_n;
    if (z >= splits) return;  // uniform over the workgroup, before any barrier
  } else {
    rg = s % row_groups;
    cg = (s / row_groups) * 8 + xcd;
    z = (int)blockIdx.z;
    splits = (int)gridDim.z;
    if (cg >= col_groups) return;  // uniform over the workgroup, before any barrier
  }
  const int ra_blocks = (Ra + 15) >> 4, rb_blocks = (Rb + 15) >> 4;
  const int kb0 = (int)((long long)n_kb * z / splits);
  const int kb1 = (int)((long long)n_kb * (z + 1) / splits);
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
    // B's imaginary planes negated once per step ((-a) b = a (-b) exactly:
    // the per-wave kernel's products, bit for bit)
    bf16x8 nbih[2], nbil[2];
#pragma unroll
    for (int j = 0; j < 2; j++) {
      nbih[j] = neg_bf(b[j][2]);
      nbil[j] = neg_bf(b[j][3]);
    }
    // A fragments in two register sets: tile i + 1's LDS reads are issued
    // before tile i's MFMAs
    u32x4 a_cur[4], a_nxt[4];
#pragma unroll
    for (int q = 0; q < 4; q++) a_cur[q] = lds[(wr * 4) * 256 + q * 64 + lane];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      if (i < 3) {
#pragma unroll
        for (int q = 0; q < 4; q++) a_nxt[q] = lds[(wr * 4 + i + 1) * 256 + q * 64 + lane];
      }
      bf16x8 arh = as_bf(a_cur[0]), arl = as_bf(a_cur[1]), aih = as_bf(a_cur[2]), ail = as_bf(a_cur[3]);
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
          re = MFMA(aih, nbih[j], re);
          re = MFMA(aih, nbil[j], re);
          re = MFMA(ail, nbih[j], re);
        }
        im = MFMA(aih, brh, im);
        im = MFMA(aih, brl, im);
        im = MFMA(ail, brh, im);
        if (CONJB) {
          im = MFMA(arh, nbih[j], im);
          im = MFMA(arh, nbil[j], im);
          im = MFMA(arl, nbih[j], im);
        } else {
          im = MFMA(arh, bih, im);
          im = MFMA(arh, bil, im);
          im = MFMA(arl, bih, im);
        }
        acc_re[i][j] = re;
        acc_im[i][j] = im;
      }
      if (i < 3) {
#pragma unroll
        for (int q = 0; q < 4; q++) a_cur[q] = a_nxt[q];
      }
    }
    // SCHED: have the scheduler issue the step's LDS reads among its first
    // MFMAs instead of before each A tile's (the LDS latency is then hidden)
    if (SCHED) __builtin_amdgcn_iglp_opt(0);
  }
  float *out = splits > 1 ? part + (long long)z * 2 * (long long)Ra * Rb : C;
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

// Three-multiplication workgroup kernel (conj_b bit 3 of ams_cgemm_frag): the
// complex product of a tile pair as three real products,
//   P1 = ar br,  P2 = ai bi,  P3 = (ar + ai)(br +/- bi),
//   re = P1 -/+ P2,  im = P3 - P1 -/+ P2  (lower signs: A . conj(B)^T),
// each split-bf16 (hi.hi + hi.lo + lo.hi): 9 MFMAs per tile pair instead of
// 12.  The sum planes come from the operands' own split values, formed once
// per workgroup while a step is staged: (hi + lo) of the real and imaginary
// part are exact in FP32, their sum is rounded once and split again
// (relative error ~1e-5 per element, as the stored planes).  LDS holds six
// planes per block (72 KB per workgroup); the mapping of workgroups to
// tiles and XCDs is k_cgemm_wg's.
__device__ __forceinline__ float bf_lo(unsigned w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf_hi(unsigned w) { return __uint_as_float(w & 0xffff0000u); }

// the sum planes (hi, lo) of one lane's 8 elements: s = (rh + rl) + sgn (ih + il)
__device__ __forceinline__ void sum_planes(u32x4 rh, u32x4 rl, u32x4 ih, u32x4 il, float sgn, u32x4 &sh,
                                           u32x4 &sl) {
#pragma unroll
  for (int q = 0; q < 4; q++) {
    unsigned h2 = 0, l2 = 0;
#pragma unroll
    for (int half = 0; half < 2; half++) {
      const float r = half ? bf_hi(rh[q]) + bf_hi(rl[q]) : bf_lo(rh[q]) + bf_lo(rl[q]);
      const float i = half ? bf_hi(ih[q]) + bf_hi(il[q]) : bf_lo(ih[q]) + bf_lo(il[q]);
      const float v = fmaf(sgn, i, r);
      const unsigned short h = ams::bf16_rne(v);
      const unsigned short l = ams::bf16_rne(v - ams::bf16_to_f(h));
      h2 |= (unsigned)h << (16 * half);
      l2 |= (unsigned)l << (16 * half);
    }
    sh[q] = h2;
    sl[q] = l2;
  }
}

constexpr int kWg3Planes = 6, kWg3Items = 3;  // (block, lane) items staged per thread and step

template <bool CONJB>
__global__ void __launch_bounds__(256) k_cgemm_wg3(const u32x4 *__restrict__ A, const u32x4 *__restrict__ B,
                                                   int Ra, int Rb, int n_kb, int row_groups, int col_groups,
                                                   int split_n, float *__restrict__ C, long long ldc,
                                                   float *__restrict__ part) {
  __shared__ u32x4 lds[kWgBlocks * kWg3Planes * 64];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int xcd = (int)(blockIdx.x & 7), s = (int)(blockIdx.x >> 3);
  int rg, cg, z, splits;
  if (split_n > 0) {
    const int per = row_groups * col_groups;
    const int zz = s / per, rem = s - zz * per;
    z = zz * 8 + xcd;
    rg = rem % row_groups;
    cg = rem / row_groups;
    splits = split_n;
    if (z >= splits) return;  // uniform over the workgroup, before any barrier
  } else {
    rg = s % row_groups;
    cg = (s / row_groups) * 8 + xcd;
    z = (int)blockIdx.z;
    splits = (int)gridDim.z;
    if (cg >= col_groups) return;  // uniform over the workgroup, before any barrier
  }
  const int ra_blocks = (Ra + 15) >> 4, rb_blocks = (Rb + 15) >> 4;
  const int kb0 = (int)((long long)n_kb * z / splits);
  const int kb1 = (int)((long long)n_kb * (z + 1) / splits);
  // item t = tid + 256 j: block t >> 6 (A blocks 0..7, then B blocks 0..3), lane t & 63
  long long base[kWg3Items];
#pragma unroll
  for (int j = 0; j < kWg3Items; j++) {
    const int blk = (tid + 256 * j) >> 6;
    if (blk < kWgA) {
      const int rb = rg * kWgA + blk;
      base[j] = (long long)(rb < ra_blocks ? rb : ra_blocks - 1) * n_kb * 256 + lane;
    } else {
      const int cb = cg * kWgB + (blk - kWgA);
      base[j] = (long long)(cb < rb_blocks ? cb : rb_blocks - 1) * n_kb * 256 + lane;
    }
  }
  u32x4 pf[kWg3Items][4];
  auto fetch = [&](int kb) {
#pragma unroll
    for (int j = 0; j < kWg3Items; j++) {
      const u32x4 *src = (j < 2 ? A : B) + base[j] + (long long)kb * 256;
#pragma unroll
      for (int q = 0; q < 4; q++) pf[j][q] = src[q * 64];
    }
  };
  if (kb0 < kb1) fetch(kb0);
  f32x4 acc1[4][2], acc2[4][2], acc3[4][2];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) {
      acc1[i][j] = (f32x4){0, 0, 0, 0};
      acc2[i][j] = (f32x4){0, 0, 0, 0};
      acc3[i][j] = (f32x4){0, 0, 0, 0};
    }
  const int wr = w & 1, wc = w >> 1;
  for (int kb = kb0; kb < kb1; kb++) {
    __syncthreads();  // every wave is done reading the previous step
#pragma unroll
    for (int j = 0; j < kWg3Items; j++) {
      const int blk = (tid + 256 * j) >> 6;
      u32x4 *d = lds + blk * kWg3Planes * 64 + lane;
#pragma unroll
      for (int q = 0; q < 4; q++) d[q * 64] = pf[j][q];
      u32x4 sh, sl;
      sum_planes(pf[j][0], pf[j][1], pf[j][2], pf[j][3], (j == 2 && CONJB) ? -1.0f : 1.0f, sh, sl);
      d[4 * 64] = sh;
      d[5 * 64] = sl;
    }
    if (kb + 1 < kb1) fetch(kb + 1);
    __syncthreads();
    u32x4 b[2][kWg3Planes];
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int q = 0; q < kWg3Planes; q++) b[j][q] = lds[((kWgA + wc * 2 + j) * kWg3Planes + q) * 64 + lane];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      u32x4 a[kWg3Planes];
#pragma unroll
      for (int q = 0; q < kWg3Planes; q++) a[q] = lds[((wr * 4 + i) * kWg3Planes + q) * 64 + lane];
#pragma unroll
      for (int j = 0; j < 2; j++) {
        f32x4 p1 = acc1[i][j], p2 = acc2[i][j], p3 = acc3[i][j];
        p1 = MFMA(as_bf(a[0]), as_bf(b[j][0]), p1);
        p1 = MFMA(as_bf(a[0]), as_bf(b[j][1]), p1);
        p1 = MFMA(as_bf(a[1]), as_bf(b[j][0]), p1);
        p2 = MFMA(as_bf(a[2]), as_bf(b[j][2]), p2);
        p2 = MFMA(as_bf(a[2]), as_bf(b[j][3]), p2);
        p2 = MFMA(as_bf(a[3]), as_bf(b[j][2]), p2);
        p3 = MFMA(as_bf(a[4]), as_bf(b[j][4]), p3);
        p3 = MFMA(as_bf(a[4]), as_bf(b[j][5]), p3);
        p3 = MFMA(as_bf(a[5]), as_bf(b[j][4]), p3);
        acc1[i][j] = p1;
        acc2[i][j] = p2;
        acc3[i][j] = p3;
      }
    }
  }
  float *out = splits > 1 ? part + (long long)z * 2 * (long long)Ra * Rb : C;
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
          const float p1 = acc1[i][j][v], p2 = acc2[i][j][v], p3 = acc3[i][j][v];
          float2 val = CONJB ? make_float2(p1 + p2, (p3 - p1) + p2) : make_float2(p1 - p2, (p3 - p1) - p2);
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
                     float *part, hipStream_t st, bool split_major, bool four_mult, bool sched) {
  int row_groups = ((Ra + 15) / 16 + kWgA - 1) / kWgA;
  int col_groups = ((Rb + 15) / 16 + kWgB - 1) / kWgB;
  const bool sm = split_major && splits >= 8;  // each XCD takes whole splits
  dim3 grid = sm ? dim3((unsigned)(8LL * ((splits + 7) / 8) * row_groups * col_groups), 1, 1)
                 : dim3((unsigned)(8LL * row_groups * ((col_groups + 7) / 8)), 1, (unsigned)splits);
  const int sn = sm ? splits : 0;
  if (four_mult && sched)
    k_cgemm_wg<CJ, true><<<grid, 256, 0, st>>>(A, B, Ra, Rb, n_kb, row_groups, col_groups, sn, C, ldc, part);
  else if (four_mult)
    k_cgemm_wg<CJ, false><<<grid, 256, 0, st>>>(A, B, Ra, Rb, n_kb, row_groups, col_groups, sn, C, ldc, part);
  else
    k_cgemm_wg3<CJ><<<grid, 256, 0, st>>>(A, B, Ra, Rb, n_kb, row_groups, col_groups, sn, C, ldc, part);
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
  const bool split_major = (conj_b & 4) == 0;
  const bool four_mult = (conj_b & 8) == 0;
  const bool sched = (conj_b & 16) == 0;
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
    if (conj_b)
      launch_cgemm_wg<true>(a, b, Ra, Rb, n_kb, splits, C, ldc, part, st, split_major, four_mult, sched);
    else
      launch_cgemm_wg<false>(a, b, Ra, Rb, n_kb, splits, C, ldc, part, st, split_major, four_mult, sched);
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
