// ams_frag.h -- the split-bf16 fragment layout of the MFMA contraction,
// shared by the kernels that write it.
//
// Fragment layout of a complex matrix X[R][Kd] (row r, reduction index k):
//   frag[rb][kb][plane][lane][8]  (bf16), rb = r/16, kb = k/32,
//   lane l holds X[rb*16 + (l&15)][kb*32 + 8*(l>>4) + e], e = 0..7,
//   planes: 0 re_hi, 1 re_lo, 2 im_hi, 3 im_lo (x = hi + lo, each bf16,
//   rounded to nearest even).  One (rb, kb) block = 4 KB.  Rows and columns
//   past the matrix hold zeros.
#pragma once
#include <hip/hip_runtime.h>

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

namespace ams {

__device__ __forceinline__ unsigned short bf16_rne(float x) {
  unsigned int u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f(unsigned short h) { return __uint_as_float(((unsigned int)h) << 16); }

// Work item t of a fragment writer: one lane of one (rb, kb) block, i.e. the
// 8 elements (r, k0 .. k0 + 7).
struct FragItem {
  long long blk;  // rb * n_kb + kb
  int lane;
  long long r;
  int k0;
};
__device__ __forceinline__ FragItem frag_item(long long t, int n_kb) {
  FragItem it;
  it.lane = (int)(t & 63);
  it.blk = t >> 6;
  const int kb = (int)(it.blk % n_kb);
  const long long rb = it.blk / n_kb;
  it.r = rb * 16 + (it.lane & 15);
  it.k0 = kb * 32 + 8 * (it.lane >> 4);
  return it;
}

// Store the 8 complex values of one item as its 4 planes of 8 bf16.
__device__ __forceinline__ void frag_store8(u32x4 *__restrict__ frag, const FragItem &it, const float2 (&v)[8]) {
  unsigned short p[4][8];
#pragma unroll
  for (int e = 0; e < 8; e++) {
    unsigned short rh = bf16_rne(v[e].x), ih = bf16_rne(v[e].y);
    p[0][e] = rh;
    p[1][e] = bf16_rne(v[e].x - bf16_to_f(rh));
    p[2][e] = ih;
    p[3][e] = bf16_rne(v[e].y - bf16_to_f(ih));
  }
  u32x4 *dst = frag + it.blk * 256 + it.lane;  // 4 planes x 64 lanes of 16 B
#pragma unroll
  for (int q = 0; q < 4; q++) {
    u32x4 w;
    w.x = p[q][0] | ((unsigned)p[q][1] << 16);
    w.y = p[q][2] | ((unsigned)p[q][3] << 16);
    w.z = p[q][4] | ((unsigned)p[q][5] << 16);
    w.w = p[q][6] | ((unsigned)p[q][7] << 16);
    dst[q * 64] = w;
  }
}

}  // namespace ams
