// ams_common.h -- layouts shared by the host side and the gfx950 kernels of
// the channel tracer.  Everything here is plain data: 32-byte BVH nodes,
// 48-byte triangle records, 16-byte sphere records and the 128-bit path key.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define AMS_MAX_HIT 5          // optixPathTracer.h:5 (MAX_HIT_N)
#define AMS_NO_PLANE 0xFFFFFu  // unused plane slot in a path key
#define AMS_TMIN 1e-5f         // optixPathTracer.cu:98 (tmin)
#define AMS_ACC_MIN 1e-4f      // floor of the Tx/Rx acceptance radius (m)
#define AMS_LEAF_MAX 4
#define AMS_NODE4_EMPTY ((int)0x80000000)

// BVH node, 32 bytes, depth-first preorder.  A hit continues at i + 1 for an
// inner node; a miss (or a finished leaf) continues at `miss`.  `info` is 0
// for an inner node, (first << 3) | count for a leaf (count 1..7).
struct __align__(16) AmsNode {
  float lo[3];
  int32_t miss;
  float hi[3];
  int32_t info;
};

// Four-wide node, 128 bytes, collapsed from the binary hierarchy: the boxes
// of up to four children in structure-of-arrays form and their references
// (>= 0: a four-wide node; < 0: a leaf, -1 - ((first << 3) | count) into the
// leaf-ordered triangles; AMS_NODE4_EMPTY: an unused slot, never entered).
struct __align__(16) AmsNode4 {
  float4 lox, loy, loz, hix, hiy, hiz;
  int4 child;
  int4 pad;
};

// Triangle record, 48 bytes, stored in BVH leaf order.
//   a = (v0.xyz, original triangle index)  b = (v1.xyz, plane id)
//   c = (v2.xyz, material index)
struct __align__(16) AmsTri {
  float4 a, b, c;
};

// Per original triangle: unit normal normalize(cross(v1-v0, v2-v0)) and the
// longest edge, both in FP32 as the reference computes them
// (optixPathTracer.cu:355, :369-370).
struct __align__(16) AmsTriAux {
  float4 n_maxedge;
};

// Plane of a plane id (indexed by the id = the smallest triangle index of the
// plane): canonical unit normal and offset in FP64, used for image points.
struct __align__(16) AmsPlane {
  double n[3];
  double d;
};

// Axis-aligned bounds of all triangles of a plane id (FP64 rounded outward).
struct __align__(16) AmsPlaneBox {
  float4 lo, hi;
};

// Path key, compared as (hi, lo):
//   hi = h:3 | b:5 | p0:20 | p1:20 | p2[19:4]:16
//   lo = p2[3:0]:4 | p3:20 | p4:20 | tile:20
// h = interactions, b = reflect mask (bit k-1 = k-th interaction reflects),
// p0.. = plane ids of the reflections in forward order.  The reference keys
// on the float image transform instead (optixPathTracer.h:9-13, :28-38).
struct AmsKey {
  uint64_t hi, lo;
};

__host__ __device__ inline void ams_key_pack(int h, int b, const uint32_t *pl, uint32_t tile,
                                             uint64_t *hi, uint64_t *lo) {
  *hi = ((uint64_t)h << 61) | ((uint64_t)b << 56) | ((uint64_t)pl[0] << 36) |
        ((uint64_t)pl[1] << 16) | ((uint64_t)pl[2] >> 4);
  *lo = ((uint64_t)(pl[2] & 0xF) << 60) | ((uint64_t)pl[3] << 40) | ((uint64_t)pl[4] << 20) |
        (uint64_t)(tile & 0xFFFFF);
}

__host__ __device__ inline void ams_key_unpack(uint64_t hi, uint64_t lo, int *h, int *b, uint32_t *pl,
                                               uint32_t *tile) {
  *h = (int)(hi >> 61);
  *b = (int)((hi >> 56) & 31);
  pl[0] = (uint32_t)((hi >> 36) & 0xFFFFF);
  pl[1] = (uint32_t)((hi >> 16) & 0xFFFFF);
  pl[2] = (uint32_t)(((hi & 0xFFFF) << 4) | (lo >> 60));
  pl[3] = (uint32_t)((lo >> 40) & 0xFFFFF);
  pl[4] = (uint32_t)((lo >> 20) & 0xFFFFF);
  *tile = (uint32_t)(lo & 0xFFFFF);
}

// optixPathTracer.cu:83-90: Rx index -> tile index
__host__ __device__ inline uint32_t ams_tile_of(int j, int n1, int gs) {
  int n1g = (n1 - 1) / gs + 1;
  return (uint32_t)((j / n1) / gs * n1g + (j % n1) / gs);
}
