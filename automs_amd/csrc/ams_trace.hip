// ams_trace.hip -- gfx950 channel tracer: forward path discovery, GPU dedup
// of (path, tile) records, backward image-method evaluation with register
// accumulation per (Tx, Rx) lane, and the neighbour-expansion rounds.
//
// Reference behaviour restated (file:line in the reference tree):
//   raygen stage 1 / closest-hit forward  optixPathTracer.cu:142-190, :285-331
//   raygen stage 2 / closest-hit backward optixPathTracer.cu:191-273, :332-417
//   sphere intersection                   optixPathTracer.cu:419-444
//   beam patterns                         optixPathTracer.cu:107-136
//   record compaction, dedup, rounds      optixPathTracer.cpp:93-360
//   torch op entry                        optixPathTracer.cpp:781-861
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_set>
#include <vector>

#include "../../include/ams.h"
#include "ams_bvh.h"
#include "ams_common.h"
#include "ams_internal.h"

// ------------------------------------------------------------------------
// device helpers
// ------------------------------------------------------------------------
namespace {

__device__ __forceinline__ float sel3(float3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
__device__ __forceinline__ float3 f3(float x, float y, float z) { return make_float3(x, y, z); }
__device__ __forceinline__ float3 operator+(float3 a, float3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 operator-(float3 a, float3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 operator*(float3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// v_rsq_f32 directly: rsqrtf adds a rescaling for denormal arguments, which
// squared lengths of direction vectors (about 1) never are; for every normal
// argument the result is the same.
__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float3 norm3(float3 a) { return a * rsq(dot3(a, a)); }

// Per-ray state of the watertight triangle test (Woop, Benthin, Wald,
// "Watertight Ray/Triangle Intersection", JCGT 2013).
struct Ray {
  float3 o, d, inv;
  int kx, ky, kz;
  float sx, sy, sz;
};

__device__ __forceinline__ void ray_setup(Ray &r, float3 o, float3 d) {
  r.o = o;
  r.d = d;
  // hardware reciprocals (1 ulp): the slab test's margin covers their error,
  // and the shear below needs only to be the same for every triangle
  r.inv = f3(__builtin_amdgcn_rcpf(d.x), __builtin_amdgcn_rcpf(d.y), __builtin_amdgcn_rcpf(d.z));
  float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  int kx = kz == 2 ? 0 : kz + 1;
  int ky = kx == 2 ? 0 : kx + 1;
  if (sel3(d, kz) < 0.0f) {
    int t = kx;
    kx = ky;
    ky = t;
  }
  r.kx = kx;
  r.ky = ky;
  r.kz = kz;
  const float iz = __builtin_amdgcn_rcpf(sel3(d, kz));
  r.sx = sel3(d, kx) * iz;
  r.sy = sel3(d, ky) * iz;
  r.sz = iz;
}

// Watertight test; returns true and t if the hit lies in [tmin, tmax].
__device__ __forceinline__ bool tri_test(const Ray &r, float4 a, float4 b, float4 c, float tmin, float tmax,
                                         float &t_out) {
  float3 A = f3(a.x, a.y, a.z) - r.o, B = f3(b.x, b.y, b.z) - r.o, C = f3(c.x, c.y, c.z) - r.o;
  float Akz = sel3(A, r.kz), Bkz = sel3(B, r.kz), Ckz = sel3(C, r.kz);
  float Ax = sel3(A, r.kx) - r.sx * Akz, Ay = sel3(A, r.ky) - r.sy * Akz;
  float Bx = sel3(B, r.kx) - r.sx * Bkz, By = sel3(B, r.ky) - r.sy * Bkz;
  float Cx = sel3(C, r.kx) - r.sx * Ckz, Cy = sel3(C, r.ky) - r.sy * Ckz;
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.0f || V == 0.0f || W == 0.0f) {
    U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
    V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
    W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
  }
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  float det = U + V + W;
  if (det == 0.0f) return false;
  float T = U * (r.sz * Akz) + V * (r.sz * Bkz) + W * (r.sz * Ckz);
  if (det < 0.0f) {
    det = -det;
    T = -T;
  }
  if (T < tmin * det || T > tmax * det) return false;
  t_out = T / det;
  return t_out >= tmin && t_out <= tmax;
}

__device__ __forceinline__ bool box_test(const Ray &r, float4 lo, float4 hi, float tmin, float tmax) {
  float tx0 = (lo.x - r.o.x) * r.inv.x, tx1 = (hi.x - r.o.x) * r.inv.x;
  float ty0 = (lo.y - r.o.y) * r.inv.y, ty1 = (hi.y - r.o.y) * r.inv.y;
  float tz0 = (lo.z - r.o.z) * r.inv.z, tz1 = (hi.z - r.o.z) * r.inv.z;
  float t0 = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tmin));
  float t1 = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
  // conservative margin on the far distance (Ize 2013): each distance is off
  // by at most 4 units of the last place (subtraction, the 1-ulp reciprocal,
  // the product), so 2 x 4 units plus slack
  return t0 <= t1 * 1.000001f;
}

struct Counters {  // per-lane counts of one launch (32 bits per lane suffice)
  unsigned nodes = 0, tris = 0, sph = 0, queries = 0;
};

// Closest scene triangle in [tmin, tbest], triangles of plane skip_pl are
// ignored (a straight ray meets a plane once).  Ties on t go to the smaller
// original triangle index.  Returns the leaf-order slot or -1.
__device__ __forceinline__ int closest_tri(const AmsNode *__restrict__ nodes, int n_nodes,
                                           const AmsTri *__restrict__ tris, const Ray &r, float tmin,
                                           float &tbest, int skip_pl, int &orig_out, Counters &ct) {
  int hit = -1, horig = 0x7fffffff;
  int i = 0;
  while (i < n_nodes) {
    const float4 *np = reinterpret_cast<const float4 *>(nodes + i);
    float4 lo = np[0], hi = np[1];
    ct.nodes++;
    int miss = __float_as_int(lo.w), info = __float_as_int(hi.w);
    if (box_test(r, lo, hi, tmin, tbest)) {
      if (info == 0) {
        i++;
        continue;
      }
      int first = info >> 3, cnt = info & 7;
      for (int k = 0; k < cnt; k++) {
        const AmsTri tr = tris[first + k];
        ct.tris++;
        if (__float_as_int(tr.b.w) == skip_pl) continue;
        float t;
        if (tri_test(r, tr.a, tr.b, tr.c, tmin, tbest, t)) {
          int orig = __float_as_int(tr.a.w);
          if (t < tbest || orig < horig) {
            tbest = t;
            hit = first + k;
            horig = orig;
          }
        }
      }
    }
    i = miss;
  }
  orig_out = horig;
  return hit;
}

// Entry distance of the ray into a box within [tmin, tmax] (INFINITY when
// missed); the same conservative test as box_test.
__device__ __forceinline__ float box_entry(const Ray &r, float lx, float ly, float lz, float hx, float hy, float hz,
                                           float tmin, float tmax) {
  float tx0 = (lx - r.o.x) * r.inv.x, tx1 = (hx - r.o.x) * r.inv.x;
  float ty0 = (ly - r.o.y) * r.inv.y, ty1 = (hy - r.o.y) * r.inv.y;
  float tz0 = (lz - r.o.z) * r.inv.z, tz1 = (hz - r.o.z) * r.inv.z;
  float t0 = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tmin));
  float t1 = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
  return t0 <= t1 * 1.000001f ? t0 : INFINITY;
}
__device__ __forceinline__ void entry_swap(float &ea, int &ca, float &eb, int &cb) {
  if (eb < ea) {
    float e = ea;
    ea = eb;
    eb = e;
    int c = ca;
    ca = cb;
    cb = c;
  }
}

// Closest scene triangle through the four-wide hierarchy with a short
// register stack: a step tests the four child boxes of a node and goes on
// at the nearest one hit, the others wait on the stack.  A chain of node
// loads is three to four times shorter than the binary walk's.  Same
// triangle tests and tie rule as closest_tri, so the same hit; on a stack
// overflow (never seen on the scenes here) `overflow` is set and the caller
// repeats the query with closest_tri.
constexpr int kStack4 = 12;
__device__ __forceinline__ int closest_tri4(const AmsNode4 *__restrict__ nodes, const AmsTri *__restrict__ tris,
                                            const Ray &r, float tmin, float &tbest, int skip_pl, int &orig_out,
                                            Counters &ct, bool &overflow) {
  int hit = -1, horig = 0x7fffffff;
  int stk[kStack4];
  int sp = 0;
  int cur = 0;
  overflow = false;
  while (true) {
    int next = AMS_NODE4_EMPTY;
    if (cur >= 0) {
      const float4 *np = reinterpret_cast<const float4 *>(nodes + cur);
      const float4 lx = np[0], ly = np[1], lz = np[2], hx = np[3], hy = np[4], hz = np[5];
      const int4 ch = *reinterpret_cast<const int4 *>(np + 6);
      ct.nodes += 4;  // counted in 32-byte boxes, as the binary walk counts its nodes
      float e0 = ch.x != AMS_NODE4_EMPTY ? box_entry(r, lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, tmin, tbest) : INFINITY;
      float e1 = ch.y != AMS_NODE4_EMPTY ? box_entry(r, lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, tmin, tbest) : INFINITY;
      float e2 = ch.z != AMS_NODE4_EMPTY ? box_entry(r, lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, tmin, tbest) : INFINITY;
      float e3 = ch.w != AMS_NODE4_EMPTY ? box_entry(r, lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, tmin, tbest) : INFINITY;
      int c0 = ch.x, c1 = ch.y, c2 = ch.z, c3 = ch.w;
      entry_swap(e0, c0, e1, c1);
      entry_swap(e2, c2, e3, c3);
      entry_swap(e0, c0, e2, c2);
      entry_swap(e1, c1, e3, c3);
      entry_swap(e1, c1, e2, c2);
      if (e3 < INFINITY) {
        if (sp < kStack4) stk[sp++] = c3;
        else overflow = true;
      }
      if (e2 < INFINITY) {
        if (sp < kStack4) stk[sp++] = c2;
        else overflow = true;
      }
      if (e1 < INFINITY) {
        if (sp < kStack4) stk[sp++] = c1;
        else overflow = true;
      }
      if (e0 < INFINITY) next = c0;
    } else {
      const int code = -1 - cur;
      const int first = code >> 3, cnt = code & 7;
      for (int k = 0; k < cnt; k++) {
        const AmsTri tr = tris[first + k];
        ct.tris++;
        if (__float_as_int(tr.b.w) == skip_pl) continue;
        float t;
        if (tri_test(r, tr.a, tr.b, tr.c, tmin, tbest, t)) {
          int orig = __float_as_int(tr.a.w);
          if (t < tbest || orig < horig) {
            tbest = t;
            hit = first + k;
            horig = orig;
          }
        }
      }
    }
    if (overflow) break;
    if (next != AMS_NODE4_EMPTY) {
      cur = next;
    } else {
      if (sp == 0) break;
      cur = stk[--sp];
    }
  }
  orig_out = horig;
  return hit;
}

// The scene query of the kernels: four-wide walk when its hierarchy is given
// (nodes4 != nullptr), the binary threaded walk otherwise or after a stack
// overflow.
__device__ __forceinline__ int scene_closest(const AmsNode4 *nodes4, const AmsNode *nodes, int n_nodes,
                                             const AmsTri *tris, const Ray &r, float tmin, float &tbest, int skip_pl,
                                             int &orig_out, Counters &ct) {
  if (nodes4) {
    const float t0 = tbest;
    bool overflow;
    int h = closest_tri4(nodes4, tris, r, tmin, tbest, skip_pl, orig_out, ct, overflow);
    if (!overflow) return h;
    tbest = t0;
  }
  return closest_tri(nodes, n_nodes, tris, r, tmin, tbest, skip_pl, orig_out, ct);
}

// One reception sphere (centre, original index in .w): accepted when the
// miss distance is below the acceptance radius, the closest approach b lies
// in [tmin, tbest] and the origin is outside the sphere
// (optixPathTracer.cu:436-444).  Ties on b go to the smaller index, so the
// outcome does not depend on the order candidates are visited in.
__device__ __forceinline__ void sphere_candidate(float4 s, const Ray &r, float tmin, float &tbest, float acc2,
                                                 float r2, int &hit, int &horig) {
  float3 O = f3(s.x, s.y, s.z) - r.o;
  float b = dot3(O, r.d);
  float3 P = O - r.d * b;
  float perp2 = dot3(P, P);
  float c = dot3(O, O);
  int orig = __float_as_int(s.w);
  if (perp2 < acc2 && b >= tmin && b <= tbest && c > r2) {
    if (b < tbest || orig < horig) {
      tbest = b;
      hit = orig;
      horig = orig;
    }
  }
}

// Closest reception sphere whose closest approach b lies in [tmin, tmax],
// with miss distance below the acceptance radius and the origin outside the
// sphere (optixPathTracer.cu:436-444).  Reported at t = b.
__device__ __forceinline__ int closest_sphere(const AmsNode *__restrict__ nodes, int n_nodes,
                                              const float4 *__restrict__ sph, const Ray &r, float tmin,
                                              float &tbest, float acc2, float r2, Counters &ct) {
  int hit = -1, horig = 0x7fffffff;
  int i = 0;
  while (i < n_nodes) {
    const float4 *np = reinterpret_cast<const float4 *>(nodes + i);
    float4 lo = np[0], hi = np[1];
    ct.nodes++;
    int miss = __float_as_int(lo.w), info = __float_as_int(hi.w);
    if (box_test(r, lo, hi, tmin, tbest)) {
      if (info == 0) {
        i++;
        continue;
      }
      int first = info >> 3, cnt = info & 7;
      for (int k = 0; k < cnt; k++) {
        ct.sph++;
        sphere_candidate(sph[first + k], r, tmin, tbest, acc2, r2, hit, horig);
      }
    }
    i = miss;
  }
  return hit;
}

// Rx arrays laid out as a regular lattice c0 + i u + j v (within delta)
// skip the sphere hierarchy.  A sphere that passes sphere_candidate has its
// closest approach q within acc of its centre, hence within acc + delta of
// its lattice point: q lies in the slab |n.(q - c0)| <= W around the lattice
// plane and its lattice coordinates are within (acc + delta) / h of (i, j),
// h being the spacing of the lattice lines.  The cells met that way along
// the clipped ray segment are a superset of the passing spheres, and the
// same per-sphere test with its order-free tie rule gives the same hit.
struct RxLattice {
  int on, n0, n1;
  float3 c0, n, gi, gj;  // i(x) = gi . (x - c0), j(x) = gj . (x - c0)
  float W, Mi, Mj;       // slab half-width (m), cell margins
};

__device__ __forceinline__ void lattice_clip(float x0, float dx, float lo, float hi, float &ta, float &tb) {
  if (fabsf(dx) > 1e-30f) {
    // a 1-ulp reciprocal: its error moves the clip by far less than the
    // lattice's 0.1 mm slack
    const float r = __builtin_amdgcn_rcpf(dx);
    float t1 = (lo - x0) * r, t2 = (hi - x0) * r;
    ta = fmaxf(ta, fminf(t1, t2));
    tb = fminf(tb, fmaxf(t1, t2));
  } else if (x0 < lo || x0 > hi) {
    tb = -INFINITY;
  }
}
__device__ __forceinline__ int lattice_lo(float v, int n) {
  return (int)ceilf(fminf(fmaxf(v, -1.0f), (float)n));
}
__device__ __forceinline__ int lattice_hi(float v, int n) {
  return (int)floorf(fminf(fmaxf(v, -1.0f), (float)n));
}

// The part [ta, tb] of the ray o + t d that lies in the lattice slab within
// the lattice's extent (margins included), and its lattice coordinates
// i(t) = i0 + t di, j(t) = j0 + t dj; false when that part is empty.
__device__ __forceinline__ bool lattice_segment(const RxLattice &L, float3 o, float3 d, float &ta, float &tb,
                                                float &i0, float &di, float &j0, float &dj) {
  const float3 w0 = o - L.c0;
  const float s0 = dot3(L.n, w0), sd = dot3(L.n, d);
  lattice_clip(s0, sd, -L.W, L.W, ta, tb);
  i0 = dot3(L.gi, w0);
  di = dot3(L.gi, d);
  j0 = dot3(L.gj, w0);
  dj = dot3(L.gj, d);
  lattice_clip(i0, di, -L.Mi, (float)(L.n0 - 1) + L.Mi, ta, tb);
  lattice_clip(j0, dj, -L.Mj, (float)(L.n1 - 1) + L.Mj, ta, tb);
  return ta <= tb;
}

__device__ __forceinline__ int closest_sphere_lattice(const RxLattice &L, const float4 *__restrict__ sph_o,
                                                      const Ray &r, float tmin, float &tbest, float acc2, float r2,
                                                      Counters &ct) {
  int hit = -1, horig = 0x7fffffff;
  float ta = tmin, tb = fminf(tbest, 1e30f), i0, di, j0, dj;
  if (!lattice_segment(L, r.o, r.d, ta, tb, i0, di, j0, dj)) return -1;
  // walk the axis along which the segment is longer, one lattice line at a time
  const bool imaj = fabsf(di) >= fabsf(dj);
  const float a0 = imaj ? i0 : j0, da = imaj ? di : dj, b0 = imaj ? j0 : i0, db = imaj ? dj : di;
  const float Ma = imaj ? L.Mi : L.Mj, Mb = imaj ? L.Mj : L.Mi;
  const int na = imaj ? L.n0 : L.n1, nb = imaj ? L.n1 : L.n0;
  const float aA = a0 + ta * da, aB = a0 + tb * da;
  const int alo = max(0, lattice_lo(fminf(aA, aB) - Ma, na)), ahi = min(na - 1, lattice_hi(fmaxf(aA, aB) + Ma, na));
  // lines in the order the ray meets them: once a line's part of the ray
  // starts beyond the best hit, so do all later ones
  const int step = da >= 0.0f ? 1 : -1;
  const int kend = da >= 0.0f ? ahi + 1 : alo - 1;
  for (int ka = da >= 0.0f ? alo : ahi; ka != kend; ka += step) {
    float tl = ta, th = tb;
    lattice_clip(a0, da, (float)ka - Ma, (float)ka + Ma, tl, th);
    if (!(tl <= th)) continue;
    if (tl > tbest) break;
    const float bA = b0 + tl * db, bB = b0 + th * db;
    const int blo = max(0, lattice_lo(fminf(bA, bB) - Mb, nb)), bhi = min(nb - 1, lattice_hi(fmaxf(bA, bB) + Mb, nb));
    for (int kb = blo; kb <= bhi; kb++) {
      const int idx = imaj ? ka * L.n1 + kb : kb * L.n1 + ka;
      ct.sph++;
      sphere_candidate(sph_o[idx], r, tmin, tbest, acc2, r2, hit, horig);
    }
  }
  return hit;
}

__device__ __forceinline__ int theta_bin(float3 d, float3 N) {
  // optixPathTracer.cu:304-306
  float c = dot3(d, N) * rsq(dot3(N, N) * dot3(d, d));
  float th = acosf(fminf(1.0f, fabsf(c)));
  int id = (int)((th / (float)(M_PI / 2) * 36.0f + 1.0f) / 2.0f);
  return id > 18 ? 18 : id;
}

// optixPathTracer.cu:107-136
__device__ __forceinline__ float beam_pattern(float3 a, float3 b, int type) {
  if (type == 0) return 1.0f;
  float c = dot3(a, b) * rsq(dot3(a, a) * dot3(b, b));
  switch (type) {
  case 1:
    return c < 0.0f ? 0.0f : c * 2.4495f;
  case 2: {
    float th = acosf(fminf(1.0f, fmaxf(-1.0f, c)));
    if (th > (float)(M_PI * 2 / 3)) return 0.1f;
    if (th > (float)(M_PI * 5 / 12))
      return powf(10.0f, ((float)(M_PI * 2 / 3) - th) / (float)M_PI * 4.0f * 0.41300368f - 1.0f) * 2.4495f;
    return c * 2.4495f;
  }
  case 3:
    return c < 0.0f ? 0.0f : 1.4142f;
  default:
    return 0.0f;  // rejected on the host before launch
  }
}

__device__ __forceinline__ void key_set_plane(uint64_t &hi, uint64_t &lo, int k, uint32_t p) {
  switch (k) {
  case 0:
    hi = (hi & ~(0xFFFFFull << 36)) | ((uint64_t)p << 36);
    break;
  case 1:
    hi = (hi & ~(0xFFFFFull << 16)) | ((uint64_t)p << 16);
    break;
  case 2:
    hi = (hi & ~0xFFFFull) | ((uint64_t)p >> 4);
    lo = (lo & ~(0xFull << 60)) | ((uint64_t)(p & 0xF) << 60);
    break;
  case 3:
    lo = (lo & ~(0xFFFFFull << 40)) | ((uint64_t)p << 40);
    break;
  default:
    lo = (lo & ~(0xFFFFFull << 20)) | ((uint64_t)p << 20);
    break;
  }
}

__device__ __forceinline__ uint32_t key_get_plane(uint64_t hi, uint64_t lo, int k) {
  switch (k) {
  case 0:
    return (uint32_t)((hi >> 36) & 0xFFFFF);
  case 1:
    return (uint32_t)((hi >> 16) & 0xFFFFF);
  case 2:
    return (uint32_t)(((hi & 0xFFFF) << 4) | (lo >> 60));
  case 3:
    return (uint32_t)((lo >> 40) & 0xFFFFF);
  default:
    return (uint32_t)((lo >> 20) & 0xFFFFF);
  }
}

// Wave-aggregated append slot: one atomic per wave for the active lanes.
__device__ __forceinline__ unsigned long long wave_append(unsigned long long *counter) {
  unsigned long long mask = __ballot(1);
  int lane = __lane_id();
  int leader = __ffsll((long long)mask) - 1;
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(counter, (unsigned long long)__popcll(mask));
  base = __shfl(base, leader);
  return base + (unsigned long long)__popcll(mask & ((1ull << lane) - 1ull));
}

// Per-workgroup filter of discovery records in LDS: neighbouring rays keep
// finding the same (path, tile) keys, and only the first copy a workgroup
// finds needs to reach the global record list (its sort and dedup produce
// the unique set).  A key is dropped only when the same key is already in
// the table and was written out; a slot another lane is filling, or a
// crowded neighbourhood, lets the record through, so every unique key is
// written at least once whatever the interleaving.
constexpr int kDedup = 512;  // entries (10 KB of LDS)
struct DedupTable {
  unsigned long long hi[kDedup], lo[kDedup];
  unsigned state[kDedup];  // 0 empty, 1 being filled, 2 holds a key already written out
};
__device__ __forceinline__ void dedup_init(DedupTable &t) {
  for (int k = threadIdx.x; k < kDedup; k += blockDim.x) t.state[k] = 0u;
  __syncthreads();
}
__device__ __forceinline__ bool dedup_first(DedupTable &t, uint64_t hi, uint64_t lo) {
  uint64_t m = (hi ^ (lo * 0x9E3779B97F4A7C15ull)) * 0xBF58476D1CE4E5B9ull;
  unsigned h = (unsigned)(m >> 40);
  for (int p = 0; p < 8; p++) {
    unsigned slot = (h + p) & (kDedup - 1);
    unsigned st = atomicCAS(&t.state[slot], 0u, 1u);
    if (st == 0u) {
      t.hi[slot] = hi;
      t.lo[slot] = lo;
      __threadfence_block();
      atomicExch(&t.state[slot], 2u);
      return true;
    }
    if (st == 1u) return true;
    if (t.hi[slot] == hi && t.lo[slot] == lo) return false;
  }
  return true;
}

// Launch counters: kCtrSlots slots of kCtrN 64-bit counters, one 128-byte
// line per slot.  A wave adds into the slot of its global wave index, so
// the waves of a launch do not queue their atomics on one line (a launch of
// 10^8 short waves otherwise spends seconds there); the host sums the slots.
constexpr int kCtrN = 16, kCtrSlots = 256;
__device__ __forceinline__ unsigned long long *ctr_slot(unsigned long long *c) {
  const unsigned w = ((blockIdx.y * gridDim.x + blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  return c + kCtrN * (w & (kCtrSlots - 1));
}

// Counters 0..N-1 of a workgroup's waves (at most WAVES) summed in LDS: one
// atomic per counter and workgroup, into the workgroup's slot.  Every
// thread of the workgroup must call it (a barrier).
template <int N, int WAVES>
__device__ __forceinline__ void block_add_counters(unsigned long long *c, const unsigned long long (&v)[N]) {
  __shared__ unsigned long long part[WAVES][N];
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < N; k++) {
    unsigned long long x = v[k];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off);
    if (__lane_id() == 0) part[w][k] = x;
  }
  __syncthreads();
  if (threadIdx.x < N) {
    unsigned long long x = 0;
#pragma unroll
    for (int q = 0; q < WAVES; q++)
      if (q < (int)(blockDim.x >> 6)) x += part[q][threadIdx.x];  // the block's own waves
    if (x) atomicAdd(c + kCtrN * ((blockIdx.y * gridDim.x + blockIdx.x) & (kCtrSlots - 1)) + threadIdx.x, x);
  }
}

__device__ __forceinline__ void wave_add_counters(unsigned long long *dst, const Counters &c) {
  unsigned long long v[4] = {(unsigned long long)c.nodes, (unsigned long long)c.tris, (unsigned long long)c.sph,
                             (unsigned long long)c.queries};
#pragma unroll
  for (int k = 0; k < 4; k++) {
    unsigned long long x = v[k];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off);
    if (__lane_id() == 0 && x) atomicAdd(dst + k, x);
  }
}

// ------------------------------------------------------------------------
// forward discovery
// ------------------------------------------------------------------------
// Fibonacci direction k of n in FP64 with the azimuth range-reduced in
// cycles (the same formula the oracle uses).
__device__ __forceinline__ void fib_dir(long long k, long long n, double &x, double &y, double &z) {
  const double g = 0.38196601125010515179541316563436;  // (3 - sqrt 5) / 2
  z = 1.0 - (2.0 * (double)k + 1.0) / (double)n;
  double rr = sqrt(fmax(0.0, 1.0 - z * z));
  double c = (double)k * g;
  double f = c - floor(c);
  double sn, cs;
  sincos(2.0 * M_PI * f, &sn, &cs);
  x = rr * cs;
  y = rr * sn;
}

// Direction table of the Fibonacci ray set in ray-slot order: slot s holds
// fib_dir(perm[s]) rounded to f32, exactly the direction discovery traces.
__global__ void k_fib_dirs(long long n, const unsigned *__restrict__ perm, float4 *__restrict__ dirs) {
  long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  double x, y, z;
  fib_dir(perm[s], n, x, y, z);
  dirs[s] = make_float4((float)x, (float)y, (float)z, 0.0f);
}

// Hilbert index of (x, y) on a 2^16 grid.
__device__ __forceinline__ unsigned hilbert16(unsigned x, unsigned y) {
  unsigned d = 0;
  for (unsigned s = 1u << 15; s > 0; s >>= 1) {
    unsigned rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
    d += s * s * ((3u * rx) ^ ry);
    if (ry == 0) {
      if (rx == 1) {
        x = s - 1 - x;
        y = s - 1 - y;
      }
      unsigned t = x;
      x = y;
      y = t;
    }
  }
  return d;
}

// Sort key of every direction: Hilbert order of its octahedral map, so that
// a wave's lanes get neighbouring directions (coherent traversal).
__global__ void k_fib_keys(long long n, unsigned *__restrict__ keys, unsigned *__restrict__ idx) {
  long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double x, y, z;
  fib_dir(k, n, x, y, z);
  double l1 = fabs(x) + fabs(y) + fabs(z);
  double u = x / l1, v = y / l1;
  if (z < 0) {
    double uu = (1.0 - fabs(v)) * (u >= 0 ? 1.0 : -1.0);
    double vv = (1.0 - fabs(u)) * (v >= 0 ? 1.0 : -1.0);
    u = uu;
    v = vv;
  }
  unsigned qx = (unsigned)fmin(65535.0, fmax(0.0, (u * 0.5 + 0.5) * 65536.0));
  unsigned qy = (unsigned)fmin(65535.0, fmax(0.0, (v * 0.5 + 0.5) * 65536.0));
  keys[k] = hilbert16(qx, qy);
  idx[k] = (unsigned)k;
}

// Scenes whose threaded BVH and leaf-ordered triangles fit in kSceneLds
// bytes are staged in LDS by every workgroup (the kernels' LDS template
// instances): the traversal is a chain of dependent node loads, and from LDS
// each costs a fraction of an L2 hit.  Larger scenes are read from L2/HBM.
constexpr int kSceneLds = 32 * 1024;
// Discovery launch shapes.  The binary walk: workgroups of kDiscBlock
// threads with kDiscRays rays each (the workgroup's record filter sees 2048
// rays) at 72 registers, 7 waves per SIMD.  The four-wide walk keeps its
// stack in registers: one ray per thread in workgroups of 256.
constexpr int kDiscBlock = 512;
constexpr int kDiscRays = 4;
constexpr int kDiscWideBlock = 256;
constexpr int kDiscWideRays = 1;
// Evaluation: at most kEvalBlock lanes per workgroup (256 measured 7% faster
// than 512 and 16% faster than 128 on the room traces)
constexpr int kEvalBlock = 256;
// records of one round the task export may hold (52 bytes each)
constexpr size_t kMaxTaskInfo = size_t(1) << 26;
// candidate lists one evaluation launch may hold
constexpr size_t kBeamBytes = size_t(1) << 30;
static_assert(sizeof(ams_task_info) == 52, "ams_task_info keeps the 52-byte layout of the reference's record");
static_assert(sizeof(AmsNode) == 2 * sizeof(float4) && sizeof(AmsNode4) == 8 * sizeof(float4) &&
                  sizeof(AmsTri) == 3 * sizeof(float4),
              "stage_scene copies nodes and triangles as float4");

__device__ __forceinline__ void stage_scene(const float4 *__restrict__ sn, int n4, const AmsTri *__restrict__ tris,
                                            int n_tris, float4 *smem) {
  const float4 *st = reinterpret_cast<const float4 *>(tris);
  const int m4 = 3 * n_tris;
  for (int k = threadIdx.x; k < n4; k += blockDim.x) smem[k] = sn[k];
  for (int k = threadIdx.x; k < m4; k += blockDim.x) smem[n4 + k] = st[k];
  __syncthreads();
}

struct DiscoverArgs {
  const AmsNode *tnodes;
  int n_tnodes;
  const AmsNode4 *tnodes4;  // four-wide hierarchy, or nullptr for the binary walk
  int n_tnodes4;
  const AmsTri *tris;
  int n_tris;
  const AmsTriAux *aux;
  const float *mat_prop;  // [n_mat][19]
  const AmsNode *snodes;
  int n_snodes;
  const float4 *sph;
  const float4 *sph_o;  // Rx centres and indices in array order (lattice path)
  RxLattice lat;
  const float *src;
  int n_src;
  int n1_dst, gs;
  float acc2, r2;
  int K;
  long long fib_n;
  int per_element;
  long long ray_begin, ray_end;
  const unsigned *perm;  // slot -> direction (per-element) or ray (distributed) index
  const float4 *dirs;    // slot -> f32 direction of fib_dir(perm[slot])
  int flags;             // ams_trace_opts.flags
  uint64_t *out_hi, *out_lo;
  unsigned long long *out_count;
  unsigned long long cap;
  unsigned long long *counters;  // nodes, tris, spheres, queries, records(raw)
};

// Scene data the discovery kernels read per query: the hierarchy walked
// (four-wide when given) and the triangles, copied into LDS once per
// workgroup by the LDS instances; otherwise the views point at global
// memory (the normals and the material table are always read from there).
struct SceneView {
  const AmsNode *tnodes;
  const AmsNode4 *tnodes4;
  const AmsTri *tris;
};
// WIDE: 1 the four-wide walk (a.tnodes4 set), 0 the binary walk (a.tnodes4
// null), -1 either, by a.tnodes4.  With the walk known at compile time the
// staged nodes are known to lie in LDS (LDS loads, not flat ones).
template <bool LDS, int WIDE = -1>
__device__ __forceinline__ SceneView discovery_scene(const DiscoverArgs &a, float4 *smem) {
  SceneView v{a.tnodes, a.tnodes4, a.tris};
  if (LDS) {
    const bool four = WIDE < 0 ? a.tnodes4 != nullptr : WIDE == 1;
    const int n4 = four ? 8 * a.n_tnodes4 : 2 * a.n_tnodes;
    stage_scene(four ? reinterpret_cast<const float4 *>(a.tnodes4) : reinterpret_cast<const float4 *>(a.tnodes), n4,
                a.tris, a.n_tris, smem);
    if (WIDE == 1) {
      v.tnodes4 = reinterpret_cast<const AmsNode4 *>(smem);
    } else if (WIDE == 0) {
      v.tnodes = reinterpret_cast<const AmsNode *>(smem);
      v.tnodes4 = nullptr;
    } else if (four) {
      v.tnodes4 = reinterpret_cast<const AmsNode4 *>(smem);
    } else {
      v.tnodes = reinterpret_cast<const AmsNode *>(smem);
    }
    v.tris = reinterpret_cast<const AmsTri *>(smem + n4);
  }
  return v;
}

__device__ __forceinline__ int rx_query(const DiscoverArgs &a, const Ray &ray, float &ts, Counters &ct) {
  return a.lat.on ? closest_sphere_lattice(a.lat, a.sph_o, ray, AMS_TMIN, ts, a.acc2, a.r2, ct)
                  : closest_sphere(a.snodes, a.n_snodes, a.sph, ray, AMS_TMIN, ts, a.acc2, a.r2, ct);
}

// Can the ray o + t d (t >= tmin) still hit a reception sphere?  A branch
// with no reflection left (ttl == 1) records nothing else, so when the
// answer is no its remaining pass-through queries are skipped.  Exact: a hit
// needs the closest approach inside the lattice slab and extent, or inside
// the sphere hierarchy's root box.
__device__ __forceinline__ bool rx_reachable(const DiscoverArgs &a, float3 o, float3 d) {
  if (a.lat.on) {
    float ta = AMS_TMIN, tb = 1e30f, i0, di, j0, dj;
    return lattice_segment(a.lat, o, d, ta, tb, i0, di, j0, dj);
  }
  Ray ray;
  ray_setup(ray, o, d);
  const float4 *np = reinterpret_cast<const float4 *>(a.snodes);
  return box_test(ray, np[0], np[1], AMS_TMIN, INFINITY);
}

struct Branch {
  float3 o, d;
  int ttl, b, nref, last_pl;
  uint64_t khi, klo;
};

template <bool LDS, bool WIDE>
__device__ __forceinline__ void discover_rays(const DiscoverArgs &a, const int rays_per_thread) {
  extern __shared__ float4 ams_smem[];
  const SceneView sv = discovery_scene<LDS, WIDE ? 1 : 0>(a, ams_smem);
  const AmsNode *tnodes = sv.tnodes;
  const AmsNode4 *tnodes4 = sv.tnodes4;
  const AmsTri *tris = sv.tris;
  __shared__ DedupTable dd;
  dedup_init(dd);
  Counters ct;
  unsigned nrec = 0;
  // rays_per_thread passes, each over blockDim.x consecutive rays
  for (int pass = 0; pass < rays_per_thread; pass++) {
    const long long r = a.ray_begin + ((long long)blockIdx.x * rays_per_thread + pass) * blockDim.x + threadIdx.x;
    const bool active = r < a.ray_end;
    if (active) {
      long long e, slot;
      if (a.per_element) {
        e = r / a.fib_n;
        slot = r % a.fib_n;
      } else {
        e = a.perm[r] % a.n_src;
        slot = r;
      }
      const float4 dv = a.dirs[slot];
      Branch st[AMS_MAX_HIT];
      int sp = 0;
      Branch cur;
      cur.o = f3(a.src[3 * e], a.src[3 * e + 1], a.src[3 * e + 2]);
      cur.d = f3(dv.x, dv.y, dv.z);
      cur.ttl = a.K + 1;
      cur.b = 0;
      cur.nref = 0;
      cur.last_pl = -1;
      cur.khi = (0xFFFFFull << 36) | (0xFFFFFull << 16) | 0xFFFFull;
      cur.klo = (0xFull << 60) | (0xFFFFFull << 40) | (0xFFFFFull << 20);
      uint64_t last_hi = ~0ull, last_lo = ~0ull;  // previous record of this ray
      bool more = true;
      while (more) {
        // After an Rx hit the ray goes on in the same direction, so its next
        // scene hit is the one already found (tri_rem further on); only the
        // sphere query is repeated (flags bit 3 repeats both).
        bool tri_cached = false;
        float tri_rem = 0.0f;
        int tri_slot = -1, tri_orig = 0;
        while (cur.ttl > 0) {
          Ray ray;
          ray_setup(ray, cur.o, cur.d);
          float tt;
          int torig, slot;
          if (tri_cached) {
            tt = tri_rem;
            slot = tri_slot;
            torig = tri_orig;
          } else {
            // With no reflection left the scene is pass-through: a scene hit
            // changes what follows only when the ray's origin then lies in a
            // reception sphere (optixPathTracer.cu:440 rejects that sphere),
            // i.e. inside the Rx lattice's slab.  Such a branch queries the
            // scene on the part of the ray in the slab only (flags bit 18
            // queries the whole ray); before and after it the pass-through
            // hits move the origin and nothing else.
            float tlo = AMS_TMIN;
            tt = INFINITY;
            if (cur.ttl == 1 && a.lat.on && !(a.flags & 262144)) {
              float ta = AMS_TMIN, tb = 1e30f, i0, di, j0, dj;
              if (lattice_segment(a.lat, cur.o, cur.d, ta, tb, i0, di, j0, dj)) {
                tlo = ta;
                tt = tb;
              } else {
                tlo = 1.0f;  // the ray misses the slab: nothing to query
                tt = 0.0f;
              }
            }
            if (tlo <= tt) {
              slot = WIDE ? scene_closest(tnodes4, tnodes, a.n_tnodes, tris, ray, tlo, tt, cur.last_pl, torig, ct)
                          : closest_tri(tnodes, a.n_tnodes, tris, ray, tlo, tt, cur.last_pl, torig, ct);
            } else {
              slot = -1;
              torig = 0;
              tt = INFINITY;
            }
          }
          float ts = tt;
          int rx = rx_query(a, ray, ts, ct);
          ct.queries++;
          tri_cached = false;
          if (rx >= 0 && ts < tt) {
            if (!(a.flags & 8)) {
              tri_cached = true;
              tri_rem = tt - ts;
              tri_slot = slot;
              tri_orig = torig;
            }
            // Rx hit: record (key, tile) and continue from the closest approach
            int h = a.K + 1 - cur.ttl;
            uint32_t tile = ams_tile_of(rx, a.n1_dst, a.gs);
            uint64_t hi = (cur.khi & ((1ull << 56) - 1)) | ((uint64_t)h << 61) | ((uint64_t)cur.b << 56);
            uint64_t lo = (cur.klo & ~0xFFFFFull) | tile;
            if ((hi != last_hi || lo != last_lo) && dedup_first(dd, hi, lo)) {
              unsigned long long idx = wave_append(a.out_count);
              nrec++;
              if (idx < a.cap) {
                a.out_hi[idx] = hi;
                a.out_lo[idx] = lo;
              }
            }
            last_hi = hi;
            last_lo = lo;
            cur.o = cur.o + cur.d * ts;
            cur.last_pl = -1;
            continue;
          }
          if (slot < 0) break;  // miss, optixPathTracer.cu:276-279
          const AmsTri tr = tris[slot];
          int pl = __float_as_int(tr.b.w);
          cur.o = cur.o + cur.d * tt;
          cur.last_pl = pl;
          if (cur.ttl > 1) {
            // optixPathTracer.cu:294-328
            float4 nm = a.aux[torig].n_maxedge;
            float3 N0 = f3(nm.x, nm.y, nm.z);
            float3 N = dot3(N0, cur.d) > 0.0f ? N0 * -1.0f : N0;  // faceforward(N0, -d, N0)
            int m = __float_as_int(tr.c.w);
            int th = theta_bin(cur.d, N);
            bool pen = a.mat_prop[th + 19 * m] < 0.0f;
            cur.ttl--;
            int h = a.K + 1 - cur.ttl;
            if (pen && sp < AMS_MAX_HIT) st[sp++] = cur;  // transmitted parent continues later
            cur.b |= 1 << (h - 1);
            cur.d = norm3(cur.d - N * (2.0f * dot3(cur.d, N)));
            if (cur.nref < AMS_MAX_HIT) key_set_plane(cur.khi, cur.klo, cur.nref, (uint32_t)pl);
            cur.nref++;
          }
          // ttl == 1: scene is pass-through (optixPathTracer.cu:294)
          if (cur.ttl == 1 && !(a.flags & 128) && !rx_reachable(a, cur.o, cur.d)) break;
        }
        if (sp > 0) cur = st[--sp];
        else more = false;
      }
    }
  }
  wave_add_counters(ctr_slot(a.counters), ct);
  unsigned long long x = nrec;
  for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off);
  if (__lane_id() == 0 && x) atomicAdd(ctr_slot(a.counters) + 4, x);
}

// Forward discovery (optixPathTracer.cu:142-190, one sample per ray): each
// thread walks the branch trees of its rays with a local stack; the two
// launch shapes above.
template <bool LDS>
__global__ void __launch_bounds__(kDiscBlock) __attribute__((amdgpu_waves_per_eu(7, 8)))
k_discover(DiscoverArgs a) {
  discover_rays<LDS, false>(a, kDiscRays);
}
template <bool LDS>
__global__ void __launch_bounds__(kDiscWideBlock) k_discover_wide(DiscoverArgs a) {
  discover_rays<LDS, true>(a, kDiscWideRays);
}

// Wavefront-persistent form of the same discovery: the grid is sized to the
// machine, every loop iteration performs one closest-hit query per lane, and
// a lane whose ray has ended takes the next ray index from a global counter
// (one atomic per wave per refill).  Lanes therefore stay busy however many
// events their rays have; each wave leaves the loop once the counter is past
// the end of the range and none of its lanes holds a ray.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6, 8)))
k_discover_persistent(DiscoverArgs a, unsigned long long *next_ray) {
  Counters ct;
  unsigned nrec = 0;
  Branch st[AMS_MAX_HIT];
  int sp = 0;
  Branch cur;
  bool have = false;
  uint64_t last_hi = ~0ull, last_lo = ~0ull;  // the lane's previous record (consecutive duplicates)
  bool tri_cached = false;                     // scene hit kept across an Rx hit (see k_discover)
  float tri_rem = 0.0f;
  int tri_slot = -1, tri_orig = 0;
  const int lane = __lane_id();
  while (true) {
    unsigned long long need = __ballot(!have);
    if (need) {
      int leader = __ffsll((long long)need) - 1;
      unsigned long long base = 0;
      if (lane == leader) base = atomicAdd(next_ray, (unsigned long long)__popcll(need));
      base = __shfl(base, leader);
      if (!have) {
        long long r = a.ray_begin + (long long)(base + __popcll(need & ((1ull << lane) - 1ull)));
        if (r < a.ray_end) {
          long long e, k;
          if (a.per_element) {
            e = (r / a.fib_n);
            k = a.perm[r % a.fib_n];
          } else {
            long long rr2 = a.perm[r];
            e = rr2 % a.n_src;
            k = rr2;
          }
          double dx, dy, dz;
          fib_dir(k, a.fib_n, dx, dy, dz);
          cur.o = f3(a.src[3 * e], a.src[3 * e + 1], a.src[3 * e + 2]);
          cur.d = f3((float)dx, (float)dy, (float)dz);
          cur.ttl = a.K + 1;
          cur.b = 0;
          cur.nref = 0;
          cur.last_pl = -1;
          cur.khi = (0xFFFFFull << 36) | (0xFFFFFull << 16) | 0xFFFFull;
          cur.klo = (0xFull << 60) | (0xFFFFFull << 40) | (0xFFFFFull << 20);
          last_hi = ~0ull;
          last_lo = ~0ull;
          tri_cached = false;
          have = true;
        }
      }
    }
    if (__ballot(have) == 0) break;
    if (have) {
      Ray ray;
      ray_setup(ray, cur.o, cur.d);
      float tt;
      int torig, slot;
      if (tri_cached) {
        tt = tri_rem;
        slot = tri_slot;
        torig = tri_orig;
      } else {
        tt = INFINITY;
        slot = closest_tri(a.tnodes, a.n_tnodes, a.tris, ray, AMS_TMIN, tt, cur.last_pl, torig, ct);
      }
      float ts = tt;
      int rx = rx_query(a, ray, ts, ct);
      ct.queries++;
      bool ended = false;
      tri_cached = false;
      if (rx >= 0 && ts < tt) {
        if (!(a.flags & 8)) {
          tri_cached = true;
          tri_rem = tt - ts;
          tri_slot = slot;
          tri_orig = torig;
        }
        int h = a.K + 1 - cur.ttl;
        uint32_t tile = ams_tile_of(rx, a.n1_dst, a.gs);
        uint64_t hi = (cur.khi & ((1ull << 56) - 1)) | ((uint64_t)h << 61) | ((uint64_t)cur.b << 56);
        uint64_t lo = (cur.klo & ~0xFFFFFull) | tile;
        if (hi != last_hi || lo != last_lo) {
          last_hi = hi;
          last_lo = lo;
          unsigned long long idx = wave_append(a.out_count);
          nrec++;
          if (idx < a.cap) {
            a.out_hi[idx] = hi;
            a.out_lo[idx] = lo;
          }
        }
        cur.o = cur.o + cur.d * ts;
        cur.last_pl = -1;
      } else if (slot < 0) {
        ended = true;  // miss, optixPathTracer.cu:276-279
      } else {
        const AmsTri tr = a.tris[slot];
        int pl = __float_as_int(tr.b.w);
        cur.o = cur.o + cur.d * tt;
        cur.last_pl = pl;
        if (cur.ttl > 1) {
          float4 nm = a.aux[torig].n_maxedge;
          float3 N0 = f3(nm.x, nm.y, nm.z);
          float3 N = dot3(N0, cur.d) > 0.0f ? N0 * -1.0f : N0;
          int m = __float_as_int(tr.c.w);
          int th = theta_bin(cur.d, N);
          bool pen = a.mat_prop[th + 19 * m] < 0.0f;
          cur.ttl--;
          int h = a.K + 1 - cur.ttl;
          if (pen && sp < AMS_MAX_HIT) st[sp++] = cur;
          cur.b |= 1 << (h - 1);
          cur.d = norm3(cur.d - N * (2.0f * dot3(cur.d, N)));
          if (cur.nref < AMS_MAX_HIT) key_set_plane(cur.khi, cur.klo, cur.nref, (uint32_t)pl);
          cur.nref++;
        }
        if (cur.ttl == 1 && !(a.flags & 128) && !rx_reachable(a, cur.o, cur.d)) ended = true;
      }
      if (ended || cur.ttl <= 0) {
        tri_cached = false;
        if (sp > 0) cur = st[--sp];
        else have = false;
      }
    }
  }
  wave_add_counters(ctr_slot(a.counters), ct);
  unsigned long long x = nrec;
  for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off);
  if (__lane_id() == 0 && x) atomicAdd(ctr_slot(a.counters) + 4, x);
}

// ---- reference-compatible discovery (optixPathTracer.cu:142-190) ----------
// TEA with 4 rounds (Wheeler and Needham 1994; key words as GPU samplers use
// them) seeds a per-thread Numerical-Recipes LCG; 24-bit uniforms.
__device__ __forceinline__ unsigned tea4(unsigned v0, unsigned v1) {
  unsigned s0 = 0;
  for (int n = 0; n < 4; n++) {
    s0 += 0x9e3779b9u;
    v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
    v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
  }
  return v0;
}
__device__ __forceinline__ float lcg_uniform(unsigned &st) {
  st = 1664525u * st + 1013904223u;
  return (float)(st & 0x00FFFFFFu) / (float)0x01000000u;
}

struct RefArgs {
  DiscoverArgs d;
  int launch_x, launch_y, samples, reserve, n_grid;
  int *min_hc;  // [threads][n_grid], zeroed
};

// One launch index per thread; the 32-slot branch tree is walked in slot
// order (a reflection child lands in a higher slot than its parent), so the
// order of Rx hits -- and with it the per-tile min_hc gating and the
// reserve cap -- follows the reference.
template <bool LDS, bool WIDE>
__global__ void __launch_bounds__(256) k_discover_ref(RefArgs ra) {
  const DiscoverArgs &a = ra.d;
  extern __shared__ float4 ams_smem[];
  const SceneView sv = discovery_scene<LDS, WIDE ? 1 : 0>(a, ams_smem);
  const AmsNode *tnodes = sv.tnodes;
  const AmsNode4 *tnodes4 = sv.tnodes4;
  const AmsTri *tris = sv.tris;
  __shared__ DedupTable dd;
  dedup_init(dd);
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  Counters ct;
  unsigned nrec = 0;
  if (t < ra.launch_x * ra.launch_y) {
    const int ix = t / ra.launch_y, iy = t % ra.launch_y;
    unsigned seed = tea4((unsigned)t, 0u);
    int *min_hc = ra.min_hc + (size_t)t * ra.n_grid;
    int res_cnt = 0;
    Branch slot[1 << AMS_MAX_HIT];
    for (int smp = 0; smp < ra.samples; smp++) {
      int src = (int)(lcg_uniform(seed) * (float)a.n_src);
      if (src >= a.n_src) src = a.n_src - 1;
      // rand_dir, optixPathTracer.cu:56-68
      float r1 = (lcg_uniform(seed) + (float)ix) / (float)ra.launch_x;
      float r2 = (lcg_uniform(seed) + (float)iy) / (float)ra.launch_y;
      float th = 2.0f * (float)M_PI * r1, rr = sqrtf(r2);
      float u = rr * cosf(th), v = rr * sinf(th), q = sqrtf(fmaxf(0.0f, 1.0f - rr * rr));
      for (int k = 0; k < (1 << AMS_MAX_HIT); k++) slot[k].ttl = 0;
      Branch &s0 = slot[0];
      s0.o = f3(a.src[3 * src], a.src[3 * src + 1], a.src[3 * src + 2]);
      s0.d = f3(2.0f * u * q, 2.0f * v * q, 1.0f - 2.0f * rr * rr);
      s0.ttl = a.K + 1;
      s0.b = 0;
      s0.nref = 0;
      s0.last_pl = -1;
      s0.khi = (0xFFFFFull << 36) | (0xFFFFFull << 16) | 0xFFFFull;
      s0.klo = (0xFull << 60) | (0xFFFFFull << 40) | (0xFFFFFull << 20);
      for (int i = 0; i < (1 << AMS_MAX_HIT); i++) {
        Branch cur = slot[i];
        bool tri_cached = false;  // scene hit kept across an Rx hit (see k_discover)
        float tri_rem = 0.0f;
        int tri_slot = -1, tri_orig = 0;
        while (cur.ttl > 0) {
          Ray ray;
          ray_setup(ray, cur.o, cur.d);
          float tt;
          int torig, sl;
          if (tri_cached) {
            tt = tri_rem;
            sl = tri_slot;
            torig = tri_orig;
          } else {
            tt = INFINITY;
            sl = scene_closest(WIDE ? tnodes4 : nullptr, tnodes, a.n_tnodes, tris, ray, AMS_TMIN, tt, cur.last_pl,
                               torig, ct);
          }
          float ts = tt;
          int rx = rx_query(a, ray, ts, ct);
          ct.queries++;
          tri_cached = false;
          if (rx >= 0 && ts < tt) {
            if (!(a.flags & 8)) {
              tri_cached = true;
              tri_rem = tt - ts;
              tri_slot = sl;
              tri_orig = torig;
            }
            if (res_cnt < ra.reserve) {  // optixPathTracer.cu:173-185
              int grid = (int)ams_tile_of(rx, a.n1_dst, a.gs);
              int h = a.K + 1 - cur.ttl;
              if (min_hc[grid] <= h) {
                min_hc[grid] = h + 1;
                uint64_t hi = (cur.khi & ((1ull << 56) - 1)) | ((uint64_t)h << 61) | ((uint64_t)cur.b << 56);
                uint64_t lo = (cur.klo & ~0xFFFFFull) | (uint64_t)grid;
                if (dedup_first(dd, hi, lo)) {
                  unsigned long long idx = wave_append(a.out_count);
                  nrec++;
                  if (idx < a.cap) {
                    a.out_hi[idx] = hi;
                    a.out_lo[idx] = lo;
                  }
                }
                res_cnt++;
              }
            }
            cur.o = cur.o + cur.d * ts;
            cur.last_pl = -1;
            continue;
          }
          if (sl < 0) break;
          const AmsTri tr = tris[sl];
          int pl = __float_as_int(tr.b.w);
          cur.o = cur.o + cur.d * tt;
          cur.last_pl = pl;
          if (cur.ttl > 1) {
            float4 nm = a.aux[torig].n_maxedge;
            float3 N0 = f3(nm.x, nm.y, nm.z);
            float3 N = dot3(N0, cur.d) > 0.0f ? N0 * -1.0f : N0;
            int m = __float_as_int(tr.c.w);
            int thb = theta_bin(cur.d, N);
            bool pen = a.mat_prop[thb + 19 * m] < 0.0f;
            cur.ttl--;
            int h = a.K + 1 - cur.ttl;
            int cs = cur.b | (1 << (h - 1));  // optixPathTracer.cu:313
            Branch ch = cur;
            ch.b = cs;
            ch.d = norm3(cur.d - N * (2.0f * dot3(cur.d, N)));
            if (ch.nref < AMS_MAX_HIT) key_set_plane(ch.khi, ch.klo, ch.nref, (uint32_t)pl);
            ch.nref++;
            slot[cs] = ch;
            if (!pen) break;
          }
          if (cur.ttl == 1 && !(a.flags & 128) && !rx_reachable(a, cur.o, cur.d)) break;
        }
      }
    }
  }
  wave_add_counters(ctr_slot(a.counters), ct);
  unsigned long long x = nrec;
  for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off);
  if (__lane_id() == 0 && x) atomicAdd(ctr_slot(a.counters) + 4, x);
}

// ------------------------------------------------------------------------
// backward evaluation
// ------------------------------------------------------------------------
// H is accumulated in 64-bit fixed point, 2^56 units per 1 (resolution
// 1.4e-17, |H| < 2^7); a lane's sum of magnitude kHLimit or more is counted
// in counters[15] and the trace is refused.
constexpr double kHScale = 72057594037927936.0;
constexpr double kHLimit = 64.0;

struct EvalArgs {
  const AmsNode *tnodes;
  int n_tnodes;
  const AmsNode4 *tnodes4;  // four-wide hierarchy (k_evaluate_wide only)
  const AmsTri *tris;
  int n_tris;
  const AmsTriAux *aux;
  const float *mat_prop;
  const float *refl_phase;
  const AmsPlane *planes;
  const float4 *planes_f;  // the planes in f32, (n, d)
  const AmsPlaneBox *pbox;
  const AmsPlane *tplane;  // each triangle's own plane, by original index
  const float *src;
  int n_src;
  const float *dst;
  int n0_dst, n1_dst, n_dst, gs, n1g, flags;
  double inv_wave_len_d;  // 1 / wave_len in FP64
  float wave_len;
  float src_r2, src_acc2;
  float3 src_dir, dst_dir;
  int src_bptype, dst_bptype;
  const uint64_t *path_hi, *path_lo;
  const int *task_path;       // tasks of this round, grouped by tile
  const int *unit_tile;       // work unit -> tile
  const int *unit_begin;      // work unit -> first task
  const int *unit_end;        // work unit -> end task
  int items;                  // gs*gs*n_src
  long long *H;               // [n_src][n_dst][2], fixed point (kHScale)
  unsigned char *task_ok;
  unsigned long long *counters;  // nodes, tris, spheres, queries, chains, successes
  ams_task_info *info;           // task export of this round (EXPORT instances), [src][task][gs2]
  long long info_nt;             // tasks of this round
  int gs2;                       // Rx per tile
  // candidate lists of the launch's (task, Rx) pairs (k_beam), or null:
  // [unit][task in unit][Rx in tile][beam_stride] ints
  const int *beam;
  int beam_stride, beam_slice;
};

// Beam candidate lists.  For a (task, Rx) pair of a path whose interactions
// are all reflections, every Tx element's backward chain runs leg by leg
// inside one convex beam per leg: the hull of where the leg starts and ends
// for the corners of the Tx array's bounding box (the reflection points of
// an element inside that box lie inside the hull of the corners'
// reflection points, a central projection of the unfolded images).  k_beam
// lists, per leg, the triangles that come within a margin of that beam; a
// lane then finds its leg's closest hit among them instead of walking the
// hierarchy.  Header: 5 bits per leg (leg 0 = the Rx side), the candidate
// count or kBeamWalk (walk the hierarchy: too many candidates, a
// transmission on the path, a beam the geometry does not close).
constexpr int kBeamCand = 24;
constexpr unsigned kBeamWalk = 31u;
constexpr int kBeamBits = 5;
constexpr int kBeamBlock = 128;
#ifndef AMS_BEAM_WAVES
#define AMS_BEAM_WAVES 6  // waves per SIMD of the evaluation with candidate lists
#endif

// First int of a (task, Rx) pair's lists: task tl of the launch's unit
// blockIdx.y, Rx q of the tile.
__device__ __forceinline__ size_t beam_index(const EvalArgs &a, int tl, int q) {
  return (((size_t)blockIdx.y * a.beam_slice + tl) * a.gs2 + q) * a.beam_stride;
}

// Wave-uniform reads through the constant address space, so the compiler
// issues scalar loads (the lists, the scene, the path keys, the tasks'
// paths and the planes are not written by the kernels that read them).
typedef const __attribute__((address_space(4))) int cint4_t;
typedef const __attribute__((address_space(4))) float cfloat4_t;
typedef const __attribute__((address_space(4))) double cdouble4_t;
typedef const __attribute__((address_space(4))) uint64_t cu64_4_t;
__device__ __forceinline__ int uni_load(const int *p) { return *(cint4_t *)p; }
__device__ __forceinline__ float4 uni_f4(const float4 *p) {
  cfloat4_t *q = (cfloat4_t *)p;
  return make_float4(q[0], q[1], q[2], q[3]);
}
__device__ __forceinline__ AmsPlane uni_plane(const AmsPlane *p) {
  cdouble4_t *q = (cdouble4_t *)p;
  AmsPlane P;
  P.n[0] = q[0];
  P.n[1] = q[1];
  P.n[2] = q[2];
  P.d = q[3];
  return P;
}
__device__ __forceinline__ AmsTri uni_tri(const AmsTri *p) {
  cfloat4_t *q = (cfloat4_t *)p;
  AmsTri t;
  t.a = make_float4(q[0], q[1], q[2], q[3]);
  t.b = make_float4(q[4], q[5], q[6], q[7]);
  t.c = make_float4(q[8], q[9], q[10], q[11]);
  return t;
}

// Closest hit among a leg's candidate triangles (leaf-order slots), with
// closest_tri's rules: [tmin, tbest], plane skip_pl ignored, ties on t to the
// smaller original index.  Every triangle a ray of the beam can hit is a
// candidate, so the result is the walk's.
// SCALAR: cand and the triangles are wave-uniform and in global memory.
template <bool SCALAR>
__device__ __forceinline__ int cand_closest(const AmsTri *__restrict__ tris, const int *__restrict__ cand, int n,
                                            const Ray &r, float tmin, float &tbest, int skip_pl, int &orig_out,
                                            Counters &ct) {
  int hit = -1, horig = 0x7fffffff;
  for (int k = 0; k < n; k++) {
    const int slot = SCALAR ? uni_load(cand + k) : cand[k];
    const AmsTri tr = SCALAR ? uni_tri(tris + slot) : tris[slot];
    ct.tris++;
    if (__float_as_int(tr.b.w) == skip_pl) continue;
    float t;
    if (tri_test(r, tr.a, tr.b, tr.c, tmin, tbest, t)) {
      int orig = __float_as_int(tr.a.w);
      if (t < tbest || orig < horig) {
        tbest = t;
        hit = slot;
        horig = orig;
      }
    }
  }
  orig_out = horig;
  return hit;
}

// closest_tri for a wave whose active lanes walk together (their rays leave
// one Rx in nearly one direction): the wave visits a node when any lane's
// box test passes, so the node index is wave-uniform and the node and its
// triangles are read with scalar loads.  A lane's own walk visits a subset
// of these nodes and tests every triangle it could hit, so its closest hit
// (ties to the smaller original index) is closest_tri's.
__device__ __forceinline__ int closest_tri_wave(const AmsNode *__restrict__ nodes, int n_nodes,
                                                const AmsTri *__restrict__ tris, const Ray &r, float tmin,
                                                float &tbest, int skip_pl, int &orig_out, Counters &ct) {
  int hit = -1, horig = 0x7fffffff;
  int i = 0;
  while (i < n_nodes) {
    const float4 lo = uni_f4(reinterpret_cast<const float4 *>(nodes + i));
    const float4 hi = uni_f4(reinterpret_cast<const float4 *>(nodes + i) + 1);
    ct.nodes++;
    const int miss = __float_as_int(lo.w), info = __float_as_int(hi.w);
    const bool h = box_test(r, lo, hi, tmin, tbest);
    if (__ballot(h)) {
      if (info == 0) {
        i++;
        continue;
      }
      if (h) {
        const int first = info >> 3, cnt = info & 7;
        for (int k = 0; k < cnt; k++) {
          const AmsTri tr = uni_tri(tris + first + k);
          ct.tris++;
          if (__float_as_int(tr.b.w) == skip_pl) continue;
          float t;
          if (tri_test(r, tr.a, tr.b, tr.c, tmin, tbest, t)) {
            int orig = __float_as_int(tr.a.w);
            if (t < tbest || orig < horig) {
              tbest = t;
              hit = first + k;
              horig = orig;
            }
          }
        }
      }
    }
    i = miss;
  }
  orig_out = horig;
  return hit;
}

// Path length of a successful chain along its reflection triangles' own
// planes, in FP64: the length the chain traced (optixPathTracer.cu:344 sums
// the traced segments), for a chain whose reflection triangles are only
// nearly on the planes of their plane ids.  The image of the Tx element
// across the key's planes gives the first leg's direction, as in the walk;
// rt holds the chain's reflection triangles (original indices) in the order
// the walk meets them from the Rx, stride kEvalBlock.
__device__ __forceinline__ double own_plane_length(const EvalArgs &a, uint64_t khi, uint64_t klo, int nref, float3 s3,
                                                float3 r3, const int *rt) {
  double Ix = s3.x, Iy = s3.y, Iz = s3.z;
  for (int k = 0; k < nref; k++) {
    const AmsPlane P = a.planes[key_get_plane(khi, klo, k)];
    double f = 2.0 * (Ix * P.n[0] + Iy * P.n[1] + Iz * P.n[2] - P.d);
    Ix -= f * P.n[0];
    Iy -= f * P.n[1];
    Iz -= f * P.n[2];
  }
  double ox = r3.x, oy = r3.y, oz = r3.z;
  double dx = Ix - ox, dy = Iy - oy, dz = Iz - oz;
  const double il = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
  dx *= il;
  dy *= il;
  dz *= il;
  double L = 0.0;
  for (int k = 0; k < nref; k++) {
    const AmsPlane P = a.tplane[rt[k * kEvalBlock]];
    const double nd = P.n[0] * dx + P.n[1] * dy + P.n[2] * dz;
    const double t = (P.d - (P.n[0] * ox + P.n[1] * oy + P.n[2] * oz)) / nd;
    ox += dx * t;
    oy += dy * t;
    oz += dz * t;
    L += t;
    dx -= 2.0 * nd * P.n[0];
    dy -= 2.0 * nd * P.n[1];
    dz -= 2.0 * nd * P.n[2];
  }
  // the closest approach to the Tx element on the last leg
  return L + ((double)s3.x - ox) * dx + ((double)s3.y - oy) * dy + ((double)s3.z - oz) * dz;
}

// One lane per (Rx in the tile, Tx element), looping over the tasks of its
// work unit.  WIDE: the scene query walks the four-wide hierarchy (global
// memory scenes: fewer dependent L2 loads per query).
// UNI (with BEAM): every wave's lanes share one Rx of the tile (the Tx count
// is a multiple of 64), so a chain's lists and candidate triangles are
// wave-uniform: scalar loads, one branch per wave.
template <bool LDS, bool EXPORT, bool WIDE, bool BEAM, bool UNI>
__device__ __forceinline__ void evaluate_body(const EvalArgs &a) {
  extern __shared__ float4 ams_smem[];
  // reflection triangles of each lane's current chain
  __shared__ int refl_tri[AMS_MAX_HIT * kEvalBlock];
  // each lane's candidate-list header of its current chain (BEAM)
  __shared__ unsigned beam_hdr[BEAM ? kEvalBlock : 1];
  const AmsNode *tnodes = a.tnodes;
  const AmsTri *tris = a.tris;
  if (LDS) {
    stage_scene(reinterpret_cast<const float4 *>(a.tnodes), 2 * a.n_tnodes, a.tris, a.n_tris, ams_smem);
    tnodes = reinterpret_cast<const AmsNode *>(ams_smem);
    tris = reinterpret_cast<const AmsTri *>(ams_smem + 2 * a.n_tnodes);
  }
  int unit = blockIdx.y;
  int item = blockIdx.x * blockDim.x + threadIdx.x;
  int g = a.unit_tile[unit];
  int tb = a.unit_begin[unit], te = a.unit_end[unit];
  int q = item / a.n_src, i = item % a.n_src;
  int gi = g / a.n1g, gj = g % a.n1g;
  int di = gi * a.gs + q / a.gs, dj = gj * a.gs + q % a.gs;
  int j = di * a.n1_dst + dj;
  // optixPathTracer.cu:211-220 (flags bit 0 keeps the reference's row wrap)
  bool active = item < a.items && j < a.n_dst && (dj < a.n1_dst || (a.flags & 1));
  Counters ct;
  unsigned long long nchain = 0, nsucc = 0;
  double acc_re = 0.0, acc_im = 0.0;
  // element positions are f32 inputs: kept in f32 registers, widened exactly
  // to FP64 where the image and the path length are formed
  float3 s3 = f3(0, 0, 0), r3 = f3(0, 0, 0);
  if (active) {
    s3 = f3(a.src[3 * i], a.src[3 * i + 1], a.src[3 * i + 2]);
    r3 = f3(a.dst[3 * j], a.dst[3 * j + 1], a.dst[3 * j + 2]);
  }
  for (int t = tb; t < te; t++) {
    bool ok = false;
    float x_loss = 1.0f, x_launch = 0.0f, x_phase = 0.0f, x_recv = 0.0f;  // task export of the chain
    int x_hits = 0;
    int x_idx[AMS_MAX_HIT] = {0, 0, 0, 0, 0};
    if (active) {
      const int p = UNI ? uni_load(a.task_path + t) : a.task_path[t];
      const uint64_t khi = UNI ? *(cu64_4_t *)(a.path_hi + p) : a.path_hi[p];
      const uint64_t klo = UNI ? *(cu64_4_t *)(a.path_lo + p) : a.path_lo[p];
      int h = (int)(khi >> 61), bm = (int)((khi >> 56) & 31);
      int nref = __popc(bm & ((1 << h) - 1));
      // image of the Tx element across the reflection planes (FP64)
      double3 I = make_double3(s3.x, s3.y, s3.z);
      for (int k = 0; k < nref; k++) {
        const uint32_t pid = key_get_plane(khi, klo, k);
        const AmsPlane P = UNI ? uni_plane(a.planes + pid) : a.planes[pid];
        double s = 2.0 * (I.x * P.n[0] + I.y * P.n[1] + I.z * P.n[2] - P.d);
        I.x -= s * P.n[0];
        I.y -= s * P.n[1];
        I.z -= s * P.n[2];
      }
      double ex = I.x - (double)r3.x, ey = I.y - (double)r3.y, ez = I.z - (double)r3.z;
      double L = sqrt(ex * ex + ey * ey + ez * ez);
      const double il = 1.0 / L;
      float3 d = f3((float)(ex * il), (float)(ey * il), (float)(ez * il));
      float3 o = r3;
      float A = beam_pattern(d, a.dst_dir, a.dst_bptype);
      float D = 0.0f, Dres = 0.0f, far = 1e6f, psi = 0.0f;
      // task export (optixPathTracer.cu:243-250): cosines as the reference forms them
      if (EXPORT) x_recv = dot3(d, a.dst_dir) / (sqrtf(dot3(d, d)) * sqrtf(dot3(a.dst_dir, a.dst_dir)));
      int ttl = h + 1, ri = 0, last_pl = -1;
      bool inexact = false;  // a reflection on a triangle off its plane id's plane
      nchain++;
      // the pair's candidate-list header (k_beam): in LDS per lane, or
      // wave-uniform (UNI)
      unsigned bhdr_u = kBeamWalk;
      if (BEAM && UNI) bhdr_u = (unsigned)uni_load(a.beam + beam_index(a, t - tb, __builtin_amdgcn_readfirstlane(q)));
      else if (BEAM) beam_hdr[threadIdx.x] = (unsigned)a.beam[beam_index(a, t - tb, q)];
      // Unfold the path analytically before any query: from the Rx towards
      // the image, every reflection point must lie ahead and inside the
      // bounds of its plane's triangles, and the Tx ahead of the last one.
      // A chain failing that (with 0.1 mm / 1 mm of slack over the FP32
      // walk) fails the traced test as well, so it is skipped unqueried.
      if (!(a.flags & 256)) {
        float3 po = o, pd = d;
        bool geo = true;
        for (int q = 0; q < nref && geo; q++) {
          const uint32_t pid = key_get_plane(khi, klo, nref - 1 - q);
          const float4 P = UNI ? uni_f4(a.planes_f + pid) : a.planes_f[pid];
          float3 pn = f3(P.x, P.y, P.z);
          float nd = dot3(pn, pd);
          // 1-ulp reciprocal: far inside the check's 0.1 mm / 1 mm slack
          float tp = nd != 0.0f ? (P.w - dot3(pn, po)) * __builtin_amdgcn_rcpf(nd) : -1.0f;
          if (!(tp > -1e-4f)) {
            geo = false;
            break;
          }
          float3 X = po + pd * tp;
          AmsPlaneBox bx;
          if (UNI) {
            bx.lo = uni_f4(&a.pbox[pid].lo);
            bx.hi = uni_f4(&a.pbox[pid].hi);
          } else {
            bx = a.pbox[pid];
          }
          float mg = 1e-3f + 1e-6f * fmaxf(fmaxf(fabsf(X.x), fabsf(X.y)), fabsf(X.z));
          if (X.x < bx.lo.x - mg || X.x > bx.hi.x + mg || X.y < bx.lo.y - mg || X.y > bx.hi.y + mg ||
              X.z < bx.lo.z - mg || X.z > bx.hi.z + mg) {
            geo = false;
            break;
          }
          pd = norm3(pd - pn * (2.0f * dot3(pd, pn)));
          po = X;
        }
        if (geo && !(dot3(s3 - po, pd) > -1e-4f)) geo = false;
        if (!geo) ttl = 0;  // no leg is traced; the chain counts as failed
      }
      while (ttl > 0) {
        // own Tx sphere (other Tx spheres are pass-through: optixPathTracer.cu:396)
        float3 O = s3 - o;
        float bs = dot3(O, d);
        float3 Pp = O - d * bs;
        bool sh = dot3(Pp, Pp) < a.src_acc2 && bs >= AMS_TMIN && dot3(O, O) > a.src_r2;
        // Bound the query by what the path allows: on the last leg nothing may
        // lie before the Tx; before a reflection the hit must lie on the key's
        // plane, so anything beyond that plane cannot be the interaction.
        // Both bounds leave the outcome of the unbounded query unchanged.
        float tlim = INFINITY;
        if (ttl == 1) {
          if (!sh) break;  // the ray passes the Tx by: no success is possible
          tlim = bs;
        } else if ((bm >> (h + 1 - ttl)) & 1) {
          if (ri >= nref) break;
          const uint32_t pid = key_get_plane(khi, klo, nref - 1 - ri);
          const float4 P = UNI ? uni_f4(a.planes_f + pid) : a.planes_f[pid];
          float3 pn = f3(P.x, P.y, P.z);
          float nd = dot3(pn, d);
          float tp = nd != 0.0f ? (P.w - dot3(pn, o)) * __builtin_amdgcn_rcpf(nd) : -1.0f;
          if (!(tp > 0.0f)) break;  // the plane is not ahead
          tlim = tp * 1.001f + 1e-3f;
        }
        Ray ray;
        ray_setup(ray, o, d);
        float tt = tlim;
        int torig;
        int slot;
        const int leg = UNI ? __builtin_amdgcn_readfirstlane(h + 1 - ttl) : h + 1 - ttl;
        const unsigned nc = !BEAM ? kBeamWalk
                            : ((UNI ? bhdr_u : beam_hdr[threadIdx.x]) >> (kBeamBits * leg)) & kBeamWalk;
        if (nc != kBeamWalk) {
          const int *cand =
              a.beam + beam_index(a, t - tb, UNI ? __builtin_amdgcn_readfirstlane(q) : q) + 1 + leg * kBeamCand;
          slot = cand_closest<UNI && !LDS>(tris, cand, (int)nc, ray, AMS_TMIN, tt, last_pl, torig, ct);
        } else {
          slot = WIDE ? scene_closest(a.tnodes4, tnodes, a.n_tnodes, tris, ray, AMS_TMIN, tt, last_pl, torig, ct)
                 : (UNI && !LDS) ? closest_tri_wave(tnodes, a.n_tnodes, tris, ray, AMS_TMIN, tt, last_pl, torig, ct)
                                 : closest_tri(tnodes, a.n_tnodes, tris, ray, AMS_TMIN, tt, last_pl, torig, ct);
          if (BEAM) ct.sph++;  // a walk of the hierarchy
        }
        ct.queries++;
        if (slot < 0) tt = INFINITY;
        if (sh && bs < tt && bs <= tlim) {
          if (ttl == 1) {
            // optixPathTracer.cu:398-406; the phase is added after the walk
            D += bs;
            A *= beam_pattern(d * -1.0f, a.src_dir, a.src_bptype) * a.wave_len /
                 (4.0f * (float)M_PI * (Dres + D));
            if (EXPORT) {  // optixPathTracer.cu:407-412
              x_loss *= a.wave_len / (4.0f * (float)M_PI * (Dres + D));
              x_launch = -dot3(d, a.src_dir) / (sqrtf(dot3(d, d)) * sqrtf(dot3(a.src_dir, a.src_dir)));
            }
            ok = true;
          }
          break;
        }
        if (slot < 0) break;  // miss
        const AmsTri tr = tris[slot];
        int pl = __float_as_int(tr.b.w);
        o = o + d * tt;
        D += tt;
        last_pl = pl;
        if (ttl == 1) break;  // blocked, optixPathTracer.cu:348-349
        float4 nm = a.aux[torig].n_maxedge;
        float3 N0 = f3(nm.x, nm.y, nm.z);
        float3 N = dot3(N0, d) > 0.0f ? N0 * -1.0f : N0;
        int m = __float_as_int(tr.c.w);
        int th = theta_bin(d, N);
        float loss = fabsf(a.mat_prop[th + 19 * m]);
        if (EXPORT && x_hits < AMS_MAX_HIT) x_idx[x_hits++] = th + 19 * m;  // optixPathTracer.cu:378-383
        ttl--;
        int k = h + 1 - ttl;
        bool refl = (bm >> (k - 1)) & 1;
        if (refl) {
          if (ri >= nref || (uint32_t)pl != key_get_plane(khi, klo, nref - 1 - ri)) break;
          refl_tri[ri * kEvalBlock + threadIdx.x] = torig;  // in walk order (Rx side first)
          inexact |= nm.w < 0.0f;
          ri++;
        }
        // optixPathTracer.cu:369-391
        float fspl = 1.0f;
        if (D > far) {
          fspl = a.wave_len / (4.0f * (float)M_PI * Dres);
          Dres = 0.0f;
        }
        A *= loss * fspl;
        if (EXPORT) x_loss *= fspl;
        Dres += D;
        float fk = 2.0f * nm.w * nm.w / a.wave_len;
        far = D > fk ? fk : 1e6f;
        D = 0.0f;
        if (refl) {
          d = norm3(d - N * (2.0f * dot3(d, N)));
          psi += a.refl_phase[m];
        }
      }
      if (ok) {
        // the path length in FP64, range-reduced in cycles before the FP32 sincos
        if (inexact) L = own_plane_length(a, khi, klo, nref, s3, r3, refl_tri + threadIdx.x);
        double cyc = L * a.inv_wave_len_d;
        float ang = (float)(2.0 * M_PI * (cyc - floor(cyc))) - psi;
        if (EXPORT) x_phase = (float)(2.0 * M_PI * cyc) - psi;
        float sn, cs;
        sincosf(ang, &sn, &cs);
        acc_re += (double)(A * cs);
        acc_im += (double)(-A * sn);
      }
    }
    if (EXPORT && active) {
      ams_task_info &rec = a.info[((long long)i * a.info_nt + t) * a.gs2 + q];
      rec.successful = ok ? 1 : 0;
      rec.launch_angle = x_launch;
      rec.receive_angle = x_recv;
      rec.loss = x_loss;
      rec.phase = x_phase;
      rec.src = i;
      rec.dst = j;
      rec.hit_cnt = x_hits;
      for (int h2 = 0; h2 < AMS_MAX_HIT; h2++) rec.hit_loss_idx[h2] = x_idx[h2];
    }
    if (ok) nsucc++;
    unsigned long long any = __ballot(ok);
    if (any && __lane_id() == __ffsll((long long)any) - 1) a.task_ok[t] = 1;
  }
  if (active && (acc_re != 0.0 || acc_im != 0.0)) {
    // the lane's FP64 sum in fixed point: integer additions in any order
    // give the same H, so H does not depend on the order the units finish in
    unsigned long long *hp = reinterpret_cast<unsigned long long *>(a.H + 2 * ((size_t)i * a.n_dst + j));
    if (fabs(acc_re) < kHLimit && fabs(acc_im) < kHLimit) {
      atomicAdd(hp, (unsigned long long)__double2ll_rn(acc_re * kHScale));
      atomicAdd(hp + 1, (unsigned long long)__double2ll_rn(acc_im * kHScale));
    } else {
      atomicAdd(ctr_slot(a.counters) + 15, 1ull);  // refused on the host
    }
  }
  unsigned long long v[6] = {ct.nodes, ct.tris, ct.sph, ct.queries, nchain, nsucc};
  block_add_counters<6, kEvalBlock / 64>(a.counters, v);
}

// Backward legs are short bounded segments: the binary walk, whose loop fits
// the 80 registers of this kernel (the four-wide one spills at that budget).
template <bool LDS, bool EXPORT, bool BEAM = false, bool UNI = false>
__global__ void __launch_bounds__(kEvalBlock) __attribute__((amdgpu_waves_per_eu(BEAM ? AMS_BEAM_WAVES : 6, 8)))
k_evaluate(EvalArgs a) {
  evaluate_body<LDS, EXPORT, false, BEAM, UNI>(a);
}

// Scenes in global memory with flags bit 11: the four-wide walk, with the
// register budget it needs (4 waves per SIMD and up).
template <bool EXPORT>
__global__ void __launch_bounds__(kEvalBlock) __attribute__((amdgpu_waves_per_eu(4, 8))) k_evaluate_wide(EvalArgs a) {
  evaluate_body<false, EXPORT, true, false, false>(a);
}

// ---- beam candidate lists -------------------------------------------------
struct BeamArgs {
  int *beam;              // [unit][task in unit][Rx in tile][stride]
  int stride, slice, legs;
  float hx[8], hy[8], hz[8];  // corners of a box holding the Tx elements (the first 4 for a planar array)
  float margin0, margin_rel;  // candidate margin: margin0 + margin_rel * path length (m)
};

// The beam of one leg: NC start and NC end points, their extent along six
// axes (x, y, z and the leg direction crossed with each), with a margin.
template <int NC>
struct BeamLeg {
  float px[2 * NC], py[2 * NC], pz[2 * NC];
  float ax[6], ay[6], az[6];
  float lo[6], hi[6];
  float cx, cy, cz;  // leg direction (unit)
  float margin;
};

template <int NC>
__device__ __forceinline__ void beam_extent(const BeamLeg<NC> &B, float x, float y, float z, float &mn, float &mx) {
  mn = INFINITY;
  mx = -INFINITY;
#pragma unroll
  for (int m = 0; m < 2 * NC; m++) {
    const float v = B.px[m] * x + B.py[m] * y + B.pz[m] * z;
    mn = fminf(mn, v);
    mx = fmaxf(mx, v);
  }
}

template <int NC>
__device__ __forceinline__ void beam_axes(BeamLeg<NC> &B) {
  float sx = 0, sy = 0, sz = 0, ex = 0, ey = 0, ez = 0;
#pragma unroll
  for (int m = 0; m < NC; m++) {
    sx += B.px[m];
    sy += B.py[m];
    sz += B.pz[m];
    ex += B.px[NC + m];
    ey += B.py[NC + m];
    ez += B.pz[NC + m];
  }
  float cx = ex - sx, cy = ey - sy, cz = ez - sz;
  const float cl = sqrtf(cx * cx + cy * cy + cz * cz);
  if (cl > 0.0f) {
    cx /= cl;
    cy /= cl;
    cz /= cl;
  } else {
    cx = 1.0f;
    cy = 0.0f;
    cz = 0.0f;
  }
  B.cx = cx;
  B.cy = cy;
  B.cz = cz;
  // x, y, z, then c x (x, y, z); a near-zero cross product falls back to its axis
  const float e[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    B.ax[k] = e[k][0];
    B.ay[k] = e[k][1];
    B.az[k] = e[k][2];
    float vx = cy * e[k][2] - cz * e[k][1], vy = cz * e[k][0] - cx * e[k][2], vz = cx * e[k][1] - cy * e[k][0];
    const float vl = sqrtf(vx * vx + vy * vy + vz * vz);
    if (vl > 0.05f) {
      vx /= vl;
      vy /= vl;
      vz /= vl;
    } else {
      vx = e[k][0];
      vy = e[k][1];
      vz = e[k][2];
    }
    B.ax[3 + k] = vx;
    B.ay[3 + k] = vy;
    B.az[3 + k] = vz;
  }
#pragma unroll
  for (int k = 0; k < 6; k++) beam_extent(B, B.ax[k], B.ay[k], B.az[k], B.lo[k], B.hi[k]);
}

// A node box apart from the beam along one of the six axes (by more than the
// margin).  NaN compares false: a box is only ever rejected on a finite gap.
template <int NC>
__device__ __forceinline__ bool beam_box_apart(const BeamLeg<NC> &B, float4 lo, float4 hi) {
  const float cx = 0.5f * (lo.x + hi.x), cy = 0.5f * (lo.y + hi.y), cz = 0.5f * (lo.z + hi.z);
  const float hx = 0.5f * (hi.x - lo.x), hy = 0.5f * (hi.y - lo.y), hz = 0.5f * (hi.z - lo.z);
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const float c = cx * B.ax[k] + cy * B.ay[k] + cz * B.az[k];
    const float r = hx * fabsf(B.ax[k]) + hy * fabsf(B.ay[k]) + hz * fabsf(B.az[k]);
    if (c - r > B.hi[k] + B.margin || c + r < B.lo[k] - B.margin) return true;
  }
  return false;
}

// Triangle and beam apart along the unit axis (x, y, z), whose beam extent
// is [bl, bh].
template <int NC>
__device__ __forceinline__ bool tri_apart_along(const BeamLeg<NC> &B, float x, float y, float z, float bl, float bh,
                                                const float3 (&v)[3]) {
  const float t0 = v[0].x * x + v[0].y * y + v[0].z * z, t1 = v[1].x * x + v[1].y * y + v[1].z * z,
              t2 = v[2].x * x + v[2].y * y + v[2].z * z;
  const float tl = fminf(fminf(t0, t1), t2), th = fmaxf(fmaxf(t0, t1), t2);
  return tl > bh + B.margin || th < bl - B.margin;
}

// The same along an axis that is not one of the beam's six (normalised here;
// a degenerate axis separates nothing).
template <int NC>
__device__ __forceinline__ bool tri_apart_new(const BeamLeg<NC> &B, float x, float y, float z,
                                              const float3 (&v)[3]) {
  const float l = sqrtf(x * x + y * y + z * z);
  if (!(l > 1e-6f)) return false;
  x /= l;
  y /= l;
  z /= l;
  float bl, bh;
  beam_extent(B, x, y, z, bl, bh);
  return tri_apart_along(B, x, y, z, bl, bh, v);
}

// A triangle apart from the beam along a separating axis: the six beam axes,
// the triangle's normal, and the leg direction crossed with each edge.
// (A subset of the separating-axis test: a pair it cannot separate stays a
// candidate.)
template <int NC>
__device__ __forceinline__ bool beam_tri_apart(const BeamLeg<NC> &B, const AmsTri &tr) {
  const float3 v[3] = {make_float3(tr.a.x, tr.a.y, tr.a.z), make_float3(tr.b.x, tr.b.y, tr.b.z),
                       make_float3(tr.c.x, tr.c.y, tr.c.z)};
#pragma unroll
  for (int k = 0; k < 6; k++)
    if (tri_apart_along(B, B.ax[k], B.ay[k], B.az[k], B.lo[k], B.hi[k], v)) return true;
  const float3 e0 = v[1] - v[0], e1 = v[2] - v[1], e2 = v[0] - v[2];
  if (tri_apart_new(B, e0.y * e1.z - e0.z * e1.y, e0.z * e1.x - e0.x * e1.z, e0.x * e1.y - e0.y * e1.x, v))
    return true;
  const float3 es[3] = {e0, e1, e2};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float3 e = es[k];
    if (tri_apart_new(B, B.cy * e.z - B.cz * e.y, B.cz * e.x - B.cx * e.z, B.cx * e.y - B.cy * e.x, v)) return true;
  }
  return false;
}

struct BeamCount {
  unsigned long long pairs = 0, nodes = 0, tris = 0, walk = 0, cand = 0;
};

// The candidate lists of one (task, Rx) pair, walked from the Rx as
// k_evaluate walks them.  Leg l (l < nref) ends on the key's plane
// nref-1-l; its end points continue past the plane by the evaluation's
// query bound (0.1% + 1 mm); the last leg ends at the Tx corners.  Triangles
// of the plane a leg starts on are left out (the lane's query skips that
// plane).  The corners' images and crossings are formed in FP64 and kept in
// FP32 between legs (1e-6 relative, far inside the margin).  Returns the
// header.
template <int NC>
__device__ unsigned beam_pair(const EvalArgs &a, const BeamArgs &b, int p, int j, int *out, BeamCount &bc) {
  const unsigned walk_all = 0xFFFFFFFFu;
  const uint64_t khi = a.path_hi[p], klo = a.path_lo[p];
  const int h = (int)(khi >> 61), bm = (int)((khi >> 56) & 31);
  const int nref = __popc(bm & ((1 << h) - 1));
  if (h > AMS_MAX_HIT || nref != h || h + 1 > b.legs) return walk_all;  // transmissions: every leg walks
  bc.pairs++;
  const float rx = a.dst[3 * j], ry = a.dst[3 * j + 1], rz = a.dst[3 * j + 2];
  // per corner: the current start of its leg, its remaining image, the path length so far
  float ox[NC], oy[NC], oz[NC], ix[NC], iy[NC], iz[NC], len[NC];
#pragma unroll
  for (int m = 0; m < NC; m++) {
    double x = b.hx[m], y = b.hy[m], z = b.hz[m];
    for (int k = 0; k < nref; k++) {
      const AmsPlane P = a.planes[key_get_plane(khi, klo, k)];
      const double f = 2.0 * (x * P.n[0] + y * P.n[1] + z * P.n[2] - P.d);
      x -= f * P.n[0];
      y -= f * P.n[1];
      z -= f * P.n[2];
    }
    ix[m] = (float)x;
    iy[m] = (float)y;
    iz[m] = (float)z;
    ox[m] = rx;
    oy[m] = ry;
    oz[m] = rz;
    len[m] = 0.0f;
  }
  unsigned hdr = 0u;
  BeamLeg<NC> B;
  bool closed = true;  // once a leg's beam does not close, every later leg walks
  for (int l = 0; l <= nref; l++) {
    float lmax = 0.0f;
    AmsPlane P;
    if (l < nref) P = a.planes[key_get_plane(khi, klo, nref - 1 - l)];
#pragma unroll
    for (int m = 0; m < NC; m++) {
      B.px[m] = ox[m];
      B.py[m] = oy[m];
      B.pz[m] = oz[m];
      if (l < nref) {
        const double sx = ox[m], sy = oy[m], sz = oz[m];
        const double ux = ix[m] - sx, uy = iy[m] - sy, uz = iz[m] - sz;
        const double nd = ux * P.n[0] + uy * P.n[1] + uz * P.n[2];
        const double tp = (P.d - (sx * P.n[0] + sy * P.n[1] + sz * P.n[2])) / nd;
        if (!(tp > 0.0 && tp < 1.0)) closed = false;  // the plane is not between start and image
        const double ul = sqrt(ux * ux + uy * uy + uz * uz);
        const double dist = tp * ul;
        const double ext = tp + (1e-3 * dist + 1e-3) / ul;  // the evaluation's bound past the plane
        B.px[NC + m] = (float)(sx + ux * ext);
        B.py[NC + m] = (float)(sy + uy * ext);
        B.pz[NC + m] = (float)(sz + uz * ext);
        len[m] += (float)dist;
        // the image loses this plane's reflection; the next leg starts at the crossing
        const double f = 2.0 * ((double)ix[m] * P.n[0] + (double)iy[m] * P.n[1] + (double)iz[m] * P.n[2] - P.d);
        ix[m] = (float)(ix[m] - f * P.n[0]);
        iy[m] = (float)(iy[m] - f * P.n[1]);
        iz[m] = (float)(iz[m] - f * P.n[2]);
        ox[m] = (float)(sx + ux * tp);
        oy[m] = (float)(sy + uy * tp);
        oz[m] = (float)(sz + uz * tp);
      } else {
        B.px[NC + m] = b.hx[m];
        B.py[NC + m] = b.hy[m];
        B.pz[NC + m] = b.hz[m];
        const float ux = b.hx[m] - ox[m], uy = b.hy[m] - oy[m], uz = b.hz[m] - oz[m];
        len[m] += sqrtf(ux * ux + uy * uy + uz * uz);
      }
      lmax = fmaxf(lmax, len[m]);
    }
    unsigned nc = 0;
    if (closed) {
      B.margin = b.margin0 + b.margin_rel * lmax;
      beam_axes(B);
      const int skip_pl = l == 0 ? -1 : (int)key_get_plane(khi, klo, nref - l);
      int *cand = out + 1 + l * kBeamCand;
      int i = 0;
      while (i < a.n_tnodes && nc != kBeamWalk) {
        const float4 *np = reinterpret_cast<const float4 *>(a.tnodes + i);
        const float4 lo = np[0], hi = np[1];
        bc.nodes++;
        const int miss = __float_as_int(lo.w), info = __float_as_int(hi.w);
        if (!beam_box_apart(B, lo, hi)) {
          if (info == 0) {
            i++;
            continue;
          }
          const int first = info >> 3, cnt = info & 7;
          for (int k = 0; k < cnt; k++) {
            const AmsTri tr = a.tris[first + k];
            bc.tris++;
            if (__float_as_int(tr.b.w) == skip_pl || beam_tri_apart(B, tr)) continue;
            if (nc == (unsigned)kBeamCand) {
              nc = kBeamWalk;
              break;
            }
            cand[nc++] = first + k;
          }
        }
        i = miss;
      }
    } else {
      nc = kBeamWalk;
    }
    if (nc == kBeamWalk) bc.walk++;
    else bc.cand += nc;
    hdr |= nc << (kBeamBits * l);
  }
  for (int l = nref + 1; l * kBeamBits < 32; l++) hdr |= kBeamWalk << (kBeamBits * l);
  return hdr;
}

// One thread per (task, Rx in tile) of a launch's units (blockIdx.y = unit).
// NC: 4 box corners for a planar Tx array, 8 otherwise.
template <int NC>
__global__ void __launch_bounds__(kBeamBlock) __attribute__((amdgpu_waves_per_eu(NC == 4 ? 4 : 2, 8))) k_beam(EvalArgs a, BeamArgs b) {
  const int unit = blockIdx.y;
  const int tb = a.unit_begin[unit], te = a.unit_end[unit], g = a.unit_tile[unit];
  const int gs2 = a.gs * a.gs;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  BeamCount bc;
  if (idx < (te - tb) * gs2) {
    const int tl = idx / gs2, q = idx - tl * gs2;
    int *out = b.beam + (((size_t)unit * b.slice + tl) * gs2 + q) * b.stride;
    const int gi = g / a.n1g, gj = g % a.n1g;
    const int di = gi * a.gs + q / a.gs, dj = gj * a.gs + q % a.gs;
    const int j = di * a.n1_dst + dj;
    unsigned hdr = 0xFFFFFFFFu;
    if (j < a.n_dst && (dj < a.n1_dst || (a.flags & 1))) hdr = beam_pair<NC>(a, b, a.task_path[tb + tl], j, out, bc);
    out[0] = (int)hdr;
  }
  // counters 6..10: beam pairs, beam node tests, beam triangle tests, legs
  // left to the walk, candidates listed
  unsigned long long v[5] = {bc.pairs, bc.nodes, bc.tris, bc.walk, bc.cand};
#pragma unroll
  for (int k = 0; k < 5; k++) {
    unsigned long long x = v[k];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off);
    if (__lane_id() == 0 && x) atomicAdd(ctr_slot(a.counters) + 6 + k, x);
  }
}

// ---- neighbour expansion on the device (optixPathTracer.cpp:264-284) ----
// The traced (path, tile) pairs and the pairs new in the next round are two
// tile-major bit maps, bit g * n_paths + p.  A round's tasks are grouped by
// tile, paths ascending within a tile; tile_start / unit_start give each
// tile's tasks and its units of at most `slice` tasks.

// Every successful task (p, g) marks (p, g +- 1) and (p, g +- n1g) within the
// linear bounds (the reference's g +- 1 wraps across tile rows); a pair it is
// the first to mark goes into the new map.  round_nt: the round's task count.
__global__ void __launch_bounds__(256) k_expand(const int *__restrict__ task_path, const int *__restrict__ task_tile,
                                                const unsigned char *__restrict__ ok,
                                                const int *__restrict__ round_nt, int n_grid, int n1g, int n_paths,
                                                unsigned *__restrict__ traced, unsigned *__restrict__ fresh) {
  const int nt = *round_nt;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += gridDim.x * blockDim.x) {
    if (!ok[t]) continue;
    const int p = task_path[t], g = task_tile[t];
    const int nb[4] = {g - 1, g + 1, g - n1g, g + n1g};
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int h = nb[k];
      if (h < 0 || h >= n_grid) continue;
      const long long b = (long long)h * n_paths + p;
      const unsigned m = 1u << (b & 31);
      if (!(atomicOr(&traced[b >> 5], m) & m)) atomicOr(&fresh[b >> 5], m);
    }
  }
}

// the bits of tile g in word w of a tile-major map
__device__ __forceinline__ unsigned tile_word_mask(long long b0, long long b1, long long w) {
  const long long lo = b0 > 32 * w ? b0 - 32 * w : 0, hi = b1 < 32 * w + 32 ? b1 - 32 * w : 32;
  if (hi <= lo) return 0u;
  const unsigned top = hi >= 32 ? 0xffffffffu : ((1u << hi) - 1u);
  return top & ~((1u << lo) - 1u);
}

// Block-wide exclusive prefix of v over 256 threads; *total gets the sum.
__device__ __forceinline__ int block_excl_scan256(int v, int *sh, int *total) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int x = t >= off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += x;
    __syncthreads();
  }
  const int incl = sh[t];
  *total = sh[255];
  __syncthreads();
  return incl - v;
}

// New tasks of each tile: the population count of its bits of the new map.
// One workgroup per tile.
__global__ void __launch_bounds__(256) k_tile_count(const unsigned *__restrict__ fresh, int n_paths,
                                                    int *__restrict__ tile_cnt) {
  __shared__ int sh[256];
  const long long g = blockIdx.x, b0 = g * n_paths, b1 = b0 + n_paths;
  int c = 0;
  for (long long w = (b0 >> 5) + threadIdx.x; w <= ((b1 - 1) >> 5); w += 256)
    c += __popc(fresh[w] & tile_word_mask(b0, b1, w));
  int total;
  block_excl_scan256(c, sh, &total);
  if (threadIdx.x == 0) tile_cnt[g] = total;
}

// Exclusive prefix over the tiles (one workgroup of 256: a workgroup of
// 1024 waited up to 2 ms for a free CU while concurrent traces filled the
// GPU): tile_start and unit_start of the next round; info = {tasks, units}.
// Thread t sums a run of consecutive tiles, the runs are scanned per wave
// and across the four waves.
__global__ void __launch_bounds__(256) k_tile_scan(const int *__restrict__ tile_cnt, int n_grid, int slice,
                                                   int *__restrict__ tile_start, int *__restrict__ unit_start,
                                                   int *__restrict__ info) {
  __shared__ int wt[4], wu[4];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, chunk = (n_grid + 255) / 256;
  const int g0 = min(n_grid, t * chunk), g1 = min(n_grid, g0 + chunk);
  int ct = 0, cu = 0;
  for (int g = g0; g < g1; g++) {
    const int c = tile_cnt[g];
    ct += c;
    cu += (c + slice - 1) / slice;
  }
  int it = ct, iu = cu;  // inclusive scans within the wave
  for (int off = 1; off < 64; off <<= 1) {
    const int xt = __shfl_up(it, off), xu = __shfl_up(iu, off);
    if (lane >= off) {
      it += xt;
      iu += xu;
    }
  }
  if (lane == 63) {
    wt[w] = it;
    wu[w] = iu;
  }
  __syncthreads();
  int bt = 0, bu = 0;
  for (int q = 0; q < w; q++) {
    bt += wt[q];
    bu += wu[q];
  }
  int rt = bt + it - ct, ru = bu + iu - cu;
  for (int g = g0; g < g1; g++) {
    const int c = tile_cnt[g];
    tile_start[g] = rt;
    unit_start[g] = ru;
    rt += c;
    ru += (c + slice - 1) / slice;
  }
  if (t == 255) {
    tile_start[n_grid] = bt + it;
    unit_start[n_grid] = bu + iu;
    info[0] = bt + it;
    info[1] = bu + iu;
  }
}

// The next round's tasks of each tile from the new map, paths ascending;
// the tile's bits are cleared and the tasks' success flags zeroed.  One
// workgroup per tile.
__global__ void __launch_bounds__(256) k_tile_emit(unsigned *__restrict__ fresh, int n_paths,
                                                   const int *__restrict__ tile_start, int *__restrict__ task_path,
                                                   int *__restrict__ task_tile, unsigned char *__restrict__ ok) {
  __shared__ int sh[256];
  const long long g = blockIdx.x, b0 = g * n_paths, b1 = b0 + n_paths;
  const long long w_first = b0 >> 5, w_last = (b1 - 1) >> 5;
  int base = tile_start[g];
  for (long long wc = w_first; wc <= w_last; wc += 256) {
    const long long w = wc + threadIdx.x;
    unsigned bits = 0u;
    if (w <= w_last) {
      const unsigned m = tile_word_mask(b0, b1, w);
      bits = fresh[w] & m;
      if (bits) atomicAnd(&fresh[w], ~m);
    }
    int total;
    int pos = base + block_excl_scan256(__popc(bits), sh, &total);
    while (bits) {
      const int k = __ffs(bits) - 1;
      bits &= bits - 1u;
      task_path[pos] = (int)(32 * w + k - b0);
      task_tile[pos] = (int)g;
      ok[pos] = 0;
      pos++;
    }
    base += total;
  }
}

// The units of a round from its tile and unit offsets: unit u of tile g
// covers tasks tile_start[g] + slice (u - unit_start[g]) onwards, at most
// slice of them.  Thread per tile.
__global__ void __launch_bounds__(256) k_units(const int *__restrict__ tile_start, const int *__restrict__ unit_start,
                                               int n_grid, int slice, int *__restrict__ unit_tile,
                                               int *__restrict__ unit_begin, int *__restrict__ unit_end) {
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < n_grid; g += gridDim.x * blockDim.x) {
    const int t0 = tile_start[g], t1 = tile_start[g + 1];
    for (int u = unit_start[g]; u < unit_start[g + 1]; u++) {
      const int b = t0 + (u - unit_start[g]) * slice;
      unit_tile[u] = g;
      unit_begin[u] = b;
      unit_end[u] = min(b + slice, t1);
    }
  }
}

// Path keys of small scenes packed losslessly into 64 bits: every field
// keeps its order (plane ids < 2^pb - 1, the "no plane" value 0xFFFFF maps
// to 2^pb - 1, tiles < 2^tb), so a single keys-only radix sort over
// 8 + 5 pb + tb bits orders them as the 128-bit comparison does.
struct KeyBits {
  int pb, tb;  // bits per plane id, bits of the tile index; pb = 0: no packing
};
__device__ __forceinline__ uint64_t key_compact(uint64_t hi, uint64_t lo, KeyBits kb) {
  const uint32_t none = (1u << kb.pb) - 1u;
  uint64_t c = (hi >> 56);  // h:3 | b:5
  for (int k = 0; k < AMS_MAX_HIT; k++) {
    uint32_t p = key_get_plane(hi, lo, k);
    c = (c << kb.pb) | (p == 0xFFFFFu ? none : p);
  }
  return (c << kb.tb) | (lo & 0xFFFFFull);
}
__global__ void k_key_compact(const uint64_t *__restrict__ hi, const uint64_t *__restrict__ lo, size_t n, KeyBits kb,
                              uint64_t *__restrict__ out) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = key_compact(hi[k], lo[k], kb);
}
__global__ void k_key_expand(const uint64_t *__restrict__ c, size_t n, KeyBits kb, uint64_t *__restrict__ hi,
                             uint64_t *__restrict__ lo) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  uint64_t v = c[k];
  const uint32_t tile = (uint32_t)(v & ((1ull << kb.tb) - 1ull));
  v >>= kb.tb;
  const uint32_t none = (1u << kb.pb) - 1u;
  uint32_t pl[AMS_MAX_HIT];
  for (int q = AMS_MAX_HIT - 1; q >= 0; q--) {
    uint32_t p = (uint32_t)(v & none);
    pl[q] = p == none ? 0xFFFFFu : p;
    v >>= kb.pb;
  }
  uint64_t h, l;
  ams_key_pack((int)(v >> 5), (int)(v & 31), pl, tile, &h, &l);
  hi[k] = h;
  lo[k] = l;
}
__global__ void k_unique_flags64(const uint64_t *__restrict__ c, size_t n, unsigned char *__restrict__ flag) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) flag[k] = (k == 0) || c[k] != c[k - 1];
}

__global__ void k_add_h(const long long *__restrict__ Hfix, float *__restrict__ out, size_t n) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] += (float)((double)Hfix[k] * (1.0 / kHScale));
}

__global__ void k_unique_flags(const uint64_t *__restrict__ hi, const uint64_t *__restrict__ lo, size_t n,
                               unsigned char *__restrict__ flag) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) flag[k] = (k == 0) || hi[k] != hi[k - 1] || lo[k] != lo[k - 1];
}

}  // namespace

// ------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------
#define HIPCHK(x)                                                               \
  do {                                                                          \
    hipError_t e_ = (x);                                                        \
    if (e_ != hipSuccess) {                                                     \
      ams::set_error(std::string(#x) + ": " + hipGetErrorString(e_));          \
      return AMS_ERR_HIP;                                                       \
    }                                                                           \
  } while (0)

namespace ams {

static inline float bits_f(int x) {
  float f;
  std::memcpy(&f, &x, 4);
  return f;
}

int upload_scene(ams_scene *s, int n_tri, const float *tri, const int32_t *mat, int32_t n_mat,
                 const float *mat_prop, const float *refl_phase) {
  s->n_tri = n_tri;
  s->n_mat = n_mat;
  std::vector<int> plane(n_tri > 0 ? n_tri : 1);
  std::vector<double> pl_n(3 * (size_t)(n_tri > 0 ? n_tri : 1)), pl_d(n_tri > 0 ? n_tri : 1);
  std::vector<double> own(4 * (size_t)(n_tri > 0 ? n_tri : 1));
  if (n_tri > 0) ams_plane_ids(n_tri, tri, plane.data(), pl_n.data(), pl_d.data(), own.data());
  s->plane.assign(plane.begin(), plane.begin() + n_tri);
  std::vector<double> lo(3 * (size_t)n_tri), hi(3 * (size_t)n_tri);
  std::vector<AmsTriAux> aux(n_tri > 0 ? n_tri : 1);
  std::vector<AmsPlane> planes(n_tri > 0 ? n_tri : 1);
  for (int i = 0; i < n_tri; i++) {
    const float *t = tri + 9 * (size_t)i;
    for (int k = 0; k < 3; k++) {
      lo[3 * i + k] = std::min((double)t[k], std::min((double)t[3 + k], (double)t[6 + k]));
      hi[3 * i + k] = std::max((double)t[k], std::max((double)t[3 + k], (double)t[6 + k]));
    }
    // Unit normal normalize(cross(v1 - v0, v2 - v0)) (optixPathTracer.cu:299-300),
    // formed in FP64 from the f32 vertices and rounded to f32: the cross
    // product of a sliver triangle's nearly parallel edges loses most of its
    // digits in FP32, and its reflections would leave in another direction.
    double e1[3] = {(double)t[3] - t[0], (double)t[4] - t[1], (double)t[5] - t[2]};
    double e2[3] = {(double)t[6] - t[0], (double)t[7] - t[1], (double)t[8] - t[2]};
    double cd[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    double l = std::sqrt(cd[0] * cd[0] + cd[1] * cd[1] + cd[2] * cd[2]);
    float c[3] = {0.0f, 0.0f, 1.0f};
    if (l > 0)
      for (int k = 0; k < 3; k++) c[k] = (float)(cd[k] / l);
    auto el = [&](int a0, int b0) {
      float dx = t[3 * a0] - t[3 * b0], dy = t[3 * a0 + 1] - t[3 * b0 + 1], dz = t[3 * a0 + 2] - t[3 * b0 + 2];
      return std::sqrt(dx * dx + dy * dy + dz * dz);
    };
    float me = std::max(std::max(el(0, 1), el(1, 2)), el(2, 0));
    // A triangle whose own plane is not bitwise the plane of its plane id (a
    // nearly coplanar triangle grouped with others) carries a negative
    // longest edge (the kernels use its square): a chain reflecting on it
    // takes its phase length from its own plane.
    bool inexact = own[4 * (size_t)i] != pl_n[3 * i] || own[4 * (size_t)i + 1] != pl_n[3 * i + 1] ||
                   own[4 * (size_t)i + 2] != pl_n[3 * i + 2] || own[4 * (size_t)i + 3] != pl_d[i];
    s->n_inexact += inexact;
    aux[i].n_maxedge = make_float4(c[0], c[1], c[2], inexact ? -me : me);
    planes[i].n[0] = pl_n[3 * i];
    planes[i].n[1] = pl_n[3 * i + 1];
    planes[i].n[2] = pl_n[3 * i + 2];
    planes[i].d = pl_d[i];
  }
  // bounds of every plane id's triangles, for the evaluation's path pre-check
  std::vector<double> pb(6 * (size_t)std::max(n_tri, 1));
  for (int i = 0; i < n_tri; i++) {
    double *b = &pb[6 * (size_t)s->plane[i]];
    if (s->plane[i] == i) {
      for (int k = 0; k < 3; k++) {
        b[k] = INFINITY;
        b[3 + k] = -INFINITY;
      }
    }
  }
  for (int i = 0; i < n_tri; i++) {
    double *b = &pb[6 * (size_t)s->plane[i]];
    for (int k = 0; k < 3; k++) {
      b[k] = std::min(b[k], lo[3 * i + k]);
      b[3 + k] = std::max(b[3 + k], hi[3 * i + k]);
    }
  }
  std::vector<AmsPlaneBox> pbox(std::max(n_tri, 1));
  for (int i = 0; i < n_tri; i++) {
    const double *b = &pb[6 * (size_t)i];
    if (s->plane[i] != i) continue;
    pbox[i].lo = make_float4(std::nextafter((float)b[0], -INFINITY), std::nextafter((float)b[1], -INFINITY),
                             std::nextafter((float)b[2], -INFINITY), 0.0f);
    pbox[i].hi = make_float4(std::nextafter((float)b[3], INFINITY), std::nextafter((float)b[4], INFINITY),
                             std::nextafter((float)b[5], INFINITY), 0.0f);
  }
  AmsBvhBuild B;
  ams_bvh_build(n_tri, lo.data(), hi.data(), AMS_LEAF_MAX, &B);
  std::vector<AmsTri> tris(n_tri > 0 ? n_tri : 1);
  for (int s2 = 0; s2 < n_tri; s2++) {
    int i = B.order[s2];
    const float *t = tri + 9 * (size_t)i;
    AmsTri r;
    r.a = make_float4(t[0], t[1], t[2], bits_f(i));
    r.b = make_float4(t[3], t[4], t[5], bits_f(plane[i]));
    r.c = make_float4(t[6], t[7], t[8], bits_f(mat[i]));
    tris[s2] = r;
  }
  if (B.nodes.empty()) {
    AmsNode n;  // an empty tree: one leaf that no ray enters
    for (int k = 0; k < 3; k++) {
      n.lo[k] = 1.0f;
      n.hi[k] = -1.0f;
    }
    n.miss = 1;
    n.info = 0;
    B.nodes.push_back(n);
  }
  s->n_tnodes = (int)B.nodes.size();
  std::vector<AmsNode4> N4;
  s->depth4 = ams_bvh_collapse4(B, &N4);
  s->n_tnodes4 = (int)N4.size();
  std::vector<AmsPlane> tplane(std::max(n_tri, 1));
  for (int i = 0; i < n_tri; i++) {
    for (int k = 0; k < 3; k++) tplane[i].n[k] = own[4 * (size_t)i + k];
    tplane[i].d = own[4 * (size_t)i + 3];
  }
  std::vector<float4> planes_f(planes.size());
  for (size_t k = 0; k < planes.size(); k++)
    planes_f[k] = make_float4((float)planes[k].n[0], (float)planes[k].n[1], (float)planes[k].n[2], (float)planes[k].d);
  std::vector<float> mp(19 * (size_t)std::max(n_mat, 1), 0.0f), rp(std::max(n_mat, 1), 0.0f);
  for (int k = 0; k < 19 * n_mat; k++) mp[k] = mat_prop[k];
  for (int k = 0; k < n_mat; k++) rp[k] = refl_phase ? refl_phase[k] : 0.0f;
  if (dev_upload(&s->d_tnodes, B.nodes.data(), sizeof(AmsNode) * B.nodes.size()) ||
      dev_upload(&s->d_tnodes4, N4.data(), sizeof(AmsNode4) * N4.size()) ||
      dev_upload(&s->d_tris, tris.data(), sizeof(AmsTri) * tris.size()) ||
      dev_upload(&s->d_aux, aux.data(), sizeof(AmsTriAux) * aux.size()) ||
      dev_upload(&s->d_planes, planes.data(), sizeof(AmsPlane) * planes.size()) ||
      dev_upload(&s->d_planes_f, planes_f.data(), sizeof(float4) * planes_f.size()) ||
      dev_upload(&s->d_pbox, pbox.data(), sizeof(AmsPlaneBox) * pbox.size()) ||
      dev_upload(&s->d_tplane, tplane.data(), sizeof(AmsPlane) * tplane.size()) ||
      dev_upload(&s->d_mat_prop, mp.data(), sizeof(float) * mp.size()) ||
      dev_upload(&s->d_refl_phase, rp.data(), sizeof(float) * rp.size()))
    return AMS_ERR_HIP;
  return AMS_OK;
}

// (hi, lo) sort + unique in place on device; returns the unique count.
// Small device-to-host read-back through the scene's page-locked buffer (a
// DMA copy: it does not wait for compute units held by other streams'
// kernels), then a wait on the stream.
static hipError_t read_back(Scratch &sc, void *dst, const void *src, size_t bytes, hipStream_t st) {
  void *h = sc.host(bytes);
  if (!h) return hipErrorOutOfMemory;
  hipError_t e = hipMemcpyAsync(h, src, bytes, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e == hipSuccess) std::memcpy(dst, h, bytes);
  return e;
}

// The launch counters summed over their slots.
static hipError_t read_counters(Scratch &sc, unsigned long long *hctr, const unsigned long long *d_ctr,
                                hipStream_t st) {
  std::vector<unsigned long long> all((size_t)kCtrN * kCtrSlots);
  hipError_t e = read_back(sc, all.data(), d_ctr, sizeof(unsigned long long) * all.size(), st);
  for (int k = 0; k < kCtrN; k++) {
    hctr[k] = 0;
    for (int s = 0; s < kCtrSlots; s++) hctr[k] += all[(size_t)s * kCtrN + k];
  }
  return e;
}

static int sort_unique(Scratch &sc, uint64_t *hi, uint64_t *lo, size_t n, hipStream_t st, size_t *n_out,
                       KeyBits kb = KeyBits{0, 0}) {
  if (n == 0) {
    *n_out = 0;
    return AMS_OK;
  }
  if (kb.pb > 0) {
    // packed keys: one keys-only sort over the used bits, unique, unpack
    const int bits = 8 + AMS_MAX_HIT * kb.pb + kb.tb;
    uint64_t *c1 = (uint64_t *)sc.get("su_c1", n * 8), *c2 = (uint64_t *)sc.get("su_c2", n * 8);
    unsigned char *flag = (unsigned char *)sc.get("su_flag", n);
    size_t *cnt = (size_t *)sc.get("su_cnt", 16);
    if (!c1 || !c2 || !flag || !cnt) return AMS_ERR_ALLOC;
    const unsigned g = (unsigned)((n + 255) / 256);
    k_key_compact<<<g, 256, 0, st>>>(hi, lo, n, kb, c1);
    HIPCHK(hipGetLastError());
    size_t tb = 0, tb3 = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, c1, c2, (int)n, 0, bits, st));
    HIPCHK(hipcub::DeviceSelect::Flagged(nullptr, tb3, c2, flag, c1, cnt, (int)n, st));
    void *tmp = sc.get("su_tmp", std::max(tb, tb3));
    if (!tmp) return AMS_ERR_ALLOC;
    HIPCHK(hipcub::DeviceRadixSort::SortKeys(tmp, tb, c1, c2, (int)n, 0, bits, st));
    k_unique_flags64<<<g, 256, 0, st>>>(c2, n, flag);
    HIPCHK(hipGetLastError());
    HIPCHK(hipcub::DeviceSelect::Flagged(tmp, tb3, c2, flag, c1, cnt, (int)n, st));
    size_t m = 0;
    HIPCHK(read_back(sc, &m, cnt, sizeof(m), st));
    if (m) k_key_expand<<<(unsigned)((m + 255) / 256), 256, 0, st>>>(c1, m, kb, hi, lo);
    HIPCHK(hipGetLastError());
    *n_out = m;
    return AMS_OK;
  }
  uint64_t *hi2 = (uint64_t *)sc.get("su_hi2", n * 8), *lo2 = (uint64_t *)sc.get("su_lo2", n * 8);
  unsigned char *flag = (unsigned char *)sc.get("su_flag", n);
  size_t *cnt = (size_t *)sc.get("su_cnt", 16);
  if (!hi2 || !lo2 || !flag || !cnt) return AMS_ERR_ALLOC;
  size_t tb = 0, tb2 = 0, tb3 = 0;
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, lo, lo2, hi, hi2, (int)n, 0, 64, st));
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb2, hi2, hi, lo2, lo, (int)n, 0, 64, st));
  HIPCHK(hipcub::DeviceSelect::Flagged(nullptr, tb3, hi2, flag, hi, cnt, (int)n, st));
  size_t tbm = std::max(tb, std::max(tb2, tb3));
  void *tmp = sc.get("su_tmp", tbm);
  if (!tmp) return AMS_ERR_ALLOC;
  // pass 1: by lo (values hi); pass 2: stable by hi (values lo)
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, lo, lo2, hi, hi2, (int)n, 0, 64, st));
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(tmp, tb2, hi2, hi, lo2, lo, (int)n, 0, 64, st));
  k_unique_flags<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(hi, lo, n, flag);
  HIPCHK(hipGetLastError());
  HIPCHK(hipcub::DeviceSelect::Flagged(tmp, tb3, hi, flag, hi2, cnt, (int)n, st));
  HIPCHK(hipcub::DeviceSelect::Flagged(tmp, tb3, lo, flag, lo2, cnt + 1, (int)n, st));
  size_t hc[2];
  HIPCHK(read_back(sc, hc, cnt, sizeof(hc), st));
  HIPCHK(hipMemcpyAsync(hi, hi2, hc[0] * 8, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(lo, lo2, hc[0] * 8, hipMemcpyDeviceToDevice, st));
  *n_out = hc[0];
  return AMS_OK;
}

static double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Corners of a box that holds every Tx element: the box's axes follow the
// array's first row and first column (orthonormalised, FP64), or x, y, z for
// a single element or a line.  Every element is a convex combination of the
// eight corners.  Returns true for a planar array (extent below 1 um across
// the row and column): its four corners at mid-depth come first and hold
// every element.
static bool tx_box_corners(const ams_array *tx, float hx[8], float hy[8], float hz[8]) {
  const int n0 = tx->n0, n1 = tx->n1, n = n0 * n1;
  const float *P = tx->pos;
  double ax[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  auto vec = [&](int i, int k, double v[3]) {
    for (int c = 0; c < 3; c++) v[c] = (double)P[3 * k + c] - (double)P[3 * i + c];
  };
  auto len = [](const double v[3]) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); };
  if (n0 > 1 && n1 > 1) {
    double u[3], v[3];
    vec(0, n1 - 1, u);
    vec(0, (n0 - 1) * n1, v);
    const double lu = len(u);
    if (lu > 0) {
      for (int c = 0; c < 3; c++) u[c] /= lu;
      const double d = u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
      for (int c = 0; c < 3; c++) v[c] -= d * u[c];
      const double lv = len(v);
      if (lv > 1e-9 * lu) {
        for (int c = 0; c < 3; c++) v[c] /= lv;
        const double w[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
        for (int c = 0; c < 3; c++) {
          ax[0][c] = u[c];
          ax[1][c] = v[c];
          ax[2][c] = w[c];
        }
      }
    }
  }
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = 0; i < n; i++) {
    for (int k = 0; k < 3; k++) {
      const double s = ax[k][0] * P[3 * i] + ax[k][1] * P[3 * i + 1] + ax[k][2] * P[3 * i + 2];
      lo[k] = std::min(lo[k], s);
      hi[k] = std::max(hi[k], s);
    }
  }
  const bool planar = hi[2] - lo[2] < 1e-6;
  for (int m = 0; m < 8; m++) {
    const double c0 = (m & 1) ? hi[0] : lo[0], c1 = (m & 2) ? hi[1] : lo[1];
    const double c2 = planar ? 0.5 * (lo[2] + hi[2]) : ((m & 4) ? hi[2] : lo[2]);
    hx[m] = (float)(c0 * ax[0][0] + c1 * ax[1][0] + c2 * ax[2][0]);
    hy[m] = (float)(c0 * ax[0][1] + c1 * ax[1][1] + c2 * ax[2][1]);
    hz[m] = (float)(c0 * ax[0][2] + c1 * ax[1][2] + c2 * ax[2][2]);
  }
  return planar;
}

// Fit c0 + i u + j v to an Rx array (u, v from its corner elements, FP64)
// and return the lattice query's parameters when every element lies within
// a quarter of the line spacing of its lattice point.
static RxLattice rx_lattice(const ams_array *rx, float acc, bool allow) {
  RxLattice L;
  memset(&L, 0, sizeof L);
  const int n0 = rx->n0, n1 = rx->n1;
  if (!allow || n0 < 2 || n1 < 2) return L;
  auto P = [&](int i, int j, int k) { return (double)rx->pos[3 * ((size_t)i * n1 + j) + k]; };
  double c0[3], u[3], v[3];
  for (int k = 0; k < 3; k++) {
    c0[k] = P(0, 0, k);
    u[k] = (P(n0 - 1, 0, k) - c0[k]) / (n0 - 1);
    v[k] = (P(0, n1 - 1, k) - c0[k]) / (n1 - 1);
  }
  double cr[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
  double area = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
  double lu = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]), lv = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (!(area > 1e-12) || !(lu > 0) || !(lv > 0)) return L;
  double hi = area / lv, hj = area / lu;  // spacing of the lines of constant i, of constant j
  double delta = 0.0, cmax = 0.0;
  for (int i = 0; i < n0; i++)
    for (int j = 0; j < n1; j++) {
      double e2 = 0.0;
      for (int k = 0; k < 3; k++) {
        double e = P(i, j, k) - (c0[k] + i * u[k] + j * v[k]);
        e2 += e * e;
        cmax = std::max(cmax, std::fabs(P(i, j, k)));
      }
      delta = std::max(delta, std::sqrt(e2));
    }
  if (!(delta <= 0.25 * std::min(hi, hj))) return L;
  double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2], vv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  double uv = u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
  double det = uu * vv - uv * uv;
  double gi[3], gj[3];
  for (int k = 0; k < 3; k++) {
    gi[k] = (vv * u[k] - uv * v[k]) / det;
    gj[k] = (uu * v[k] - uv * u[k]) / det;
  }
  // f32 evaluation of the slab and the lattice coordinates errs by far less
  // than the 0.1 mm and 0.01-cell allowances added here
  double reach = (double)acc + delta + 1e-4 + 1e-6 * cmax;
  L.on = 1;
  L.n0 = n0;
  L.n1 = n1;
  L.c0 = make_float3((float)c0[0], (float)c0[1], (float)c0[2]);
  L.n = make_float3((float)(cr[0] / area), (float)(cr[1] / area), (float)(cr[2] / area));
  L.gi = make_float3((float)gi[0], (float)gi[1], (float)gi[2]);
  L.gj = make_float3((float)gj[0], (float)gj[1], (float)gj[2]);
  L.W = (float)reach;
  L.Mi = (float)(reach / hi + 0.01);
  L.Mj = (float)(reach / hj + 0.01);
  return L;
}

int trace_channel(ams_scene *s, const ams_array *tx, const ams_array *rx, const ams_trace_opts *o, float *H_out,
                  ams_trace_stats *stats) {
  auto t_start = std::chrono::steady_clock::now();
  hipStream_t st = (hipStream_t)o->stream;
  const int n_src = tx->n0 * tx->n1, n_dst = rx->n0 * rx->n1;
  const int gs = o->gridsize, K = o->max_reflection;
  const int n1g = (rx->n1 - 1) / gs + 1, n0g = (rx->n0 - 1) / gs + 1;
  const long long n_grid = (long long)n0g * n1g;
  Scratch &sc = s->scratch;
  ams_trace_stats S;
  memset(&S, 0, sizeof S);
  bool beams_used = false;  // backward queries answered from beam candidate lists

  // Tx / Rx positions and the Rx sphere hierarchy
  float *d_src = (float *)sc.get("src", sizeof(float) * 3 * n_src);
  float *d_dst = (float *)sc.get("dst", sizeof(float) * 3 * n_dst);
  if (!d_src || !d_dst) return AMS_ERR_ALLOC;
  HIPCHK(hipMemcpyAsync(d_src, tx->pos, sizeof(float) * 3 * n_src, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d_dst, rx->pos, sizeof(float) * 3 * n_dst, hipMemcpyHostToDevice, st));
  const float dst_acc = std::max(rx->radius, AMS_ACC_MIN);
  // regular Rx lattice: sphere candidates found without the hierarchy (flags bit 5 keeps it)
  RxLattice lat = rx_lattice(rx, dst_acc, !(o->flags & 32));
  // The sphere hierarchy is built only for arrays that are not a lattice (a
  // host build of 4 ms for 4096 Rx and 84 ms for 65,536, on every call);
  // otherwise one node that no ray enters stands in for it.
  AmsBvhBuild SB;
  std::vector<float4> sph(std::max(n_dst, 1), make_float4(0.0f, 0.0f, 0.0f, 0.0f));
  if (!lat.on) {
    std::vector<double> slo(3 * (size_t)n_dst), shi(3 * (size_t)n_dst);
    for (int j = 0; j < n_dst; j++)
      for (int k = 0; k < 3; k++) {
        slo[3 * j + k] = (double)rx->pos[3 * j + k] - dst_acc;
        shi[3 * j + k] = (double)rx->pos[3 * j + k] + dst_acc;
      }
    ams_bvh_build(n_dst, slo.data(), shi.data(), AMS_LEAF_MAX, &SB);
    for (int q = 0; q < n_dst; q++) {
      int j = SB.order[q];
      sph[q] = make_float4(rx->pos[3 * j], rx->pos[3 * j + 1], rx->pos[3 * j + 2], bits_f(j));
    }
  }
  if (SB.nodes.empty()) {
    AmsNode n;
    for (int k = 0; k < 3; k++) {
      n.lo[k] = 1.0f;
      n.hi[k] = -1.0f;
    }
    n.miss = 1;
    n.info = 0;
    SB.nodes.push_back(n);
  }
  AmsNode *d_snodes = (AmsNode *)sc.get("snodes", sizeof(AmsNode) * SB.nodes.size());
  float4 *d_sph = (float4 *)sc.get("sph", sizeof(float4) * sph.size());
  unsigned long long *d_ctr = (unsigned long long *)sc.get("ctr", sizeof(unsigned long long) * kCtrN * kCtrSlots);
  if (!d_snodes || !d_sph || !d_ctr) return AMS_ERR_ALLOC;
  HIPCHK(hipMemcpyAsync(d_snodes, SB.nodes.data(), sizeof(AmsNode) * SB.nodes.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d_sph, sph.data(), sizeof(float4) * sph.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(d_ctr, 0, sizeof(unsigned long long) * kCtrN * kCtrSlots, st));
  float4 *d_sph_o = nullptr;
  if (lat.on) {
    std::vector<float4> sph_o(n_dst);
    for (int j = 0; j < n_dst; j++)
      sph_o[j] = make_float4(rx->pos[3 * j], rx->pos[3 * j + 1], rx->pos[3 * j + 2], bits_f(j));
    d_sph_o = (float4 *)sc.get("sph_o", sizeof(float4) * n_dst);
    if (!d_sph_o) return AMS_ERR_ALLOC;
    HIPCHK(hipMemcpyAsync(d_sph_o, sph_o.data(), sizeof(float4) * n_dst, hipMemcpyHostToDevice, st));
  }

  // key packing for the record sorts: plane ids and tiles of small scenes fit in 64 bits
  KeyBits kbits{0, 0};
  {
    int pb = 1, tb = 1;
    while (((1ll << pb) - 1) < (long long)std::max(1, s->n_tri) + 1) pb++;
    while ((1ll << tb) < n_grid) tb++;
    if (8 + AMS_MAX_HIT * pb + tb <= 64 && !(o->flags & 512)) kbits = KeyBits{pb, tb};
  }

  // ---------------- forward discovery in chunks of rays -----------------
  const long long n_rays = o->discovery_mode == 1
                               ? (long long)o->launch_x * o->launch_y * o->samples_per_launch
                               : (o->per_element ? (long long)n_src * o->fib_n : o->fib_n);
  // one launch covers up to 2^24 rays; the per-workgroup record filter keeps
  // the raw records near 0.3 per ray, and a launch whose records overflow
  // the 4-per-ray buffer (1 GB at 2^24 rays) is split and redone
  const long long chunk = std::min<long long>(n_rays, 1LL << 24);
  const unsigned long long cap = (unsigned long long)chunk * 4 + 1024;
  uint64_t *c_hi = (uint64_t *)sc.get("c_hi", cap * 8), *c_lo = (uint64_t *)sc.get("c_lo", cap * 8);
  unsigned long long *d_cnt = (unsigned long long *)sc.get("c_cnt", 8);
  unsigned long long *d_next = (unsigned long long *)sc.get("c_next", 8);
  if (!c_hi || !c_lo || !d_cnt || !d_next) return AMS_ERR_ALLOC;
  int dev = 0, n_cu = 256;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
  std::vector<uint64_t> acc_hi, acc_lo;  // running unique set (host)
  DiscoverArgs da;
  da.tnodes = s->d_tnodes;
  da.n_tnodes = s->n_tnodes;
  // Scene hierarchy of the discovery walk.  Scenes that fit in LDS (flags
  // bit 4 keeps them in global memory) walk the binary hierarchy there: its
  // node loads are cheap and its loop is the shorter one.  Larger scenes walk
  // the four-wide hierarchy, a quarter as many dependent L2 loads.  Bit 10
  // forces the binary walk, bit 11 the four-wide one.
  const size_t eval_scene_lds = (size_t)s->n_tnodes * sizeof(AmsNode) + (size_t)s->n_tri * sizeof(AmsTri);
  const bool eval_lds = eval_scene_lds <= (size_t)kSceneLds && !(o->flags & 16);
  const bool wide = s->d_tnodes4 && !(o->flags & 1024) && ((o->flags & 2048) || !eval_lds);
  da.tnodes4 = wide ? s->d_tnodes4 : nullptr;
  da.n_tnodes4 = wide ? s->n_tnodes4 : 0;
  da.tris = s->d_tris;
  da.n_tris = s->n_tri;
  const size_t scene_lds = (wide ? (size_t)s->n_tnodes4 * sizeof(AmsNode4) : (size_t)s->n_tnodes * sizeof(AmsNode)) +
                           (size_t)s->n_tri * sizeof(AmsTri);
  const bool use_lds = scene_lds <= (size_t)kSceneLds && !(o->flags & 16);
  const unsigned lds_bytes = use_lds ? (unsigned)scene_lds : 0u;
  da.aux = s->d_aux;
  da.mat_prop = s->d_mat_prop;
  da.snodes = d_snodes;
  da.n_snodes = (int)SB.nodes.size();
  da.sph = d_sph;
  da.sph_o = d_sph_o;
  da.lat = lat;
  da.src = d_src;
  da.n_src = n_src;
  da.n1_dst = rx->n1;
  da.gs = gs;
  da.acc2 = dst_acc * dst_acc;
  da.r2 = rx->radius * rx->radius;
  da.K = K;
  da.fib_n = o->fib_n;
  da.per_element = o->per_element;
  da.flags = o->flags;
  da.out_hi = c_hi;
  da.out_lo = c_lo;
  da.out_count = d_cnt;
  da.cap = cap;
  da.counters = d_ctr;
  // coherent ray order: a permutation of the Fibonacci indices, cached per fib_n
  if (!o->seed_keys && o->discovery_mode == 0 && s->perm_n != o->fib_n) {
    long long n = o->fib_n;
    unsigned *keys = (unsigned *)sc.get("perm_keys", 4 * (size_t)n), *idx = (unsigned *)sc.get("perm_idx", 4 * (size_t)n);
    unsigned *k2 = (unsigned *)sc.get("perm_k2", 4 * (size_t)n), *perm = (unsigned *)sc.get("perm", 4 * (size_t)n);
    if (!keys || !idx || !k2 || !perm) return AMS_ERR_ALLOC;
    k_fib_keys<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, keys, idx);
    HIPCHK(hipGetLastError());
    size_t tb = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys, k2, idx, perm, (int)n, 0, 32, st));
    void *tmp = sc.get("perm_tmp", tb);
    if (!tmp) return AMS_ERR_ALLOC;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, keys, k2, idx, perm, (int)n, 0, 32, st));
    float4 *dirs = (float4 *)sc.get("dirs", sizeof(float4) * (size_t)n);
    if (!dirs) return AMS_ERR_ALLOC;
    k_fib_dirs<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, perm, dirs);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    s->perm_n = n;
  }
  da.perm = (const unsigned *)sc.get("perm", 4 * (size_t)o->fib_n);
  da.dirs = (const float4 *)sc.get("dirs", sizeof(float4) * (size_t)std::max<long long>(1, o->fib_n));
  // the timing events are destroyed on every return path
  struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    ~EventPair() {
      if (a) (void)hipEventDestroy(a);
      if (b) (void)hipEventDestroy(b);
    }
  } ev;
  HIPCHK(hipEventCreate(&ev.a));
  HIPCHK(hipEventCreate(&ev.b));
  hipEvent_t e0 = ev.a, e1 = ev.b;
  float disc_ms = 0.0f;
  size_t n_unique_total = 0;
  std::vector<uint64_t> uh, ul;
  if (o->seed_keys) {
    uh.assign(o->n_seed, 0);
    ul.assign(o->n_seed, 0);
    for (long long k = 0; k < o->n_seed; k++) {
      uh[k] = o->seed_keys[2 * k];
      ul[k] = o->seed_keys[2 * k + 1];
      if ((ul[k] & 0xFFFFF) >= (uint64_t)n_grid) {
        set_error("seed key tile out of range");
        return AMS_ERR_ARG;
      }
      int h = (int)(uh[k] >> 61);
      if (h > K) {
        set_error("seed key hit count exceeds max_reflection");
        return AMS_ERR_ARG;
      }
      uint32_t pl[5], tile;
      int hh, bb;
      ams_key_unpack(uh[k], ul[k], &hh, &bb, pl, &tile);
      int nref = __builtin_popcount((unsigned)bb & ((1u << hh) - 1u));
      for (int q = 0; q < nref; q++) {
        if (pl[q] >= (uint32_t)s->n_tri) {
          set_error("seed key: a reflection has no valid plane id");
          return AMS_ERR_ARG;
        }
      }
    }
  }
  // Running unique set on the device (acc_hi/acc_lo); each chunk's records
  // are sorted and deduplicated, appended, and the set is compacted when it
  // outgrows its buffer.  A chunk whose records overflow is split and redone
  // (the kernel is deterministic, so the split gives the same records).
  size_t acc_n = 0, acc_cap = 1 << 20;
  // the running set is one sorted, duplicate-free run (one launch's unique
  // records, or a compacted set): its final sort + unique is then skipped
  bool acc_sorted = true;
  uint64_t *a_hi = (uint64_t *)sc.get("acc_hi", acc_cap * 8), *a_lo = (uint64_t *)sc.get("acc_lo", acc_cap * 8);
  if (!a_hi || !a_lo) return AMS_ERR_ALLOC;
  long long cur_chunk = chunk;
  if (o->discovery_mode == 1 && !o->seed_keys) {
    // reference-compatible rays: one launch over launch_x x launch_y threads
    long long nthr = (long long)o->launch_x * o->launch_y;
    unsigned long long rcap = (unsigned long long)nthr * (unsigned long long)std::max(0, o->reserve_size) + 1024;
    int *d_minhc = (int *)sc.get("min_hc", sizeof(int) * (size_t)nthr * (size_t)n_grid);
    uint64_t *r_hi = (uint64_t *)sc.get("r_hi", rcap * 8), *r_lo = (uint64_t *)sc.get("r_lo", rcap * 8);
    if (!d_minhc || !r_hi || !r_lo) return AMS_ERR_ALLOC;
    HIPCHK(hipMemsetAsync(d_minhc, 0, sizeof(int) * (size_t)nthr * (size_t)n_grid, st));
    HIPCHK(hipMemsetAsync(d_cnt, 0, 8, st));
    RefArgs ra;
    ra.d = da;
    ra.d.out_hi = r_hi;
    ra.d.out_lo = r_lo;
    ra.d.cap = rcap;
    ra.launch_x = o->launch_x;
    ra.launch_y = o->launch_y;
    ra.samples = o->samples_per_launch;
    ra.reserve = o->reserve_size;
    ra.n_grid = (int)n_grid;
    ra.min_hc = d_minhc;
    HIPCHK(hipEventRecord(e0, st));
    const unsigned rblocks = (unsigned)((nthr + 255) / 256);
    if (use_lds && wide) k_discover_ref<true, true><<<rblocks, 256, lds_bytes, st>>>(ra);
    else if (use_lds) k_discover_ref<true, false><<<rblocks, 256, lds_bytes, st>>>(ra);
    else if (wide) k_discover_ref<false, true><<<rblocks, 256, 0, st>>>(ra);
    else k_discover_ref<false, false><<<rblocks, 256, 0, st>>>(ra);
    HIPCHK(hipGetLastError());
    S.n_discover_launches++;
    HIPCHK(hipEventRecord(e1, st));
    unsigned long long hcnt = 0;
    HIPCHK(read_back(sc, &hcnt, d_cnt, 8, st));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    disc_ms += ms;
    if (hcnt > rcap) {
      set_error("reference-mode records exceed launch_x * launch_y * reserve_size");
      return AMS_ERR_OVERFLOW;
    }
    size_t nu = 0;
    int rc = sort_unique(sc, r_hi, r_lo, (size_t)hcnt, st, &nu, kbits);
    if (rc) return rc;
    S.records_raw += (long long)hcnt;
    if (nu > acc_cap) {
      uint64_t *nh = (uint64_t *)dmalloc(nu * 8), *nl = (uint64_t *)dmalloc(nu * 8);
      if (!nh || !nl) {
        dfree(nh);
        dfree(nl);
        set_error("device allocation failed: record set");
        return AMS_ERR_ALLOC;
      }
      sc.adopt("acc_hi", nh, nu * 8);
      sc.adopt("acc_lo", nl, nu * 8);
      a_hi = nh;
      a_lo = nl;
      acc_cap = nu;
    }
    HIPCHK(hipMemcpyAsync(a_hi, r_hi, nu * 8, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(a_lo, r_lo, nu * 8, hipMemcpyDeviceToDevice, st));
    acc_n = nu;
  }
  for (long long rb = (o->seed_keys || o->discovery_mode == 1) ? n_rays : 0; rb < n_rays;) {
    long long re = std::min(n_rays, rb + cur_chunk);
    HIPCHK(hipMemsetAsync(d_cnt, 0, 8, st));
    da.ray_begin = rb;
    da.ray_end = re;
    unsigned blocks = (unsigned)((re - rb + 255) / 256);
    const long long per_block = wide ? (long long)kDiscWideBlock * kDiscWideRays : (long long)kDiscBlock * kDiscRays;
    const unsigned dblocks = (unsigned)((re - rb + per_block - 1) / per_block);
    HIPCHK(hipEventRecord(e0, st));
    if (o->flags & 4) {
      HIPCHK(hipMemsetAsync(d_next, 0, 8, st));
      unsigned pblocks = std::min<unsigned>(blocks, (unsigned)(n_cu * 8));
      k_discover_persistent<<<pblocks, 256, 0, st>>>(da, d_next);
    } else {
      if (wide && use_lds) k_discover_wide<true><<<dblocks, kDiscWideBlock, lds_bytes, st>>>(da);
      else if (wide) k_discover_wide<false><<<dblocks, kDiscWideBlock, 0, st>>>(da);
      else if (use_lds) k_discover<true><<<dblocks, kDiscBlock, lds_bytes, st>>>(da);
      else k_discover<false><<<dblocks, kDiscBlock, 0, st>>>(da);
    }
    HIPCHK(hipGetLastError());
    S.n_discover_launches++;
    HIPCHK(hipEventRecord(e1, st));
    unsigned long long hcnt = 0;
    HIPCHK(read_back(sc, &hcnt, d_cnt, 8, st));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    disc_ms += ms;
    if (hcnt > cap) {
      if (re - rb <= 1) {
        set_error("forward records of a single ray exceed the record buffer");
        return AMS_ERR_OVERFLOW;
      }
      S.overflow++;
      cur_chunk = std::max<long long>(1, (re - rb) / 2);
      continue;  // redo [rb, rb + cur_chunk)
    }
    size_t nu = 0;
    int rc = sort_unique(sc, c_hi, c_lo, (size_t)hcnt, st, &nu, kbits);
    if (rc) return rc;
    S.records_raw += (long long)hcnt;
    if (acc_n + nu > acc_cap) {
      // compact the running set, then grow its buffers if still needed
      size_t m = 0;
      rc = sort_unique(sc, a_hi, a_lo, acc_n, st, &m, kbits);
      if (rc) return rc;
      acc_n = m;
      if (acc_n + nu > acc_cap) {
        size_t ncap = std::max(2 * acc_cap, acc_n + nu);
        uint64_t *nh = (uint64_t *)dmalloc(ncap * 8), *nl = (uint64_t *)dmalloc(ncap * 8);
        if (!nh || !nl) {
          dfree(nh);
          dfree(nl);
          set_error("device allocation failed: record set");
          return AMS_ERR_ALLOC;
        }
        HIPCHK(hipMemcpyAsync(nh, a_hi, acc_n * 8, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipMemcpyAsync(nl, a_lo, acc_n * 8, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
        sc.adopt("acc_hi", nh, ncap * 8);
        sc.adopt("acc_lo", nl, ncap * 8);
        a_hi = nh;
        a_lo = nl;
        acc_cap = ncap;
      }
    }
    HIPCHK(hipMemcpyAsync(a_hi + acc_n, c_hi, nu * 8, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(a_lo + acc_n, c_lo, nu * 8, hipMemcpyDeviceToDevice, st));
    acc_sorted = acc_n == 0;  // a second run appended: no longer sorted as a whole
    acc_n += nu;
    rb = re;
    cur_chunk = chunk;
  }
  if (o->seed_keys) {
    acc_hi.swap(uh);
    acc_lo.swap(ul);
    // seeds: sort + unique on the host (small)
    std::vector<size_t> ix(acc_hi.size());
    for (size_t k = 0; k < ix.size(); k++) ix[k] = k;
    std::sort(ix.begin(), ix.end(), [&](size_t a0, size_t b0) {
      return acc_hi[a0] != acc_hi[b0] ? acc_hi[a0] < acc_hi[b0] : acc_lo[a0] < acc_lo[b0];
    });
    std::vector<uint64_t> h2, l2;
    for (size_t k : ix) {
      if (!h2.empty() && h2.back() == acc_hi[k] && l2.back() == acc_lo[k]) continue;
      h2.push_back(acc_hi[k]);
      l2.push_back(acc_lo[k]);
    }
    acc_hi.swap(h2);
    acc_lo.swap(l2);
  } else {
    size_t m = acc_n;
    if (!acc_sorted) {
      int rc = sort_unique(sc, a_hi, a_lo, acc_n, st, &m, kbits);
      if (rc) return rc;
    }
    acc_hi.resize(m);
    acc_lo.resize(m);
    if (m) {  // through page-locked memory, as read_back
      uint64_t *h = (uint64_t *)sc.host(16 * m);
      if (!h) return AMS_ERR_ALLOC;
      HIPCHK(hipMemcpyAsync(h, a_hi, m * 8, hipMemcpyDeviceToHost, st));
      HIPCHK(hipMemcpyAsync(h + m, a_lo, m * 8, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      std::memcpy(acc_hi.data(), h, m * 8);
      std::memcpy(acc_lo.data(), h + m, m * 8);
    }
  }
  n_unique_total = acc_hi.size();
  S.rays = n_rays;
  S.records_unique = (long long)n_unique_total;
  S.t_discover_ms = disc_ms;
  unsigned long long hctr[kCtrN];
  HIPCHK(read_counters(sc, hctr, d_ctr, st));
  S.fwd_nodes = (long long)hctr[0];
  S.fwd_tris = (long long)hctr[1];
  S.fwd_spheres = (long long)hctr[2];
  S.fwd_queries = (long long)hctr[3];

  // ---------------- path list and round-1 tasks (host) -------------------
  std::vector<uint64_t> p_hi, p_lo;
  std::vector<int> rec_path(n_unique_total);
  for (size_t k = 0; k < n_unique_total; k++) {
    uint64_t h = acc_hi[k], l = acc_lo[k] & ~0xFFFFFull;
    if (p_hi.empty() || p_hi.back() != h || p_lo.back() != l) {
      p_hi.push_back(h);
      p_lo.push_back(l);
    }
    rec_path[k] = (int)p_hi.size() - 1;
  }
  const int n_paths = (int)p_hi.size();
  S.paths = n_paths;
  long long *d_H = (long long *)sc.get("H64", sizeof(long long) * 2 * (size_t)n_src * n_dst);
  if (!d_H) return AMS_ERR_ALLOC;
  HIPCHK(hipMemsetAsync(d_H, 0, sizeof(long long) * 2 * (size_t)n_src * n_dst, st));
  HIPCHK(hipMemsetAsync(d_ctr, 0, sizeof(unsigned long long) * kCtrN * kCtrSlots, st));
  // task export file, created for every call as the reference's path.dat (optixPathTracer.cpp:831)
  std::unique_ptr<FILE, int (*)(FILE *)> info_file(nullptr, &std::fclose);
  if (o->task_info_path) {
    info_file.reset(std::fopen(o->task_info_path, "wb"));
    if (!info_file) {
      set_error(std::string("task export: cannot open ") + o->task_info_path);
      return AMS_ERR_ARG;
    }
  }
  if (n_paths > 0) {
    uint64_t *d_phi = (uint64_t *)sc.get("p_hi", 8 * (size_t)n_paths), *d_plo = (uint64_t *)sc.get("p_lo", 8 * (size_t)n_paths);
    if (!d_phi || !d_plo) return AMS_ERR_ALLOC;
    HIPCHK(hipMemcpyAsync(d_phi, p_hi.data(), 8 * (size_t)n_paths, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_plo, p_lo.data(), 8 * (size_t)n_paths, hipMemcpyHostToDevice, st));
    // tasks grouped per tile (stable), optixPathTracer.cpp:287-330
    std::vector<int> tp, tt;
    {
      std::vector<long long> cnt(n_grid + 1, 0);
      for (size_t k = 0; k < n_unique_total; k++) cnt[(acc_lo[k] & 0xFFFFF) + 1]++;
      for (long long g = 0; g < n_grid; g++) cnt[g + 1] += cnt[g];
      tp.resize(n_unique_total);
      tt.resize(n_unique_total);
      for (size_t k = 0; k < n_unique_total; k++) {
        long long g = (long long)(acc_lo[k] & 0xFFFFF);
        long long pos = cnt[g]++;
        tp[pos] = rec_path[k];
        tt[pos] = (int)g;
      }
    }
    // (path, tile) pairs already traced: a bitmap over paths x tiles while it
    // stays within 128 MB (the room: 15 KB, the office: about 1 MB), else a
    // hash set.  mark() returns true for a pair not traced before.
    const uint64_t n_pairs = (uint64_t)n_paths * (uint64_t)n_grid;
    const bool traced_bits_on = n_pairs <= (1ull << 30);
    std::vector<uint64_t> traced_bits(traced_bits_on ? (size_t)((n_pairs + 63) / 64) : 0, 0);
    std::unordered_set<uint64_t> traced_set;
    auto mark = [&](int p, long long g) -> bool {
      if (traced_bits_on) {
        const uint64_t b = (uint64_t)p * (uint64_t)n_grid + (uint64_t)g;
        uint64_t &w = traced_bits[b >> 6];
        const uint64_t m = 1ull << (b & 63);
        if (w & m) return false;
        w |= m;
        return true;
      }
      return traced_set.insert(((uint64_t)p << 24) | (uint64_t)g).second;
    };
    for (size_t k = 0; k < tp.size(); k++) mark(tp[k], tt[k]);

    EvalArgs ea;
    ea.tnodes = s->d_tnodes;
    ea.n_tnodes = s->n_tnodes;
    // Evaluation walks the binary hierarchy: its legs are short bounded
    // segments.  Flags bit 11 makes scenes in global memory walk the four-wide
    // one here too (the office measured 2.60 against 2.31 s per step with it).
    const bool eval_wide = !eval_lds && s->d_tnodes4 && !(o->flags & 1024) && (o->flags & 2048);
    ea.tnodes4 = eval_wide ? s->d_tnodes4 : nullptr;
    ea.tris = s->d_tris;
    ea.n_tris = s->n_tri;
    ea.aux = s->d_aux;
    ea.mat_prop = s->d_mat_prop;
    ea.refl_phase = s->d_refl_phase;
    ea.planes = s->d_planes;
    ea.planes_f = s->d_planes_f;
    ea.pbox = s->d_pbox;
    ea.tplane = s->d_tplane;
    ea.src = d_src;
    ea.n_src = n_src;
    ea.dst = d_dst;
    ea.n0_dst = rx->n0;
    ea.n1_dst = rx->n1;
    ea.n_dst = n_dst;
    ea.gs = gs;
    ea.n1g = n1g;
    ea.flags = o->flags;
    ea.inv_wave_len_d = 1.0 / o->wave_len;
    ea.wave_len = (float)o->wave_len;
    ea.src_r2 = tx->radius * tx->radius;
    float sacc = std::max(tx->radius, AMS_ACC_MIN);
    ea.src_acc2 = sacc * sacc;
    ea.src_dir = make_float3(tx->dir[0], tx->dir[1], tx->dir[2]);
    ea.dst_dir = make_float3(rx->dir[0], rx->dir[1], rx->dir[2]);
    ea.src_bptype = tx->bptype;
    ea.dst_bptype = rx->bptype;
    ea.path_hi = d_phi;
    ea.path_lo = d_plo;
    ea.items = gs * gs * n_src;
    ea.H = d_H;
    ea.counters = d_ctr;
    const int bpt = (ea.items + kEvalBlock - 1) / kEvalBlock;     // blocks per tile slice
    const int block = ((ea.items + bpt - 1) / bpt + 63) / 64 * 64;  // multiple of the wave size
    // tasks per workgroup: longer slices when a tile spans many workgroups
    // (many Tx elements) to cut the FP64 atomics into H; short ones otherwise
    const int slice = o->tasks_per_unit > 0 ? o->tasks_per_unit : (bpt >= 8 ? 64 : 16);
    float eval_ms = 0.0f;
    ea.gs2 = gs * gs;
    ea.beam = nullptr;
    // Beam candidate lists (k_beam) ahead of every evaluation launch: by
    // default for Tx arrays of 256 elements and more, whose lanes share one
    // pair's lists (with 64, the office AP, building the lists cost more
    // than the walks they save: office H_d 114 against 70 ms); flags bit 14
    // turns them off, bit 15 on for any Tx array.  A launch covers at most
    // kBeamBytes of lists.
    const bool beams = !(o->flags & 16384) && ((o->flags & 32768) || n_src >= 256) && !info_file && !eval_wide;
    beams_used = beams;
    BeamArgs ba{};
    bool beam_planar = false;
    // a wave's 64 lanes share one Rx when the Tx count is a multiple of 64
    const bool beam_uni = n_src % 64 == 0 && !(o->flags & 65536);
    size_t beam_per_unit = 0;
    int unit_chunk = 65535;
    if (beams) {
      ba.legs = o->max_reflection + 1;
      ba.stride = 1 + ba.legs * kBeamCand;
      ba.slice = slice;
      beam_planar = tx_box_corners(tx, ba.hx, ba.hy, ba.hz);
      ba.margin0 = 2e-3f;
      ba.margin_rel = 2e-5f;
      beam_per_unit = (size_t)slice * gs * gs * ba.stride * sizeof(int);
      unit_chunk = (int)std::max<size_t>(1, std::min<size_t>(65535, kBeamBytes / beam_per_unit));
      ea.beam_stride = ba.stride;
      ea.beam_slice = slice;
    }
    // one launch of units [u0, u0 + un) of a round: the lists, then the evaluation
    auto launch_eval = [&](EvalArgs eb, unsigned un) -> int {
      if (beams) {
        ba.beam = (int *)sc.get("beam", beam_per_unit * un);
        if (!ba.beam) return AMS_ERR_ALLOC;
        eb.beam = ba.beam;
        const unsigned bx = (unsigned)((slice * gs * gs + kBeamBlock - 1) / kBeamBlock);
        if (beam_planar) k_beam<4><<<dim3(bx, un), kBeamBlock, 0, st>>>(eb, ba);
        else k_beam<8><<<dim3(bx, un), kBeamBlock, 0, st>>>(eb, ba);
        if (beam_uni) {
          if (eval_lds) k_evaluate<true, false, true, true><<<dim3(bpt, un), block, (unsigned)eval_scene_lds, st>>>(eb);
          else k_evaluate<false, false, true, true><<<dim3(bpt, un), block, 0, st>>>(eb);
        } else {
          if (eval_lds) k_evaluate<true, false, true><<<dim3(bpt, un), block, (unsigned)eval_scene_lds, st>>>(eb);
          else k_evaluate<false, false, true><<<dim3(bpt, un), block, 0, st>>>(eb);
        }
      } else if (eb.info) {
        if (eval_lds) k_evaluate<true, true><<<dim3(bpt, un), block, (unsigned)eval_scene_lds, st>>>(eb);
        else if (eval_wide) k_evaluate_wide<true><<<dim3(bpt, un), block, 0, st>>>(eb);
        else k_evaluate<false, true><<<dim3(bpt, un), block, 0, st>>>(eb);
      } else if (eval_lds) {
        k_evaluate<true, false><<<dim3(bpt, un), block, (unsigned)eval_scene_lds, st>>>(eb);
      } else if (eval_wide) {
        k_evaluate_wide<false><<<dim3(bpt, un), block, 0, st>>>(eb);
      } else {
        k_evaluate<false, false><<<dim3(bpt, un), block, 0, st>>>(eb);
      }
      HIPCHK(hipGetLastError());
      S.n_evaluate_launches++;
      return AMS_OK;
    };
    // Rounds built on the device: the expansion over the bitmap of traced
    // pairs, the grouping by tile and the next round's units are kernels; the
    // host reads back two counts per round and launches the next evaluation
    // with its exact grid (HIP has no launch from the device, and a unit loop
    // inside the evaluation kernel cost it about 24 more spilled registers).
    // The default: with candidate lists the evaluation launches are short
    // and the host's round trips show (the room step 25.0-25.7 against
    // 25.8-26.4 ms with the host loop; office H_d 77.3 against 79.5 ms per
    // trace).  Flags bit 12 forces the host loop (bit 13, the device rounds,
    // is the default); the task export and maps beyond 2^27 pairs use the
    // host loop.
    const bool dev_rounds = !info_file && traced_bits_on && n_pairs <= (1ull << 27) && !(o->flags & 4096);
    if (dev_rounds && !tp.empty()) {
      const size_t np = (size_t)n_pairs, words = (size_t)((n_pairs + 31) / 32);
      const size_t ng1 = (size_t)n_grid + 1;
      // two rounds' task lists (paths, tiles, success flags) and their tile
      // and unit offsets, the two bit maps, the counts
      int *d_path[2], *d_tile[2], *d_tstart[2], *d_ustart[2];
      unsigned char *d_okr[2];
      for (int b = 0; b < 2; b++) {
        d_path[b] = (int *)sc.get(b ? "dr_path1" : "dr_path0", 4 * np);
        d_tile[b] = (int *)sc.get(b ? "dr_tile1" : "dr_tile0", 4 * np);
        d_okr[b] = (unsigned char *)sc.get(b ? "dr_ok1" : "dr_ok0", np);
        d_tstart[b] = (int *)sc.get(b ? "dr_ts1" : "dr_ts0", 4 * ng1);
        d_ustart[b] = (int *)sc.get(b ? "dr_us1" : "dr_us0", 4 * ng1);
        if (!d_path[b] || !d_tile[b] || !d_okr[b] || !d_tstart[b] || !d_ustart[b]) return AMS_ERR_ALLOC;
      }
      unsigned *d_traced = (unsigned *)sc.get("dr_traced", 4 * words);
      unsigned *d_fresh = (unsigned *)sc.get("dr_fresh", 4 * words);
      int *d_cnt = (int *)sc.get("dr_cnt", 4 * ng1);
      int *d_info = (int *)sc.get("dr_info", 8);
      // the units of the round being evaluated (units <= tasks <= pairs);
      // the next round's are written after this round's evaluation
      int *d_units = (int *)sc.get("dr_units", 12 * np);
      if (!d_traced || !d_fresh || !d_cnt || !d_info || !d_units) return AMS_ERR_ALLOC;
      const unsigned units_grid = (unsigned)std::min<long long>(1024, (n_grid + 255) / 256);
      // round 1 from the host: tasks, offsets, the traced map (tile-major)
      const size_t nt1 = tp.size();
      std::vector<int> ts(ng1, 0), us(ng1, 0);
      for (size_t k = 0; k < nt1; k++) ts[tt[k] + 1]++;
      for (long long g = 0; g < n_grid; g++) {
        us[g + 1] = us[g] + (ts[g + 1] + slice - 1) / slice;
        ts[g + 1] += ts[g];
      }
      std::vector<unsigned> tmap(words, 0u);
      for (size_t k = 0; k < nt1; k++) {
        const uint64_t b = (uint64_t)tt[k] * (uint64_t)n_paths + (uint64_t)tp[k];
        tmap[b >> 5] |= 1u << (b & 31);
      }
      HIPCHK(hipMemcpyAsync(d_path[0], tp.data(), 4 * nt1, hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(d_tile[0], tt.data(), 4 * nt1, hipMemcpyHostToDevice, st));
      HIPCHK(hipMemsetAsync(d_okr[0], 0, nt1, st));
      HIPCHK(hipMemcpyAsync(d_tstart[0], ts.data(), 4 * ng1, hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(d_ustart[0], us.data(), 4 * ng1, hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(d_traced, tmap.data(), 4 * words, hipMemcpyHostToDevice, st));
      HIPCHK(hipMemsetAsync(d_fresh, 0, 4 * words, st));
      ea.info = nullptr;
      const bool expand = !(o->flags & 2);
      long long nt_r = (long long)nt1, nu_r = us[n_grid];
      for (int cur = 0;;) {
        S.rounds++;
        S.tasks += nt_r;
        k_units<<<units_grid, 256, 0, st>>>(d_tstart[cur], d_ustart[cur], (int)n_grid, slice, d_units,
                                             d_units + np, d_units + 2 * np);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(e0, st));
        for (long long u0 = 0; u0 < nu_r; u0 += unit_chunk) {
          const unsigned un = (unsigned)std::min<long long>(unit_chunk, nu_r - u0);
          EvalArgs eb = ea;
          eb.task_path = d_path[cur];
          eb.task_ok = d_okr[cur];
          eb.unit_tile = d_units + u0;
          eb.unit_begin = d_units + np + u0;
          eb.unit_end = d_units + 2 * np + u0;
          const int rc = launch_eval(eb, un);
          if (rc != AMS_OK) return rc;
        }
        HIPCHK(hipEventRecord(e1, st));
        if (!expand) {
          HIPCHK(hipEventSynchronize(e1));
          float ms = 0.0f;
          HIPCHK(hipEventElapsedTime(&ms, e0, e1));
          eval_ms += ms;
          break;
        }
        const int nxt = cur ^ 1;
        k_expand<<<1024, 256, 0, st>>>(d_path[cur], d_tile[cur], d_okr[cur], d_tstart[cur] + n_grid, (int)n_grid,
                                        n1g, n_paths, d_traced, d_fresh);
        k_tile_count<<<(unsigned)n_grid, 256, 0, st>>>(d_fresh, n_paths, d_cnt);
        k_tile_scan<<<1, 256, 0, st>>>(d_cnt, (int)n_grid, slice, d_tstart[nxt], d_ustart[nxt], d_info);
        k_tile_emit<<<(unsigned)n_grid, 256, 0, st>>>(d_fresh, n_paths, d_tstart[nxt], d_path[nxt], d_tile[nxt],
                                                       d_okr[nxt]);
        HIPCHK(hipGetLastError());
        int hi2[2];
        HIPCHK(read_back(sc, hi2, d_info, 8, st));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        eval_ms += ms;
        nt_r = hi2[0];
        nu_r = hi2[1];
        if (nt_r == 0) break;
        cur = nxt;
      }
      tp.clear();  // the host loop has nothing left to do
    }
    while (!tp.empty()) {
      S.rounds++;
      size_t nt = tp.size();
      S.tasks += (long long)nt;
      // work units: (tile, slice of tasks)
      std::vector<int> ut, ub, ue;
      for (size_t k = 0; k < nt;) {
        size_t e = k;
        while (e < nt && tt[e] == tt[k] && e - k < (size_t)slice) e++;
        ut.push_back(tt[k]);
        ub.push_back((int)k);
        ue.push_back((int)e);
        k = e;
      }
      int nu = (int)ut.size();
      // one upload per round from page-locked memory: task paths | units
      // (tile, begin, end) | success flags (zero); the flags come back into
      // the bytes after it
      const size_t up = 4 * nt + 12 * (size_t)nu + nt;
      char *d_round = (char *)sc.get("t_round", up);
      char *h_round = (char *)s->stage.get(up + nt);
      if (!d_round || !h_round) return AMS_ERR_ALLOC;
      std::memcpy(h_round, tp.data(), 4 * nt);
      std::memcpy(h_round + 4 * nt, ut.data(), 4 * (size_t)nu);
      std::memcpy(h_round + 4 * nt + 4 * (size_t)nu, ub.data(), 4 * (size_t)nu);
      std::memcpy(h_round + 4 * nt + 8 * (size_t)nu, ue.data(), 4 * (size_t)nu);
      std::memset(h_round + 4 * nt + 12 * (size_t)nu, 0, nt);
      HIPCHK(hipMemcpyAsync(d_round, h_round, up, hipMemcpyHostToDevice, st));
      int *d_tp = (int *)d_round;
      int *d_u = (int *)(d_round + 4 * nt);
      unsigned char *d_ok = (unsigned char *)(d_round + 4 * nt + 12 * (size_t)nu);
      ea.task_path = d_tp;
      ea.unit_tile = d_u;
      ea.unit_begin = d_u + nu;
      ea.unit_end = d_u + 2 * nu;
      ea.task_ok = d_ok;
      // task export of this round: [src][task][Rx in tile], zero-filled
      size_t n_info = 0;
      ea.info = nullptr;
      ea.info_nt = (long long)nt;
      ea.gs2 = gs * gs;
      if (info_file) {
        n_info = nt * (size_t)n_src * (size_t)(gs * gs);
        if (n_info > kMaxTaskInfo) {
          set_error("task export of one round exceeds 2^26 records (about 3.5 GB); trace fewer rays");
          return AMS_ERR_ARG;
        }
        ea.info = (ams_task_info *)sc.get("t_info", n_info * sizeof(ams_task_info));
        if (!ea.info) return AMS_ERR_ALLOC;
        HIPCHK(hipMemsetAsync(ea.info, 0, n_info * sizeof(ams_task_info), st));
      }
      HIPCHK(hipEventRecord(e0, st));
      for (int u0 = 0; u0 < nu; u0 += unit_chunk) {
        const unsigned un = (unsigned)std::min(unit_chunk, nu - u0);
        EvalArgs eb = ea;
        eb.unit_tile += u0;
        eb.unit_begin += u0;
        eb.unit_end += u0;
        const int rc = launch_eval(eb, un);
        if (rc != AMS_OK) return rc;
      }
      HIPCHK(hipEventRecord(e1, st));
      const unsigned char *ok = (const unsigned char *)h_round + up;
      HIPCHK(hipMemcpyAsync(h_round + up, d_ok, nt, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, e0, e1));
      eval_ms += ms;
      if (info_file && n_info > 0) {  // optixPathTracer.cpp:244-256
        std::vector<ams_task_info> buf(n_info);
        HIPCHK(hipMemcpy(buf.data(), ea.info, n_info * sizeof(ams_task_info), hipMemcpyDeviceToHost));
        if (std::fwrite(buf.data(), sizeof(ams_task_info), n_info, info_file.get()) != n_info) {
          set_error("task export: write failed");
          return AMS_ERR_ARG;
        }
      }
      // neighbour expansion, optixPathTracer.cpp:264-284 (linear bounds only)
      std::vector<int> np_, nt_;
      for (size_t k = 0; k < nt; k++) {
        if (!ok[k]) continue;
        long long g = tt[k];
        long long nb[4] = {g - 1, g + 1, g - n1g, g + n1g};
        for (int q = 0; q < 4; q++) {
          if (nb[q] < 0 || nb[q] >= n_grid) continue;
          if (mark(tp[k], nb[q])) {
            np_.push_back(tp[k]);
            nt_.push_back((int)nb[q]);
          }
        }
      }
      if (o->flags & 2) {
        np_.clear();
        nt_.clear();
      }
      std::vector<long long> cnt(n_grid + 1, 0);
      for (size_t k = 0; k < nt_.size(); k++) cnt[nt_[k] + 1]++;
      for (long long g = 0; g < n_grid; g++) cnt[g + 1] += cnt[g];
      tp.assign(np_.size(), 0);
      tt.assign(np_.size(), 0);
      for (size_t k = 0; k < np_.size(); k++) {
        long long pos = cnt[nt_[k]]++;
        tp[pos] = np_[k];
        tt[pos] = nt_[k];
      }
    }
    S.t_evaluate_ms = eval_ms;
  }
  HIPCHK(read_counters(sc, hctr, d_ctr, st));
  if (hctr[15]) {  // before the caller's buffer is touched
    set_error("a channel entry's partial sum reached the fixed-point range of H (|H| >= 64): " +
              std::to_string(hctr[15]) + " lanes");
    return AMS_ERR_ARG;
  }
  // accumulate into the caller's FP32 result, optixPathTracer.cu:259-264
  size_t nH = 2 * (size_t)n_src * n_dst;
  if (nH > 0) {
    k_add_h<<<(unsigned)((nH + 255) / 256), 256, 0, st>>>(d_H, H_out, nH);
    HIPCHK(hipGetLastError());
  }
  S.bwd_nodes = (long long)hctr[0];
  S.bwd_tris = (long long)hctr[1];
  S.bwd_queries = (long long)hctr[3];
  S.chains = (long long)hctr[4];
  S.successes = (long long)hctr[5];
  S.bwd_walks = (long long)(beams_used ? hctr[2] : hctr[3]);
  S.beam_pairs = (long long)hctr[6];
  S.beam_nodes = (long long)hctr[7];
  S.beam_tris = (long long)hctr[8];
  S.beam_walk_legs = (long long)hctr[9];
  S.beam_cands = (long long)hctr[10];
  S.t_total_ms = ms_since(t_start);
  if (stats) *stats = S;
  if (o->keys_out && o->keys_cap > 0) {
    size_t c = std::min<size_t>((size_t)o->keys_cap, n_unique_total);
    for (size_t k = 0; k < c; k++) {
      o->keys_out[2 * k] = acc_hi[k];
      o->keys_out[2 * k + 1] = acc_lo[k];
    }
  }
  return AMS_OK;
}

}  // namespace ams
