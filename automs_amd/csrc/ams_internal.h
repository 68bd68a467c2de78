// ams_internal.h -- library-private state: the scene handle, a named scratch
// pool reused across calls, device allocation bookkeeping and error text.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/ams.h"
#include "ams_common.h"

namespace ams {

void set_error(const std::string &msg);

// Device allocations with a pointer -> size record, so that every byte the
// library holds can be reported (the role of cudaMemManager.h:23-81).
void *dmalloc(size_t bytes);
void dfree(void *p);
int64_t outstanding();

template <typename T>
int dev_upload(T **dst, const void *src, size_t bytes) {
  *dst = (T *)dmalloc(bytes > 0 ? bytes : 16);
  if (!*dst) return AMS_ERR_ALLOC;
  if (bytes && hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) != hipSuccess) {
    set_error("hipMemcpy upload failed");
    return AMS_ERR_HIP;
  }
  return AMS_OK;
}

// Named device buffers that only grow; reused by later calls on the handle.
// Page-locked host staging for the per-round uploads and read-backs of a
// trace: the copies run on the DMA engines and need no compute units, so
// they do not queue behind other traces' kernels.  Grows as needed; the
// caller synchronises its stream before asking again.
struct PinnedStage {
  void *p = nullptr;
  size_t bytes = 0;
  void *get(size_t n) {
    if (n <= bytes) return p;
    release();
    size_t want = n + n / 4;
    if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) {
      p = nullptr;
      set_error("pinned host allocation failed");
      return nullptr;
    }
    bytes = want;
    return p;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
  }
};

struct Scratch {
  std::map<std::string, std::pair<void *, size_t>> bufs;
  void *get(const char *name, size_t bytes) {
    if (bytes == 0) bytes = 16;
    auto &b = bufs[name];
    if (b.second >= bytes) return b.first;
    if (b.first) dfree(b.first);
    size_t want = bytes + bytes / 4;
    b.first = dmalloc(want);
    b.second = b.first ? want : 0;
    if (!b.first) set_error(std::string("device allocation failed: ") + name);
    return b.first;
  }
  // take ownership of an allocation made with dmalloc under this name
  void adopt(const char *name, void *p, size_t bytes) {
    auto &b = bufs[name];
    if (b.first && b.first != p) dfree(b.first);
    b.first = p;
    b.second = bytes;
  }
  void release() {
    for (auto &kv : bufs)
      if (kv.second.first) dfree(kv.second.first);
    bufs.clear();
    pin.release();
  }
  // page-locked host bytes for small read-backs (counts, counters)
  void *host(size_t bytes) { return pin.get(bytes); }
  PinnedStage pin;
};

}  // namespace ams

struct ams_scene {
  int n_tri = 0, n_mat = 0;
  std::vector<int> plane;
  AmsNode *d_tnodes = nullptr;
  int n_tnodes = 0;
  AmsNode4 *d_tnodes4 = nullptr;  // four-wide form of the same hierarchy
  int n_tnodes4 = 0, depth4 = 0;
  AmsTri *d_tris = nullptr;
  AmsTriAux *d_aux = nullptr;
  AmsPlane *d_planes = nullptr;
  float4 *d_planes_f = nullptr;  // the same planes rounded to f32: (n, d)
  AmsPlaneBox *d_pbox = nullptr;
  AmsPlane *d_tplane = nullptr;  // each triangle's own plane (FP64), by original index
  int n_inexact = 0;             // triangles whose own plane is not bitwise their plane id's
  float *d_mat_prop = nullptr;
  float *d_refl_phase = nullptr;
  ams::Scratch scratch;
  ams::PinnedStage stage;  // host side of the per-round task uploads
  long long perm_n = -1;  // fib_n of the cached ray permutation in scratch "perm"
};

namespace ams {
int upload_scene(ams_scene *s, int n_tri, const float *tri, const int32_t *mat, int32_t n_mat,
                 const float *mat_prop, const float *refl_phase);
int trace_channel(ams_scene *s, const ams_array *tx, const ams_array *rx, const ams_trace_opts *o,
                  float *H_out, ams_trace_stats *stats);
}  // namespace ams
