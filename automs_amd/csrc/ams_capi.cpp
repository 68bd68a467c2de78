// ams_capi.cpp -- extern "C" entry points: argument validation, error text,
// scene lifetime and allocation bookkeeping.  Compute lives in the .hip files.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>

#include <vector>

#include "../../include/ams.h"
#include "ams_bvh.h"
#include "ams_internal.h"

namespace {
thread_local std::string g_err;
std::mutex g_mem_mu;
std::unordered_map<void *, size_t> g_mem;  // live device allocations
int64_t g_mem_bytes = 0;
}  // namespace

namespace ams {

void set_error(const std::string &msg) { g_err = msg; }

void *dmalloc(size_t bytes) {
  void *p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_mem_mu);
  g_mem[p] = bytes;
  g_mem_bytes += (int64_t)bytes;
  return p;
}

void dfree(void *p) {
  if (!p) return;
  {
    std::lock_guard<std::mutex> lk(g_mem_mu);
    auto it = g_mem.find(p);
    if (it != g_mem.end()) {
      g_mem_bytes -= (int64_t)it->second;
      g_mem.erase(it);
    }
  }
  (void)hipFree(p);
}

int64_t outstanding() {
  std::lock_guard<std::mutex> lk(g_mem_mu);
  return g_mem_bytes;
}

}  // namespace ams

extern "C" {

const char *ams_last_error(void) { return g_err.c_str(); }

int ams_version(void) { return 1; }

int64_t ams_mem_outstanding(void) { return ams::outstanding(); }

int ams_scene_create(const float *tri, int32_t n_tri, const int32_t *mat, int32_t n_mat, const float *mat_prop,
                     const float *refl_phase, ams_scene **out) {
  if (!out) {
    ams::set_error("ams_scene_create: out is NULL");
    return AMS_ERR_ARG;
  }
  *out = nullptr;
  if (n_tri < 0 || n_tri > 0xFFFFE || (n_tri > 0 && (!tri || !mat)) || n_mat < 0 ||
      (n_mat > 0 && !mat_prop)) {
    ams::set_error("ams_scene_create: bad triangle / material arguments (n_tri must be < 1048575)");
    return AMS_ERR_ARG;
  }
  for (int i = 0; i < n_tri; i++) {
    if (mat[i] < 0 || mat[i] >= n_mat) {
      ams::set_error("ams_scene_create: material index out of range at triangle " + std::to_string(i));
      return AMS_ERR_ARG;
    }
  }
  ams_scene *s = new ams_scene();
  int rc = ams::upload_scene(s, n_tri, tri, mat, n_mat, mat_prop, refl_phase);
  if (rc != AMS_OK) {
    ams_scene_destroy(s);
    return rc;
  }
  *out = s;
  return AMS_OK;
}

int ams_scene_destroy(ams_scene *s) {
  if (!s) return AMS_OK;
  s->scratch.release();
  s->stage.release();
  ams::dfree(s->d_tnodes);
  ams::dfree(s->d_tris);
  ams::dfree(s->d_aux);
  ams::dfree(s->d_planes);
  ams::dfree(s->d_planes_f);
  ams::dfree(s->d_pbox);
  ams::dfree(s->d_tplane);
  ams::dfree(s->d_tnodes4);
  ams::dfree(s->d_mat_prop);
  ams::dfree(s->d_refl_phase);
  delete s;
  return AMS_OK;
}

int ams_scene_plane_ids(const ams_scene *s, int32_t *out) {
  if (!s || !out) {
    ams::set_error("ams_scene_plane_ids: NULL argument");
    return AMS_ERR_ARG;
  }
  for (int i = 0; i < s->n_tri; i++) out[i] = s->plane[i];
  return AMS_OK;
}

static bool check_array(const ams_array *a, const char *what) {
  if (!a || a->n0 <= 0 || a->n1 <= 0 || !a->pos) {
    ams::set_error(std::string("ams_trace_channel: bad ") + what + " array");
    return false;
  }
  if (a->bptype < 0 || a->bptype > 3) {
    // optixPathTracer.cu:110-135 has no default case (undefined result)
    ams::set_error(std::string("ams_trace_channel: ") + what + " beam pattern type must be 0..3");
    return false;
  }
  if (!(a->radius >= 0.0f)) {
    ams::set_error(std::string("ams_trace_channel: ") + what + " radius must be >= 0");
    return false;
  }
  return true;
}

int ams_trace_channel(ams_scene *s, const ams_array *tx, const ams_array *rx, const ams_trace_opts *o,
                      float *H_out, ams_trace_stats *stats) {
  if (!s || !o || !H_out) {
    ams::set_error("ams_trace_channel: NULL scene, options or output");
    return AMS_ERR_ARG;
  }
  if (!check_array(tx, "tx") || !check_array(rx, "rx")) return AMS_ERR_ARG;
  if (o->max_reflection < 1 || o->max_reflection > AMS_MAX_HIT) {
    ams::set_error("ams_trace_channel: max_reflection must be in 1..5");
    return AMS_ERR_ARG;
  }
  if (o->discovery_mode == 1 &&
      (o->launch_x < 1 || o->launch_y < 1 || o->samples_per_launch < 1 || o->reserve_size < 0 ||
       (long long)o->launch_x * o->launch_y > (1LL << 26))) {
    ams::set_error("ams_trace_channel: reference-mode launch_x, launch_y, samples_per_launch must be positive");
    return AMS_ERR_ARG;
  }
  if (o->discovery_mode != 0 && o->discovery_mode != 1) {
    ams::set_error("ams_trace_channel: discovery_mode must be 0 or 1");
    return AMS_ERR_ARG;
  }
  if (o->gridsize < 1 || (o->discovery_mode == 0 && o->fib_n < 1) || !(o->wave_len > 0)) {
    ams::set_error("ams_trace_channel: gridsize, fib_n and wave_len must be positive");
    return AMS_ERR_ARG;
  }
  int n1g = (rx->n1 - 1) / o->gridsize + 1, n0g = (rx->n0 - 1) / o->gridsize + 1;
  if ((long long)n1g * n0g >= (1LL << 20)) {
    ams::set_error("ams_trace_channel: more than 2^20 Rx tiles");
    return AMS_ERR_ARG;
  }
  long long items = (long long)o->gridsize * o->gridsize * tx->n0 * tx->n1;
  if (items > (1LL << 30)) {
    ams::set_error("ams_trace_channel: gridsize^2 * n_tx too large");
    return AMS_ERR_ARG;
  }
  return ams::trace_channel(s, tx, rx, o, H_out, stats);
}

int ams_host_bvh_build(const double *lo, const double *hi, int32_t n, int32_t leaf_max, void *nodes_out,
                       int32_t nodes_cap, int32_t *order_out, int32_t *n_nodes_out) {
  if (!lo || !hi || n < 0 || !nodes_out || !order_out || !n_nodes_out) {
    ams::set_error("ams_host_bvh_build: bad arguments");
    return AMS_ERR_ARG;
  }
  AmsBvhBuild B;
  ams_bvh_build(n, lo, hi, leaf_max, &B);
  *n_nodes_out = (int32_t)B.nodes.size();
  if ((int32_t)B.nodes.size() > nodes_cap) {
    ams::set_error("ams_host_bvh_build: nodes_cap too small");
    return AMS_ERR_ARG;
  }
  std::memcpy(nodes_out, B.nodes.data(), sizeof(AmsNode) * B.nodes.size());
  for (int i = 0; i < n; i++) order_out[i] = B.order[i];
  return AMS_OK;
}

int ams_host_bvh4_build(const double *lo, const double *hi, int32_t n, int32_t leaf_max, void *nodes4_out,
                        int32_t nodes4_cap, int32_t *order_out, int32_t *n_nodes4_out) {
  if ((n > 0 && (!lo || !hi)) || n < 0 || !nodes4_out || !order_out || !n_nodes4_out) {
    ams::set_error("ams_host_bvh4_build: bad arguments");
    return AMS_ERR_ARG;
  }
  AmsBvhBuild B;
  if (n > 0) ams_bvh_build(n, lo, hi, leaf_max, &B);
  std::vector<AmsNode4> N4;
  ams_bvh_collapse4(B, &N4);
  *n_nodes4_out = (int32_t)N4.size();
  if ((int32_t)N4.size() > nodes4_cap) {
    ams::set_error("ams_host_bvh4_build: nodes4_cap too small");
    return AMS_ERR_ARG;
  }
  std::memcpy(nodes4_out, N4.data(), sizeof(AmsNode4) * N4.size());
  for (int i = 0; i < n; i++) order_out[i] = B.order[i];
  return AMS_OK;
}

int ams_host_plane_ids(const float *tri, int32_t n_tri, int32_t *plane_out) {
  if ((n_tri > 0 && !tri) || !plane_out || n_tri < 0) {
    ams::set_error("ams_host_plane_ids: bad arguments");
    return AMS_ERR_ARG;
  }
  std::vector<int> pl(n_tri);
  std::vector<double> pn(3 * (size_t)n_tri), pd(n_tri);
  if (n_tri > 0) ams_plane_ids(n_tri, tri, pl.data(), pn.data(), pd.data());
  for (int i = 0; i < n_tri; i++) plane_out[i] = pl[i];
  return AMS_OK;
}

}  // extern "C"
