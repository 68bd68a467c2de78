// ams_bvh.cpp -- binned SAH build of the threaded BVH used by the gfx950
// traversal kernels, plus the coplanar-triangle grouping that gives path keys
// their plane ids.  The reference delegates both to OptiX's closed builder
// (optixPathTracer.cpp:443-546); this is an independent construction.
#include "ams_bvh.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>

namespace {

constexpr int kBins = 16;

struct Box {
  double lo[3] = {INFINITY, INFINITY, INFINITY};
  double hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  void grow(const double *l, const double *h) {
    for (int k = 0; k < 3; k++) {
      lo[k] = std::min(lo[k], l[k]);
      hi[k] = std::max(hi[k], h[k]);
    }
  }
  void grow_pt(const double *p) { grow(p, p); }
  double area() const {
    double d[3];
    for (int k = 0; k < 3; k++) d[k] = std::max(0.0, hi[k] - lo[k]);
    return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
  }
};

struct Builder {
  const double *lo, *hi;
  std::vector<double> c;  // centroids
  std::vector<int> idx;
  std::vector<AmsNode> nodes;
  std::vector<int> right;  // right child of each inner node
  int leaf_max;

  static float down(double x) {
    float f = (float)x;
    if ((double)f > x) f = std::nextafter(f, -INFINITY);
    return f;
  }
  static float up(double x) {
    float f = (float)x;
    if ((double)f < x) f = std::nextafter(f, INFINITY);
    return f;
  }

  int emit(const Box &b) {
    AmsNode n;
    for (int k = 0; k < 3; k++) {
      n.lo[k] = down(b.lo[k]);
      n.hi[k] = up(b.hi[k]);
    }
    n.miss = -1;
    n.info = 0;
    nodes.push_back(n);
    right.push_back(-1);
    return (int)nodes.size() - 1;
  }

  void make_leaf(int me, int first, int count) { nodes[me].info = (first << 3) | count; }

  // Builds the subtree of idx[first, first+count) in preorder; returns its root.
  int build(int first, int count) {
    Box b, cb;
    for (int i = first; i < first + count; i++) {
      int p = idx[i];
      b.grow(lo + 3 * p, hi + 3 * p);
      cb.grow_pt(&c[3 * p]);
    }
    int me = emit(b);
    if (count <= 1) {
      make_leaf(me, first, count);
      return me;
    }
    // binned SAH over centroid bounds
    double best_cost = INFINITY;
    int best_ax = -1, best_split = -1;
    for (int ax = 0; ax < 3; ax++) {
      double ext = cb.hi[ax] - cb.lo[ax];
      if (!(ext > 0)) continue;
      Box bb[kBins];
      int bn[kBins] = {0};
      double sc = kBins / ext;
      for (int i = first; i < first + count; i++) {
        int p = idx[i];
        int k = std::min(kBins - 1, (int)((c[3 * p + ax] - cb.lo[ax]) * sc));
        bn[k]++;
        bb[k].grow(lo + 3 * p, hi + 3 * p);
      }
      double al[kBins], ar[kBins];
      int nl[kBins], nr[kBins];
      Box acc;
      int n = 0;
      for (int k = 0; k < kBins; k++) {
        if (bn[k]) acc.grow(bb[k].lo, bb[k].hi);
        n += bn[k];
        al[k] = acc.area();
        nl[k] = n;
      }
      acc = Box();
      n = 0;
      for (int k = kBins - 1; k >= 0; k--) {
        if (bn[k]) acc.grow(bb[k].lo, bb[k].hi);
        n += bn[k];
        ar[k] = acc.area();
        nr[k] = n;
      }
      for (int s = 0; s < kBins - 1; s++) {
        if (nl[s] == 0 || nr[s + 1] == 0) continue;
        double cost = al[s] * nl[s] + ar[s + 1] * nr[s + 1];
        if (cost < best_cost) {
          best_cost = cost;
          best_ax = ax;
          best_split = s;
        }
      }
    }
    double pa = b.area();
    double leaf_cost = (double)count;
    double split_cost = 0.5 + (pa > 0 ? best_cost / pa : INFINITY);
    if (count <= leaf_max && (best_ax < 0 || leaf_cost <= split_cost)) {
      make_leaf(me, first, count);
      return me;
    }
    int mid;
    if (best_ax < 0) {
      mid = first + count / 2;  // identical centroids: split by index
    } else {
      double ext = cb.hi[best_ax] - cb.lo[best_ax];
      double sc = kBins / ext;
      auto it = std::partition(idx.begin() + first, idx.begin() + first + count, [&](int p) {
        int k = std::min(kBins - 1, (int)((c[3 * p + best_ax] - cb.lo[best_ax]) * sc));
        return k <= best_split;
      });
      mid = (int)(it - idx.begin());
      if (mid == first || mid == first + count) mid = first + count / 2;
    }
    build(first, mid - first);
    right[me] = build(mid, first + count - mid);
    return me;
  }
};

}  // namespace

// Skip links: a node's miss target is the first node after its subtree.
static void set_links(std::vector<AmsNode> &nodes, const std::vector<int> &right, int root, int end) {
  // iterative: (node, end) pairs
  std::vector<std::pair<int, int>> st;
  st.push_back({root, end});
  while (!st.empty()) {
    auto [i, e] = st.back();
    st.pop_back();
    nodes[i].miss = e;
    if (nodes[i].info != 0) continue;
    int r = right[i];
    st.push_back({i + 1, r});
    st.push_back({r, e});
  }
}

void ams_bvh_build(int n, const double *lo, const double *hi, int leaf_max, AmsBvhBuild *out) {
  out->nodes.clear();
  out->order.clear();
  if (n <= 0) return;
  Builder B;
  B.lo = lo;
  B.hi = hi;
  B.leaf_max = std::max(1, std::min(7, leaf_max));
  B.c.resize(3 * (size_t)n);
  for (size_t i = 0; i < 3 * (size_t)n; i++) B.c[i] = 0.5 * (lo[i] + hi[i]);
  B.idx.resize(n);
  for (int i = 0; i < n; i++) B.idx[i] = i;
  B.nodes.reserve(2 * (size_t)n);
  B.build(0, n);
  set_links(B.nodes, B.right, 0, (int)B.nodes.size());
  out->nodes = std::move(B.nodes);
  out->order = std::move(B.idx);
}

void ams_plane_ids(int n_tri, const float *tri, int *plane_out, double *pl_n_out, double *pl_d_out, double *own_out) {
  // Quantised (normal, offset) -> smallest triangle index.  Same definition
  // as the one the path keys are compared under: unit normal from
  // cross(v1-v0, v2-v0) in FP64, sign made positive on its first component
  // above 1e-9, normal quantised to 1e-6 and offset to 1e-5 m.
  std::map<std::tuple<long long, long long, long long, long long>, int> first;
  std::vector<double> nn(3 * (size_t)n_tri);
  for (int i = 0; i < n_tri; i++) {
    const float *t = tri + 9 * (size_t)i;
    double e1[3] = {(double)t[3] - t[0], (double)t[4] - t[1], (double)t[5] - t[2]};
    double e2[3] = {(double)t[6] - t[0], (double)t[7] - t[1], (double)t[8] - t[2]};
    double cx = e1[1] * e2[2] - e1[2] * e2[1];
    double cy = e1[2] * e2[0] - e1[0] * e2[2];
    double cz = e1[0] * e2[1] - e1[1] * e2[0];
    double l = std::sqrt(cx * cx + cy * cy + cz * cz);
    double n[3] = {0, 0, 1};
    if (l > 0) {
      n[0] = cx / l;
      n[1] = cy / l;
      n[2] = cz / l;
    }
    for (int k = 0; k < 3; k++) {
      if (std::fabs(n[k]) > 1e-9) {
        if (n[k] < 0) {
          n[0] = -n[0];
          n[1] = -n[1];
          n[2] = -n[2];
        }
        break;
      }
    }
    double d = n[0] * t[0] + n[1] * t[1] + n[2] * t[2];
    if (own_out) {
      own_out[4 * (size_t)i] = n[0];
      own_out[4 * (size_t)i + 1] = n[1];
      own_out[4 * (size_t)i + 2] = n[2];
      own_out[4 * (size_t)i + 3] = d;
    }
    auto key = std::make_tuple(std::llround(n[0] * 1e6), std::llround(n[1] * 1e6),
                               std::llround(n[2] * 1e6), std::llround(d * 1e5));
    auto it = first.find(key);
    int rep = i;
    if (it == first.end()) first.emplace(key, i);
    else rep = it->second;
    plane_out[i] = rep;
    nn[3 * i] = n[0];
    nn[3 * i + 1] = n[1];
    nn[3 * i + 2] = n[2];
    if (rep == i) {
      pl_n_out[3 * i] = n[0];
      pl_n_out[3 * i + 1] = n[1];
      pl_n_out[3 * i + 2] = n[2];
      pl_d_out[i] = d;
    }
  }
  for (int i = 0; i < n_tri; i++) {
    int r = plane_out[i];
    if (r != i) {
      pl_n_out[3 * i] = pl_n_out[3 * r];
      pl_n_out[3 * i + 1] = pl_n_out[3 * r + 1];
      pl_n_out[3 * i + 2] = pl_n_out[3 * r + 2];
      pl_d_out[i] = pl_d_out[r];
    }
  }
}

namespace {
double node_area(const AmsNode &n) {
  double dx = (double)n.hi[0] - n.lo[0], dy = (double)n.hi[1] - n.lo[1], dz = (double)n.hi[2] - n.lo[2];
  if (dx < 0 || dy < 0 || dz < 0) return 0.0;
  return dx * dy + dy * dz + dz * dx;
}
}  // namespace

int ams_bvh_collapse4(const AmsBvhBuild &b, std::vector<AmsNode4> *out) {
  const std::vector<AmsNode> &N = b.nodes;
  out->clear();
  const int nn = (int)N.size();
  auto leaf_ref = [&](int i) { return -1 - N[i].info; };
  auto kids = [&](int i, int *l, int *r) {
    *l = i + 1;
    *r = N[i + 1].miss;
  };
  int depth_max = 0;
  // explicit work list: (binary inner node, four-wide slot, depth)
  struct Job {
    int bin, slot, depth;
  };
  std::vector<Job> work;
  out->emplace_back();
  if (nn <= 1 || N[0].info != 0) {
    // a single leaf, or no triangles (no nodes, or the one placeholder node of
    // an empty scene): one node with at most one child
    AmsNode4 &r = out->back();
    float *lo[3] = {&r.lox.x, &r.loy.x, &r.loz.x}, *hi[3] = {&r.hix.x, &r.hiy.x, &r.hiz.x};
    for (int c = 0; c < 4; c++)
      for (int k = 0; k < 3; k++) {
        lo[k][c] = 0.0f;
        hi[k][c] = 0.0f;
      }
    r.child = make_int4(AMS_NODE4_EMPTY, AMS_NODE4_EMPTY, AMS_NODE4_EMPTY, AMS_NODE4_EMPTY);
    r.pad = make_int4(0, 0, 0, 0);
    if (nn > 0 && N[0].info != 0) {
      for (int k = 0; k < 3; k++) {
        lo[k][0] = N[0].lo[k];
        hi[k][0] = N[0].hi[k];
      }
      r.child.x = leaf_ref(0);
    }
    return 1;
  }
  work.push_back({0, 0, 1});
  while (!work.empty()) {
    Job j = work.back();
    work.pop_back();
    depth_max = std::max(depth_max, j.depth);
    int c[4], m = 0;
    kids(j.bin, &c[0], &c[1]);
    m = 2;
    while (m < 4) {  // open the largest inner child
      int best = -1;
      double ba = -1.0;
      for (int k = 0; k < m; k++)
        if (N[c[k]].info == 0 && node_area(N[c[k]]) > ba) {
          ba = node_area(N[c[k]]);
          best = k;
        }
      if (best < 0) break;
      int l, r;
      kids(c[best], &l, &r);
      c[best] = l;
      c[m++] = r;
    }
    AmsNode4 node;
    float *lo[3] = {&node.lox.x, &node.loy.x, &node.loz.x}, *hi[3] = {&node.hix.x, &node.hiy.x, &node.hiz.x};
    int ref[4] = {AMS_NODE4_EMPTY, AMS_NODE4_EMPTY, AMS_NODE4_EMPTY, AMS_NODE4_EMPTY};
    for (int k = 0; k < 4; k++) {
      if (k < m) {
        for (int a = 0; a < 3; a++) {
          lo[a][k] = N[c[k]].lo[a];
          hi[a][k] = N[c[k]].hi[a];
        }
      } else {
        for (int a = 0; a < 3; a++) {
          lo[a][k] = 0.0f;
          hi[a][k] = 0.0f;
        }
      }
    }
    for (int k = 0; k < m; k++) {
      if (N[c[k]].info != 0) {
        ref[k] = leaf_ref(c[k]);
      } else {
        ref[k] = (int)out->size();
        out->emplace_back();
        work.push_back({c[k], ref[k], j.depth + 1});
      }
    }
    node.child = make_int4(ref[0], ref[1], ref[2], ref[3]);
    node.pad = make_int4(0, 0, 0, 0);
    (*out)[j.slot] = node;
  }
  return depth_max;
}
