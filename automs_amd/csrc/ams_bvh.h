// ams_bvh.h -- host-side binned-SAH BVH builder (FP64 bounds, FP32 output
// rounded outward) producing 32-byte threaded nodes in depth-first preorder.
#pragma once
#include <vector>

#include "ams_common.h"

struct AmsBvhBuild {
  std::vector<AmsNode> nodes;  // depth-first preorder with skip links
  std::vector<int> order;      // leaf slot -> primitive index
};

// lo/hi: per-primitive bounds [n][3] in FP64.  leaf_max <= 7.
void ams_bvh_build(int n, const double *lo, const double *hi, int leaf_max, AmsBvhBuild *out);

// Four-wide hierarchy over the same leaves: every inner node's children are
// expanded, largest surface area first, until it has four; root at index 0.
// Returns the deepest root-to-leaf chain of four-wide nodes.
int ams_bvh_collapse4(const AmsBvhBuild &b, std::vector<AmsNode4> *out);

// Plane ids: for each triangle, the smallest triangle index lying on the same
// plane (unit normal and offset equal after canonical sign, to 1e-6 / 1e-5 m).
// own_out (may be null): each triangle's own plane (canonical unit normal,
// offset), 4 doubles per triangle, before grouping
void ams_plane_ids(int n_tri, const float *tri, int *plane_out, double *pl_n_out, double *pl_d_out,
                   double *own_out = nullptr);
