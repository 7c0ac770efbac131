// Host-side BVH construction for BVHModel<OBBRSS> / BVHModel<RSS>.
#pragma once
#include <cstdint>
#include <vector>

namespace fclx {

enum BVType : int32_t { kOBBRSS = 0, kRSS = 1 };

// One fitted node: OBB and RSS share the covariance eigen-axes.
struct HostNode {
  double obb_axis[9];
  double obb_to[3];
  double obb_extent[3];
  double rss_axis[9];
  double rss_to[3];
  double rss_l[2];
  double rss_r;
  int32_t first_child;
  int32_t first_primitive;
  int32_t num_primitives;
};

struct HostBVH {
  int32_t bv_type = kOBBRSS;
  std::vector<double> vertices;   // 3 per vertex
  std::vector<int32_t> tris;      // 3 per triangle
  std::vector<HostNode> nodes;    // 2*num_tris-1, root 0
  std::vector<uint32_t> primitive_indices;
  int32_t depth = 0;

  // BV::size() of node b for this model's BV type.
  double node_size(int b) const;
};

// Builds the hierarchy the way BVHModel<BV>::endModel() does with the
// default fitter and the mean splitter.  Returns 0, or -1 for an empty model.
int build_bvh(HostBVH& out, int32_t bv_type, const double* vertices, int num_vertices,
              const int32_t* tris, int num_tris);

}  // namespace fclx
