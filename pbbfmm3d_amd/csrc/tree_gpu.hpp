// tree_gpu.hpp -- the tree of one point set built on the GPU (tree_gpu.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "kernels.hpp"

namespace pbb {
namespace dev {

struct GpuLevel {
    int64_t nb = 0, off = 0;  // non-empty boxes; global box offset (levels >= 2)
    int32_t *coord = nullptr, *map = nullptr, *child = nullptr, *order = nullptr;
    int32_t *pstart = nullptr, *seg_box = nullptr, *seg_p0 = nullptr;
    int nseg = 0;
    int up_item0 = 0, up_red0 = 0;  // the level's first direct upward work item and reduction
    int64_t class_count[8] = {0};
};

// Device arrays of one tree (the caller frees them with hipFree); the same
// contents as build_tree plus the plan's derived lists (tree.cpp, capi.cpp).
struct GpuTree {
    int Lv = 0;
    int64_t N = 0;
    double ctr[3] = {0, 0, 0};
    double *x = nullptr, *y = nullptr, *z = nullptr;  // sorted, re-centred
    int32_t *perm = nullptr;                          // sorted position -> input index
    int32_t *pt_leaf = nullptr;                       // packed leaf coordinates of every sorted point
    int32_t *leaf_start = nullptr;                    // leaves + 1
    int32_t *leaf_class = nullptr;                    // leaves grouped by near-field class
    int leaf_class_off[NEAR_CLASSES + 1] = {0};
    int4 *up_items = nullptr, *up_reds = nullptr;
    int n_up_items = 0, n_up_reds = 0, n_up_slots = 0;
    std::vector<GpuLevel> lv;                         // 0..Lv
    int64_t Btot = 0;
    int64_t m2l_pairs = 0, near_pairs = 0, max_leaf = 0, outside = 0;
    int64_t m2l_pairs_level[16] = {0};
};

// Builds the tree of N points (host AoS xyz) on stream st.  ctab: per-axis
// node centres of levels 0..levels-1 (H2_3D_Tree.cpp:939-961), dirs: the far
// directions of each octant class (class_directions).  The bbox centre is
// that of the first nbox points (the targets).  Synchronises st.
void build_tree_gpu(GpuTree &T, const double *xyz_host, int64_t N, int64_t nbox, double L, int levels,
                    const std::vector<std::vector<double>> &ctab, const int dirs[8][189][4], hipStream_t st);

}  // namespace dev
}  // namespace pbb
