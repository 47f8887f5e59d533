// bvh_build.hpp — host construction of the 64-byte wide BVH (pt_layout.hpp).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "pt_layout.hpp"

namespace fnd {

struct Box {
    float lo[3];
    float hi[3];
};

struct WideTree {
    std::vector<WideNode> nodes;   // root at 0; indices local to this tree
    std::vector<uint32_t> order;   // leaf payload order: original primitive indices
    uint32_t max_depth = 0;        // wide levels on the deepest root-to-leaf chain
    Box bounds{};
};

// Binned SAH over primitive boxes (32 bins, leaves of at most `leaf_max`
// primitives), collapsed to 4-wide nodes and quantised.  `threads` workers
// build independent subtrees; the result does not depend on `threads`.
bool build_wide_tree(const Box* boxes, uint32_t n, uint32_t leaf_max, unsigned threads, WideTree& out,
                     std::string& err);

// An object-space box in double precision (exact dequantised planes).
struct DBox {
    double lo[3];
    double hi[3];
};

// Boxes that together contain everything a tree holds, tighter than its root
// box: the dequantised boxes of the subtrees `depth` levels below the root
// (a leaf slot met higher up counts as it is), at most 4^depth boxes.
std::vector<DBox> cover_boxes(const WideTree& t, int depth);

// The box of the mapped corners of every box in `cover` under a row-major 3x4
// map (in double), rounded outward to float and padded by 2^-10 of the
// extent.  A rotated instance's world box then follows its geometry instead
// of its rotated object-space box.
Box transform_cover_conservative(const std::vector<DBox>& cover, const float* m34);

}  // namespace fnd
