// Device mesh layout shared by the drop and push kernels.
//
// Triangles live in HBM in the reference's KD evaluation order ("rank"), one
// 128-byte record of 16 doubles each:
//   [0..8]  p0, p1, p2          [9..11] unit normal (Triangle::calcNormal)
//   [12]    max z               [13..15] padding
// so that one wavefront can fetch a record with two 64-byte scalar loads.
// On top of the rank-ordered array sits an implicit 16-ary box hierarchy:
// level 0 = XY bounding box of 16 consecutive triangles, level l+1 = box of
// 16 consecutive level-l boxes.  Because consecutive ranks are leaves of one
// KD subtree, the boxes are tight, and walking the hierarchy in index order
// yields overlapping triangles in exactly the reference's order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ocl_amd {

constexpr int kRec = 16;        // doubles per triangle record
constexpr int kLeaf = 16;       // triangles per level-0 box
constexpr int kFan = 16;        // children per box above level 0
constexpr int kMaxLevels = 8;   // 16^8 triangles is far beyond HBM

struct BoxTree {
    const float4* box;     // all levels, (minx, miny, maxx, maxy), outward rounded
    const uint32_t* lvl;   // lvl[2l] = box count of level l, lvl[2l+1] = its offset in box
    int top;               // index of the highest level (count <= kFan)
};

struct MeshView {
    const double* tri;      // kRec doubles per triangle, rank order
    const uint32_t* order;  // rank -> input index
    uint32_t ntri;
    uint32_t xcd_group;     // consecutive logical blocks per XCD (see xcd_block)
    BoxTree tree;
    const double* bzmax;    // highest vertex per level-0 box (or NULL)
};

}  // namespace ocl_amd
