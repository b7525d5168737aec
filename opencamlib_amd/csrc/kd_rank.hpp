// Triangle processing order of the reference's KD-tree.
//
// The reference resolves equal-height contacts by evaluation order: triangles
// are visited in the leaf order of its KD-tree, hi child before lo child
// (src/common/kdtree.hpp:273-304), and CLPoint::liftZ only lifts on a strictly
// higher z (src/geo/clpoint.cpp:79-99).  That order does not depend on the
// query, so every triangle has a fixed position ("rank") in it.  The device
// mesh stores triangles sorted by rank; walking candidates in index order then
// reproduces the reference's tie-breaks.
//
// The split rule restated here (kdtree.hpp:141-268): the split dimension is
// the first of the active bbox dimensions with the largest spread (strict >,
// so equal spreads keep the earlier, min-side dimension); the cut is
// min + spread/2; "hi" takes bb[d] > cut; a node becomes a leaf when it holds
// <= bucket triangles or its spread is below 1e-7.  Both partitions keep the
// input order.
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

namespace ocl_amd {

// dims: 4 bbox indices into [minx maxx miny maxy minz maxz]
//   XY (drop) = {0,1,2,3}, YZ (X-fibers) = {2,3,4,5}, XZ (Y-fibers) = {0,1,4,5}
// bb: 6 doubles per triangle.  Returns order[k] = triangle at rank k.
//
// acc (optional): 4 doubles per rank describing when KDTree::search_node
// (kdtree.hpp:273-304) reaches that triangle's leaf.  The search always
// enters the lo child of a min-dimension node (bbox index 0, 2, 4) and the hi
// child of a max-dimension node (1, 3, 5); it enters the hi child of a
// min-dimension cut c only if query[d+1] >= c, and the lo child of a
// max-dimension cut c only if query[d-1] <= c.  Per search axis (dims[0]/2,
// dims[2]/2) the leaf thus needs query_max >= A and query_min <= B, stored
// as (A0, B0, A1, B1).
inline std::vector<uint32_t> kd_rank_order(const double* bb, size_t n, const int dims[4], unsigned bucket,
                                           std::vector<double>* acc = nullptr) {
    std::vector<uint32_t> items(n), tmp(n);
    for (size_t i = 0; i < n; ++i) items[i] = (uint32_t)i;
    if (acc) acc->assign(4 * n, 0.0);
    if (n == 0) return items;
    if (bucket < 1) bucket = 1;
    struct Range { size_t b, e; double A[3], B[3]; };
    // explicit stack; hi ranges are pushed last so they are split first, which
    // keeps the final array in hi-first depth-first leaf order
    std::vector<Range> stack;
    Range root;
    root.b = 0;
    root.e = n;
    for (int a = 0; a < 3; ++a) {
        root.A[a] = -INFINITY;
        root.B[a] = INFINITY;
    }
    stack.push_back(root);
    const int ax0 = dims[0] / 2, ax1 = dims[2] / 2;
    while (!stack.empty()) {
        Range r = stack.back();
        stack.pop_back();
        size_t cnt = r.e - r.b;
        double lo[6] = {0, 0, 0, 0, 0, 0}, hi[6] = {0, 0, 0, 0, 0, 0};
        for (int m = 0; m < 4; ++m) {
            const double v = bb[6 * items[r.b] + dims[m]];
            lo[dims[m]] = v;
            hi[dims[m]] = v;
        }
        for (size_t i = r.b + 1; i < r.e; ++i) {
            const double* t = bb + 6 * items[i];
            for (int m = 0; m < 4; ++m) {
                const double v = t[dims[m]];
                if (hi[dims[m]] < v) hi[dims[m]] = v;
                if (lo[dims[m]] > v) lo[dims[m]] = v;
            }
        }
        double best = 0.0;
        int pick = 0;
        for (int m = 0; m < 4; ++m) {
            const double s = hi[dims[m]] - lo[dims[m]];
            if (s > best) { best = s; pick = m; }
        }
        const int d = dims[pick];
        const double spread = hi[d] - lo[d];
        const double cut = lo[d] + spread / 2;
        if (cnt <= bucket || std::fabs(spread) < 1e-7) {  // leaf: keep order
            if (acc)
                for (size_t i = r.b; i < r.e; ++i) {
                    double* q = acc->data() + 4 * i;
                    q[0] = r.A[ax0];
                    q[1] = r.B[ax0];
                    q[2] = r.A[ax1];
                    q[3] = r.B[ax1];
                }
            continue;
        }
        size_t nh = 0;
        for (size_t i = r.b; i < r.e; ++i)
            if (bb[6 * items[i] + d] > cut) tmp[r.b + nh++] = items[i];
        size_t w = r.b + nh;
        for (size_t i = r.b; i < r.e; ++i)
            if (!(bb[6 * items[i] + d] > cut)) tmp[w++] = items[i];
        for (size_t i = r.b; i < r.e; ++i) items[i] = tmp[i];
        Range rlo = r, rhi = r;
        rlo.b = r.b + nh;
        rhi.e = r.b + nh;
        if (d % 2 == 0) rhi.A[d / 2] = std::fmax(rhi.A[d / 2], cut);  // min-dimension cut
        else rlo.B[d / 2] = std::fmin(rlo.B[d / 2], cut);             // max-dimension cut
        if (rlo.b < rlo.e) stack.push_back(rlo);  // lo, split later
        if (nh > 0) stack.push_back(rhi);         // hi, split next
    }
    return items;
}

}  // namespace ocl_amd
