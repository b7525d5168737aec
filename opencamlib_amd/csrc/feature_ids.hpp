// Feature numbering of a triangle soup (host code): vertices with identical
// coordinates get one ID, edges joining the same two vertex IDs (in either
// direction) get one ID.  A vertex or edge shared by several triangles then
// carries one ID in all of them, so the FP32 selection can tell a second
// copy of the winning contact (the same vertex or edge seen from a
// neighbouring triangle, whose height is the same exact value) from a
// different contact of nearly the same height.  Vertex IDs are 0 .. nv-1,
// edge IDs nv .. nv+ne-1, both below 2^30 (facet IDs, made from the triangle
// rank, have the top bit set).  An edge ID carries kEdgeReversed when the
// triangle runs the edge from its higher vertex ID to its lower one: a
// cutter whose contact on the edge depends on the edge's direction keeps the
// bit, so the two copies of the edge count as different contacts.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace ocl_amd {

constexpr uint32_t kEdgeReversed = 0x40000000u;

// ids: 6 per triangle in the order given (order[r] = input index of rank r,
// or identity when order is null): vertex IDs of p0, p1, p2, then edge IDs
// of (p0,p1), (p1,p2), (p2,p0).
inline std::vector<uint32_t> feature_ids(const double* tri, size_t n, const uint32_t* order) {
    std::vector<uint32_t> ids(6 * n);
    const size_t nv3 = 3 * n;
    auto src = [&](size_t r) { return order ? (size_t)order[r] : r; };
    bool finite = true;
    for (size_t i = 0; i < 9 * n && finite; ++i) finite = std::isfinite(tri[i]);
    // vertices: sort the 3n corner slots by coordinates, number equal runs
    std::vector<uint32_t> vid(nv3);
    if (finite) {
        std::vector<uint32_t> idx(nv3);
        for (size_t i = 0; i < nv3; ++i) idx[i] = (uint32_t)i;
        auto at = [&](uint32_t s) { return tri + 9 * src(s / 3) + 3 * (s % 3); };
        std::sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) {
            const double *p = at(a), *q = at(b);
            if (p[0] != q[0]) return p[0] < q[0];
            if (p[1] != q[1]) return p[1] < q[1];
            return p[2] < q[2];
        });
        uint32_t next = 0;
        for (size_t i = 0; i < nv3; ++i) {
            if (i > 0) {
                const double *p = at(idx[i - 1]), *q = at(idx[i]);
                if (!(p[0] == q[0] && p[1] == q[1] && p[2] == q[2])) ++next;
            }
            vid[idx[i]] = next;
        }
    } else {
        for (size_t i = 0; i < nv3; ++i) vid[i] = (uint32_t)i;  // no sharing across NaN coordinates
    }
    uint32_t nv = 0;
    for (uint32_t v : vid) nv = std::max(nv, v + 1);
    // edges: unordered vertex-ID pairs, numbered after the vertices
    std::vector<uint64_t> key(nv3);
    for (size_t r = 0; r < n; ++r)
        for (int j = 0; j < 3; ++j) {
            const uint64_t a = vid[3 * r + j], b = vid[3 * r + (j + 1) % 3];
            key[3 * r + j] = (std::min(a, b) << 32) | std::max(a, b);
        }
    std::vector<uint32_t> eidx(nv3);
    for (size_t i = 0; i < nv3; ++i) eidx[i] = (uint32_t)i;
    std::sort(eidx.begin(), eidx.end(), [&](uint32_t a, uint32_t b) { return key[a] < key[b]; });
    std::vector<uint32_t> eid(nv3);
    uint32_t next = nv;
    for (size_t i = 0; i < nv3; ++i) {
        if (i > 0 && key[eidx[i]] != key[eidx[i - 1]]) ++next;  // at most 3n + 3n < 2^30 IDs
        eid[eidx[i]] = next;
    }
    for (size_t r = 0; r < n; ++r)
        for (int j = 0; j < 3; ++j) {
            ids[6 * r + j] = vid[3 * r + j];
            const bool rev = vid[3 * r + j] > vid[3 * r + (j + 1) % 3];
            ids[6 * r + 3 + j] = eid[3 * r + j] | (rev ? kEdgeReversed : 0u);
        }
    return ids;
}

}  // namespace ocl_amd
