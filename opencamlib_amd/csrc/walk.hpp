// Wavefront-uniform walk of the rank-ordered box hierarchy.
//
// All 64 lanes of a wavefront share one query box (the XY bounding box of
// their CL points grown by the cutter radius), so the walk is scalar control
// flow: the hierarchy boxes are read with uniform addresses, tested with
// uniform compares, and every overlapping level-0 box is handed to `visit`
// in increasing index order, i.e. in the reference's KD evaluation order.
#pragma once

#include "mesh.hpp"

namespace ocl_amd {

__device__ inline bool box_hits(float4 b, float qx0, float qy0, float qx1, float qy1) {
    return !(b.z < qx0 || b.x > qx1 || b.w < qy0 || b.y > qy1);
}

// Wave min/max of a double (inactive lanes pass the neutral element).
__device__ inline double wave_min(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline double wave_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline float uniform_f(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ inline uint32_t uniform_u(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

template <class Visit>
__device__ inline void walk_boxes(const BoxTree& T, float qx0, float qy0, float qx1, float qy1, Visit&& visit) {
    int lvl = T.top;
    uint32_t node = 0;
    for (;;) {
        const uint32_t cnt = T.lvl[2 * lvl], off = T.lvl[2 * lvl + 1];
        if (node < cnt && box_hits(T.box[off + node], qx0, qy0, qx1, qy1)) {
            if (lvl == 0) {
                visit(uniform_u(node));
            } else {
                --lvl;
                node *= kFan;
                continue;
            }
        }
        ++node;
        while (lvl < T.top && (node % kFan == 0 || node >= T.lvl[2 * lvl])) {
            node = (node + kFan - 1) / kFan;
            ++lvl;
        }
        if (lvl == T.top && node >= T.lvl[2 * lvl]) break;
    }
}

}  // namespace ocl_amd
