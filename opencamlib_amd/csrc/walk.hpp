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
template <class T>
__device__ inline T wave_min_t(T v) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
// Min/max over the lane's row: lanes 16r .. 16r + 15 of the wavefront.
template <class T>
__device__ inline T row_min_t(T v) {
    for (int o = 8; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
template <class T>
__device__ inline T row_max_t(T v) {
    for (int o = 8; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline uint32_t wave_max_u(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
    return v;
}
__device__ inline float uniform_f(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ inline uint32_t uniform_u(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

__device__ inline uint32_t lane_id() { return threadIdx.x & 63u; }

// Orders a wavefront's LDS stores before its following LDS loads by other
// lanes of the same wavefront.
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Walk with one vector load per visited node: lanes 0..15 load the node's
// children boxes together, a ballot turns the hits into a 16-bit mask, and
// the children are taken in index order from the mask.  The pending masks of
// all levels live in two 64-bit scalars (16 bits per level), and the first
// child index of a level follows from the level below, so the walk needs no
// stack memory.  `block(base)` runs whenever the walk enters a block of
// level-0 boxes (children base..base+15 of one level-1 node; lane j may
// prefetch data of box base+j); `visit(b, j)` runs for each overlapping
// level-0 box b = base + j.
template <class Block, class Visit>
__device__ inline void walk_boxes(const BoxTree& T, float qx0, float qy0, float qx1, float qy1, Block&& block,
                                  Visit&& visit) {
    const uint32_t lane = lane_id();
    uint64_t mlo = 0, mhi = 0;
    auto get = [&](int l) -> uint32_t {
        return (uint32_t)(((l < 4) ? (mlo >> (16 * l)) : (mhi >> (16 * (l - 4)))) & 0xffffull);
    };
    auto put = [&](int l, uint32_t v) {
        if (l < 4) mlo = (mlo & ~(0xffffull << (16 * l))) | ((uint64_t)v << (16 * l));
        else mhi = (mhi & ~(0xffffull << (16 * (l - 4)))) | ((uint64_t)v << (16 * (l - 4)));
    };
    auto load_mask = [&](int l, uint32_t base) -> uint32_t {
        const uint32_t cnt = T.lvl[2 * l], off = T.lvl[2 * l + 1];
        bool hit = false;
        if (lane < (uint32_t)kFan && base + lane < cnt) hit = box_hits(T.box[off + base + lane], qx0, qy0, qx1, qy1);
        return (uint32_t)__ballot(hit) & 0xffffu;
    };
    const int top = T.top;
    const uint32_t top_cnt = T.lvl[2 * top];
    int l = top;
    uint32_t base = 0;
    put(l, load_mask(l, 0));
    if (l == 0) block(0u);
    for (;;) {
        const uint32_t M = get(l);
        if (M == 0) {
            if (l == top) {
                // the top level may hold more than kFan boxes on huge meshes
                base += kFan;
                if (base >= top_cnt) break;
                put(l, load_mask(l, base));
                if (l == 0) block(base);
                continue;
            }
            base = (base / kFan) & ~(uint32_t)(kFan - 1);
            ++l;
            continue;
        }
        const uint32_t j = (uint32_t)__builtin_ctz(M);
        put(l, M & (M - 1));
        const uint32_t node = base + j;
        if (l == 0) {
            visit(node, j);
            continue;
        }
        --l;
        base = node * kFan;
        put(l, load_mask(l, base));
        if (l == 0) block(base);
    }
}

// The same walk, resumable: visit(b, j) returns false to pause it; calling
// again with the same state continues with the next box.  w.done is set when
// the walk has ended.
struct BoxWalk {
    uint64_t mlo = 0, mhi = 0;
    int l = 0;
    uint32_t base = 0;
    bool started = false, done = false;
};

// hit(i) tests box i of the hierarchy's box arrays (all levels, as T.box)
// against the walk's query: a node is entered, and a level-0 box visited,
// only where it holds.
template <class Hit, class Block, class Visit>
__device__ inline void walk_boxes_resumable_if(const BoxTree& T, Hit&& hit, BoxWalk& w, Block&& block,
                                               Visit&& visit) {
    const uint32_t lane = lane_id();
    auto get = [&](int l) -> uint32_t {
        return (uint32_t)(((l < 4) ? (w.mlo >> (16 * l)) : (w.mhi >> (16 * (l - 4)))) & 0xffffull);
    };
    auto put = [&](int l, uint32_t v) {
        if (l < 4) w.mlo = (w.mlo & ~(0xffffull << (16 * l))) | ((uint64_t)v << (16 * l));
        else w.mhi = (w.mhi & ~(0xffffull << (16 * (l - 4)))) | ((uint64_t)v << (16 * (l - 4)));
    };
    auto load_mask = [&](int l, uint32_t base) -> uint32_t {
        const uint32_t cnt = T.lvl[2 * l], off = T.lvl[2 * l + 1];
        bool h = false;
        if (lane < (uint32_t)kFan && base + lane < cnt) h = hit(off + base + lane);
        return (uint32_t)__ballot(h) & 0xffffu;
    };
    const int top = T.top;
    const uint32_t top_cnt = T.lvl[2 * top];
    if (!w.started) {
        w.started = true;
        w.l = top;
        w.base = 0;
        put(w.l, load_mask(w.l, 0));
        if (w.l == 0) block(0u);
    }
    for (;;) {
        const uint32_t M = get(w.l);
        if (M == 0) {
            if (w.l == top) {
                w.base += kFan;
                if (w.base >= top_cnt) {
                    w.done = true;
                    return;
                }
                put(w.l, load_mask(w.l, w.base));
                if (w.l == 0) block(w.base);
                continue;
            }
            w.base = (w.base / kFan) & ~(uint32_t)(kFan - 1);
            ++w.l;
            continue;
        }
        const uint32_t j = (uint32_t)__builtin_ctz(M);
        put(w.l, M & (M - 1));
        const uint32_t node = w.base + j;
        if (w.l == 0) {
            if (!visit(node, j)) return;
            continue;
        }
        --w.l;
        w.base = node * kFan;
        put(w.l, load_mask(w.l, w.base));
        if (w.l == 0) block(w.base);
    }
}

template <class Block, class Visit>
__device__ inline void walk_boxes_resumable(const BoxTree& T, float qx0, float qy0, float qx1, float qy1, BoxWalk& w,
                                            Block&& block, Visit&& visit) {
    walk_boxes_resumable_if(
        T, [&](uint32_t i) { return box_hits(T.box[i], qx0, qy0, qx1, qy1); }, w, block, visit);
}

template <class Visit>
__device__ inline void walk_boxes(const BoxTree& T, float qx0, float qy0, float qx1, float qy1, Visit&& visit) {
    walk_boxes(T, qx0, qy0, qx1, qy1, [](uint32_t) {}, [&](uint32_t b, uint32_t) { visit(b); });
}

__device__ inline double readlane_d(double v, uint32_t j) {
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, (int)j);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), (int)j);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ inline float readlane_f(float v, uint32_t j) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)j));
}

// Workgroups are handed out round-robin over the 8 XCDs, each with its own
// L2.  xcd_block(g) renumbers the hardware block index so that every XCD runs
// runs of g consecutive logical blocks: neighbouring Morton-ordered CL tiles
// overlap the same triangles, so a run shares one L2, while the runs still
// interleave over the XCDs for balance.  g = 1 is the identity.  The grid
// must be a multiple of kXcds * g (xcd_grid).
constexpr uint32_t kXcds = 8;
__device__ inline uint32_t xcd_block(uint32_t g) {
    const uint32_t x = blockIdx.x % kXcds, r = blockIdx.x / kXcds;
    return (r / g) * (kXcds * g) + x * g + (r % g);
}
__host__ __device__ inline uint32_t xcd_grid(uint32_t blocks, uint32_t g) {
    const uint32_t q = kXcds * g;
    return (blocks + q - 1) / q * q;
}

}  // namespace ocl_amd
