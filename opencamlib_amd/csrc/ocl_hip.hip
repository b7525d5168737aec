// MI355X (gfx950) batch drop-cutter: mesh preparation, CL binning, drop
// kernels and the C ABI declared in include/ocl_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/ocl_hip.h"
#include "cutter.hpp"
#include "drop_comp.hpp"
#include "drop_exact.hpp"
#include "drop_fast.hpp"
#include "drop_grad.hpp"
#include "feature_ids.hpp"
#include "kd_rank.hpp"
#include "mesh.hpp"
#include "push_exact.hpp"
#include "radix.hpp"
#include "stl_read.hpp"
#include "walk.hpp"

namespace ocl_amd {

// ------------------------------------------------------------ error state
static thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(OCL_E_HIP, std::string(#expr ": ") + hipGetErrorString(e_));   \
    } while (0)

// --------------------------------------------------------------- buffers
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    int ensure(size_t need) {
        if (need <= bytes) return OCL_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        size_t want = need + need / 4 + 256;
        if (hipMalloc(&p, want) != hipSuccess) return fail(OCL_E_NOMEM, "hipMalloc failed");
        bytes = want;
        return OCL_OK;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// Selection records of one precision: per-triangle records, per-box cull
// caps and the mesh's largest contact error bound (read back once), for one
// cutter.  A mesh caches the set of its last cutter; a batch keeps the sets
// of its jobs alive, so a later call with another cutter builds a new set
// instead of overwriting one a batch replays.
struct RecSet {
    DevBuf rec, cap, emax;
    DevBuf capup;  // highest reachable height per hierarchy box, absolute (k_capup_*)
    DevBuf topv;   // FP32, simple cutters: highest vertex per level-0 box, box frame (k_box_topv)
    bool ok = false;
    CutterDev key{};
    double emax_h = 0.0;
};

// Workspace of the push in one fiber direction: the fibers and their order,
// the interval lists (cap entries per fiber), the row offsets and the
// gathered rows of the last push.
struct PushWork {
    DevBuf fib, perm, pool, off, n, retry, cnt, big, offs, lohi, cc;
    uint64_t pool_cap = 0;   // entries of pool for the current push
    uint32_t rcap = 0;       // list capacity of the last retry pass (stride of big), 0: none
    uint32_t nretry = 0;     // fibers of the last push that needed the retry pass
    uint32_t nf = 0;
    uint64_t total = 0;
};
// cnt of a PushWork: retry count of the first pass, overflow count of a
// retry pass, and the pool entries taken (64-bit)
constexpr size_t kPushCntBytes = 16;
constexpr uint32_t kMaxFiberList = 16384;  // longest interval list a fiber may build

// Workspace of one drop job: the per-point selection records, near-tie
// sets, escalation queues and floors, and the job's counters.  A mesh owns
// one for its single calls and each batch job owns its own, so a batch's
// captured pointers stay valid whatever other calls do on the mesh.
struct DropWork {
    DevBuf esc_list, esc_list2, esc_list3, zfloor, ntset, ntcnt, zband, selrec;
    DevBuf counters, cnt2, cnt3, qcount;
    int exact_slot = 0;  // queue-count slot read by the exact pass in the last run
    int ensure_counters() {
        int rc;
        if ((rc = counters.ensure(sizeof(uint64_t) * 8)) || (rc = cnt2.ensure(sizeof(uint64_t) * 8)) ||
            (rc = cnt3.ensure(sizeof(uint64_t) * 8)) || (rc = qcount.ensure(16)))
            return rc;
        return OCL_OK;
    }
};

}  // namespace ocl_amd

// Box hierarchy over a rank-ordered copy of the triangles (the push trees:
// YZ order with (y, z) boxes for X fibers, XZ order with (x, z) boxes for Y fibers).
struct ocl_tree {
    ocl_amd::DevBuf tri, order, boxes, levels;
    ocl_amd::DevBuf acc;      // per rank: the KD search's leaf acceptance bounds (kd_rank_order)
    ocl_amd::DevBuf kdboxes;  // acceptance boxes, same level layout as boxes
    ocl_amd::DevBuf reach;    // reach boxes (k_box_leaf_reach), same level layout
    uint32_t cnt[ocl_amd::kMaxLevels] = {};
    size_t box_off[ocl_amd::kMaxLevels] = {};
    int top = 0;
    bool built = false;
};

struct ocl_mesh {
    int device = 0;
    std::vector<double> host_tri;  // input triangles (for the push trees, built on first use)
    ocl_tree push_tree[2];
    ocl_amd::PushWork pw[2];       // per fiber direction: the last push of that direction
    int push_last = 0;             // direction of the last push (ocl_hip_push_fetch)
    ocl_amd::DevBuf push_cnt, push_queue, push_sides;
    size_t ntri = 0;
    double bbox[6] = {0, 0, 0, 0, 0, 0};
    unsigned bucket = 1;
    ocl_amd::DevBuf tri;     // kRec doubles per triangle, rank order
    ocl_amd::DevBuf order;   // rank -> input index
    ocl_amd::DevBuf rank_of; // input index -> rank
    ocl_amd::DevBuf boxes;   // all hierarchy levels
    ocl_amd::DevBuf levels;  // (count, offset) per level
    uint32_t cnt[ocl_amd::kMaxLevels] = {};
    size_t box_off[ocl_amd::kMaxLevels] = {};
    int top = 0;
    // per-call workspace
    ocl_amd::DevBuf keys, hist, cl, out_z, out_cc, out_ty, out_tri, out_ft, out_dz;
    ocl_amd::DropWork work;      // selection workspace and counters of the single calls
    ocl_amd::DevBuf origins;     // frame (BoxOrigin) per level-0 box
    ocl_amd::DevBuf fid;         // 6 feature IDs per triangle (rank order; feature_ids.hpp)
    ocl_amd::DevBuf boxzmax;     // highest vertex per level-0 box
    ocl_amd::DevBuf gr_part;     // backward: per-block partial sums of dz/dr
    ocl_amd::DevBuf gt_kv, gt_hist, gt_part;  // backward: per-vertex sums in a fixed order
    double zmid = 0, zhalf = 0, xyb = 0;
    // selection records of the last cutter, per precision (shared with the
    // batches that use them)
    std::shared_ptr<ocl_amd::RecSet> rs32, rs64;
    int batches = 0;             // live batches with a job on this mesh
    uint32_t xcd_group = 8;       // runs of 8 neighbouring CL tiles per XCD (OCL_XCD_GROUP overrides)
    uint32_t hilbert_bits = 9;   // CL binning: Hilbert curve over 2^bits x 2^bits cells
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_s0 = nullptr, ev_s1 = nullptr;

    ocl_amd::MeshView view() const {
        ocl_amd::MeshView v;
        v.tri = tri.as<const double>();
        v.order = order.as<const uint32_t>();
        v.ntri = (uint32_t)ntri;
        v.xcd_group = xcd_group;
        v.tree.box = boxes.as<const float4>();
        v.tree.lvl = levels.as<const uint32_t>();
        v.tree.top = top;
        v.bzmax = boxzmax.as<const double>();
        return v;
    }
};

namespace ocl_amd {

// --------------------------------------------------------- mesh kernels
// Triangle record: vertices, reference normal (triangle.cpp:65-71), max z.
__global__ void k_tri_prep(const double* __restrict__ src, double* __restrict__ rec, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* q = src + 9 * (size_t)i;
    const V3 a = v3(q[0], q[1], q[2]), b = v3(q[3], q[4], q[5]), c = v3(q[6], q[7], q[8]);
    const V3 nrm = vunit(vcross(vsub(a, b), vsub(a, c)));
    double* r = rec + (size_t)kRec * i;
    for (int k = 0; k < 9; ++k) r[k] = q[k];
    r[9] = nrm.x;
    r[10] = nrm.y;
    r[11] = nrm.z;
    r[12] = fmax(fmax(a.z, b.z), c.z);
    r[13] = 0.0;
    r[14] = 0.0;
    r[15] = 0.0;
}

// Push trees: the pair-bound record (push_bound_rec) in the record's
// padding, for fibers along `axis`.
__global__ void k_push_boundrec(double* __restrict__ rec, uint32_t n, int axis) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double* r = rec + (size_t)kRec * i;
    double b[3];
    push_bound_rec(r, axis, b);
    r[13] = b[0];
    r[14] = b[1];
    r[15] = b[2];
}

// Level-0 boxes: XY box of kLeaf consecutive triangles, rounded outward.
// Highest vertex of each level-0 box (the exact path's box-level cap).
__global__ void k_box_zmax(const double* __restrict__ rec, uint32_t ntri, double* __restrict__ out, uint32_t nbox) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    double z = -INFINITY;
    const uint32_t e = min(ntri, (b + 1) * kLeaf);
    for (uint32_t i = b * kLeaf; i < e; ++i) z = fmax(z, rec[(size_t)kRec * i + 12]);
    out[b] = z;
}

__global__ void k_box_leaf(const double* __restrict__ rec, uint32_t ntri, float4* __restrict__ out, uint32_t nbox) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    double x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY;
    const uint32_t e = min(ntri, (b + 1) * kLeaf);
    for (uint32_t i = b * kLeaf; i < e; ++i) {
        const double* r = rec + (size_t)kRec * i;
        for (int k = 0; k < 3; ++k) {
            x0 = fmin(x0, r[3 * k]);
            x1 = fmax(x1, r[3 * k]);
            y0 = fmin(y0, r[3 * k + 1]);
            y1 = fmax(y1, r[3 * k + 1]);
        }
    }
    out[b] = make_float4(__double2float_rd(x0), __double2float_rd(y0), __double2float_ru(x1), __double2float_ru(y1));
}

// Level-0 boxes over (coordinate cu, z) for the push trees: float4(u0, z0, u1, z1).
__global__ void k_box_leaf_uz(const double* __restrict__ rec, uint32_t ntri, float4* __restrict__ out, uint32_t nbox,
                              int cu) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    double u0 = INFINITY, z0 = INFINITY, u1 = -INFINITY, z1 = -INFINITY;
    const uint32_t e = min(ntri, (b + 1) * kLeaf);
    for (uint32_t i = b * kLeaf; i < e; ++i) {
        const double* r = rec + (size_t)kRec * i;
        for (int k = 0; k < 3; ++k) {
            u0 = fmin(u0, r[3 * k + cu]);
            u1 = fmax(u1, r[3 * k + cu]);
            z0 = fmin(z0, r[3 * k + 2]);
            z1 = fmax(z1, r[3 * k + 2]);
        }
    }
    out[b] = make_float4(__double2float_rd(u0), __double2float_rd(z0), __double2float_ru(u1), __double2float_ru(z1));
}

// Level-0 "reach boxes" of the push trees for the pair bound (push_bound):
// the extent across the fiber and the highest vertex of the box's
// well-conditioned triangles (x = lowest across, z = highest across, w =
// highest vertex, y unused = -inf); an ill-conditioned triangle, which gets
// no bound, makes the box reach everything.  A box none of whose triangles
// can add an interval to a fiber fails the fiber's reach query.
__global__ void k_box_leaf_reach(const double* __restrict__ rec, uint32_t ntri, float4* __restrict__ out,
                                 uint32_t nbox, int cu) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    double u0 = INFINITY, u1 = -INFINITY, z1 = -INFINITY;
    const uint32_t e = min(ntri, (b + 1) * kLeaf);
    for (uint32_t i = b * kLeaf; i < e; ++i) {
        const double* r = rec + (size_t)kRec * i;
        if (r[13] == 0.0) {
            u0 = -INFINITY;
            u1 = INFINITY;
            z1 = INFINITY;
        }
        for (int k = 0; k < 3; ++k) {
            u0 = fmin(u0, r[3 * k + cu]);
            u1 = fmax(u1, r[3 * k + cu]);
        }
        z1 = fmax(z1, r[12]);
    }
    out[b] = make_float4(__double2float_rd(u0), -INFINITY, __double2float_ru(u1), __double2float_ru(z1));
}

// Level-0 "acceptance boxes" of the push trees: the KD search reaches a
// triangle when its query satisfies max >= A and min <= B per axis
// (kd_rank_order), i.e. when the query interval meets [A, B]; per box the
// union [min A, max B] (rounded outward) makes the same test a box overlap.
__global__ void k_kdbox_leaf(const double* __restrict__ acc, uint32_t ntri, float4* __restrict__ out,
                             uint32_t nbox) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    double a0 = INFINITY, a1 = INFINITY, b0 = -INFINITY, b1 = -INFINITY;
    const uint32_t e = min(ntri, (b + 1) * kLeaf);
    for (uint32_t i = b * kLeaf; i < e; ++i) {
        const double* q = acc + 4 * (size_t)i;
        a0 = fmin(a0, q[0]);
        b0 = fmax(b0, q[1]);
        a1 = fmin(a1, q[2]);
        b1 = fmax(b1, q[3]);
    }
    out[b] = make_float4(__double2float_rd(a0), __double2float_rd(a1), __double2float_ru(b0), __double2float_ru(b1));
}

__global__ void k_box_up(const float4* __restrict__ in, uint32_t nin, float4* __restrict__ out, uint32_t nout) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nout) return;
    float4 r = make_float4(INFINITY, INFINITY, -INFINITY, -INFINITY);
    const uint32_t e = min(nin, (b + 1) * kFan);
    for (uint32_t i = b * kFan; i < e; ++i) {
        const float4 c = in[i];
        r.x = fminf(r.x, c.x);
        r.y = fminf(r.y, c.y);
        r.z = fmaxf(r.z, c.z);
        r.w = fmaxf(r.w, c.w);
    }
    out[b] = r;
}

// Absolute caps of the hierarchy boxes (FP32 selection): level 0 is the
// box's cap (highest reachable height, relative to its frame) plus the
// frame's height, rounded up; a higher box takes the largest of its
// children.
template <class T>
__global__ void k_capup_leaf(const T* __restrict__ cap, const BoxOrigin* __restrict__ org, uint32_t nbox,
                             float* __restrict__ out) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    const T c = cap[b];
    out[b] = (c == -INFINITY) ? -INFINITY : __double2float_ru((double)c + org[b].oz);
}
__global__ void k_capup_up(const float* __restrict__ in, uint32_t nin, float* __restrict__ out, uint32_t nout) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nout) return;
    float r = -INFINITY;
    const uint32_t e = min(nin, (b + 1) * kFan);
    for (uint32_t i = b * kFan; i < e; ++i) r = fmaxf(r, in[i]);
    out[b] = r;
}

// Highest vertex of each level-0 box in the box frame, from its FP32
// selection records (z = -inf for an empty box): the selection's vertex
// floor takes one vertex per box near the wavefront's points.
__global__ void k_box_topv(const float* __restrict__ rec, uint32_t ntri, uint32_t nbox, float4* __restrict__ out) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    float4 v = make_float4(0.0f, 0.0f, -INFINITY, 0.0f);
    const uint32_t e = min(ntri, (b + 1) * kLeaf);
    for (uint32_t i = b * kLeaf; i < e; ++i) {
        const float* r = rec + (size_t)rec_len<float>() * i + R_P0X;
        for (int k = 0; k < 3; ++k)
            if (r[3 * k + 2] > v.z) v = make_float4(r[3 * k], r[3 * k + 1], r[3 * k + 2], 0.0f);
    }
    out[b] = v;
}

// Level-0 box centres (local frame origins of the selection records).
__global__ void k_box_origin(const float4* __restrict__ box, uint32_t nbox, double2* __restrict__ org) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    const float4 q = box[b];
    org[b] = make_double2(0.5 * ((double)q.x + (double)q.z), 0.5 * ((double)q.y + (double)q.w));
}

// ------------------------------------------------------------ CL binning
// CL points are grouped so that each wavefront holds 64 points that are close
// in XY: a Hilbert key over the mesh box grown by the cutter radius, then a
// stable radix sort by key, so equal keys keep their input order and the
// point -> wavefront assignment is the same on every run.
struct BinFrame {
    double x0, y0, sx, sy;  // cell = (clamp((x - x0) * sx), clamp((y - y0) * sy)) on a 2^hb x 2^hb grid
};

// Position of cell (x, y) along the Hilbert curve through the 2^hb x 2^hb
// grid (the classic iterative xy -> d conversion: per level, take the
// quadrant and rotate/reflect the remaining bits into its frame).  Successive
// keys are always edge-adjacent cells, so 64 consecutive CL points of the
// sorted order form one compact patch: no wavefront straddles a jump across
// the domain, as a Morton order's quadrant boundaries would make it.
__device__ inline uint32_t hilbert_key(uint32_t x, uint32_t y, uint32_t hb) {
    const uint32_t side = 1u << hb;
    uint32_t d = 0;
    for (uint32_t s = side >> 1; s > 0; s >>= 1) {
        const uint32_t rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
        d += s * s * ((3u * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = side - 1u - x;
                y = side - 1u - y;
            }
            const uint32_t t = x;
            x = y;
            y = t;
        }
    }
    return d;
}

// Up to four small device ranges (counters, queue counts) that a call clears
// at its start, from the first block of its first kernel: one fill kernel
// per range would cost a launch each.
struct ZeroSet {
    uint32_t* p[4];
    uint32_t words[4];
};

__global__ void k_cl_keys(const double* __restrict__ cl, uint32_t n, BinFrame f, uint32_t hb,
                          uint32_t* __restrict__ keys, uint32_t* __restrict__ vals, ZeroSet zs) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x == 0) {
        for (int r = 0; r < 4; ++r)
            if (zs.p[r] && threadIdx.x < zs.words[r]) zs.p[r][threadIdx.x] = 0u;
    }
    if (i >= n) return;
    const double cmax = (double)((1u << hb) - 1u);
    const double fx = (cl[3 * (size_t)i] - f.x0) * f.sx, fy = (cl[3 * (size_t)i + 1] - f.y0) * f.sy;
    // NaN coordinates land in cell 0 (fmax/fmin drop the NaN)
    const uint32_t ix = (uint32_t)fmin(fmax(fx, 0.0), cmax), iy = (uint32_t)fmin(fmax(fy, 0.0), cmax);
    keys[i] = hilbert_key(ix, iy, hb);
    vals[i] = i;
}



// ------------------------------------------------------------ drop outputs
struct DropOut {
    double* z;
    double* cc;
    uint8_t* type;
    uint32_t* tri;
    uint8_t* feat;
};

struct Counters {
    unsigned long long pairs, escalated, brent, nesc, evals;  // evals: wavefront runs of the contact code
};
static_assert(sizeof(Counters) <= 64, "DropWork::ensure_counters sizes the counters at 64 bytes");

__device__ inline TriX load_tri(const double* __restrict__ base, uint32_t k) {
    const double* r = base + (size_t)kRec * k;
    TriX t;
    t.p[0] = v3(r[0], r[1], r[2]);
    t.p[1] = v3(r[3], r[4], r[5]);
    t.p[2] = v3(r[6], r[7], r[8]);
    t.n = v3(r[9], r[10], r[11]);
    t.maxz = r[12];
    return t;
}

__device__ inline void wave_add(unsigned long long* dst, unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(dst, v);
}

// Appends v of every lane with `push` to queue q (count *qn), one atomic per
// wavefront; lanes keep their order within the wavefront.
__device__ inline void wave_push(bool push, uint32_t v, uint32_t* q, uint32_t* qn) {
    const unsigned long long bal = __ballot(push);
    if (!bal) return;
    const uint32_t lead = (uint32_t)(__ffsll((long long)bal) - 1);
    uint32_t base = 0;
    if ((threadIdx.x & 63) == lead) base = atomicAdd(qn, (uint32_t)__popcll(bal));
    base = __shfl(base, (int)lead, 64);
    if (push) q[base + __popcll(bal & ((1ull << (threadIdx.x & 63)) - 1ull))] = v;
}

__device__ inline void write_drop(const DropOut& o, const uint32_t* order, uint32_t src, const DropLane& L) {
    o.z[src] = L.z;
    if (o.cc) {
        o.cc[3 * (size_t)src] = L.cc.x;
        o.cc[3 * (size_t)src + 1] = L.cc.y;
        o.cc[3 * (size_t)src + 2] = L.cc.z;
    }
    if (o.type) o.type[src] = (uint8_t)L.type;
    if (o.tri) o.tri[src] = (L.tri == 0xffffffffu) ? 0xffffffffu : order[L.tri];
    if (o.feat) o.feat[src] = (uint8_t)L.feat;
}

__device__ inline DropLane fresh_lane(double x, double y, double z) {
    DropLane L;
    L.x = x;
    L.y = y;
    L.z = z;
    L.cc = v3(0, 0, 0);
    L.type = CC_NONE;
    L.tri = 0xffffffffu;
    L.feat = FT_NONE;
    L.cur_tri = 0;
    L.cur_feat = 0;
    return L;
}

// Profile height of the cutter at XY distance sqrt(d2) from its axis,
// rounded down: every contact whose contact point lies that far from the
// axis sits at least this much below the contact point (Cyl/Ball/Bull; cones
// as prof_lower, when their length is not negative).  0 where no bound holds.
__device__ inline double prof_lower64(const CutterP& c, double d2) {
    double h = 0.0;
    if (c.kind == K_BALL) {
        h = c.radius - sqrt(fmax(c.radius * c.radius - d2, 0.0));
    } else if (c.kind == K_BULL) {
        const double w = fmax(sqrt(d2) - c.r1, 0.0);
        h = c.r2 - sqrt(fmax(c.r2 * c.r2 - w * w, 0.0));
    } else if (c.kind == K_CONE && c.length >= 0.0) {
        h = fmin(sqrt(d2) / c.tan_a, c.length);
    }
    return fmax(h - 1e-12 * fmax(c.radius, h), 0.0);
}

// Cutters whose contacts are bounded by max z - prof_lower64(distance).
__device__ inline bool has_contact_cap(const CutterP& c) {
    return c.kind == K_CYL || c.kind == K_BALL || c.kind == K_BULL || (c.kind == K_CONE && c.length >= 0.0);
}

// Squared XY distance from (x, y) to triangle t's XY box.
__device__ inline double tri_box_d2(const TriX& t, double x, double y) {
    const double bx0 = fmin(fmin(t.p[0].x, t.p[1].x), t.p[2].x), bx1 = fmax(fmax(t.p[0].x, t.p[1].x), t.p[2].x);
    const double by0 = fmin(fmin(t.p[0].y, t.p[1].y), t.p[2].y), by1 = fmax(fmax(t.p[0].y, t.p[1].y), t.p[2].y);
    const double dx = fmax(fmax(bx0 - x, x - bx1), 0.0), dy = fmax(fmax(by0 - y, y - by1), 0.0);
    return dx * dx + dy * dy;
}

// Upper bound of every contact height of triangle t for a CL point at
// (x, y): max z minus the cutter profile height at the XY distance to the
// triangle's box.
__device__ inline double contact_cap(const CutterP& c, const TriX& t, double x, double y) {
    return t.maxz - prof_lower64(c, tri_box_d2(t, x, y));
}

// Composite cutters (compositecutter.cpp:150-194 and the vertex drop): the
// contact of sub-cutter n is accepted only when its contact point lies in
// the radius window [rlim[n-1], rlim[n]] (1e-6 tolerance), and it sits
// h_n(r) + zoff[n] below the contact point, h_n being the sub-cutter's
// profile (non-decreasing in r).  So every accepted contact of a triangle at
// XY box distance d sits at least min over the windows reaching d of
// h_n(max(d, window start)) + zoff[n] below the triangle's max z; with
// negative offsets this may be negative (contacts above max z).
__device__ inline bool comp_has_cap(const CutterDev& cd) {
    if (cd.c.kind != K_COMPOSITE || cd.nsub < 1) return false;
    for (int n = 0; n < cd.nsub; ++n)
        if (!has_contact_cap(cd.sub[n])) return false;
    return true;
}
__device__ inline double comp_prof_lower64(const CutterDev& cd, double d2) {
    const double d = sqrt(d2);
    double h = INFINITY;
    for (int n = 0; n < cd.nsub; ++n) {
        const double lo = (n == 0) ? 0.0 : cd.rlim[n - 1] - 1e-6, hi = cd.rlim[n] + 1e-6;
        if (d > hi) continue;
        const double r = fmax(d, lo);
        double hn = prof_lower64(cd.sub[n], r * r) + cd.zoff[n];
        // a vertex beyond a CylCutter's or BullCutter's own radius but inside
        // its window's tolerance gets the height -1 (prof_height)
        const CutterP& s = cd.sub[n];
        if ((s.kind == K_CYL || s.kind == K_BULL) && s.radius < hi) hn = fmin(hn, -1.0 + cd.zoff[n]);
        h = fmin(h, hn - 1e-12 * (cd.c.radius + fabs(hn)));
    }
    return h;  // +inf: no window reaches d, the triangle has no accepted contact
}

// Exact FP64 drop over a list of points (identity list or escalation list).
// One lane per point (gid = position in the list); the wavefront walks the
// hierarchy with the union box.
__device__ __attribute__((always_inline)) inline void exact_points(const MeshView& m, const CutterDev& cd,
                                                                  const double* __restrict__ cl,
                                                                  const uint32_t* __restrict__ list, uint32_t n_eff,
                                                                  const DropOut& o, Counters* cnt, uint32_t gid) {
    const uint32_t wave_first = gid & ~63u;
    if (wave_first >= n_eff) return;
    const bool act = gid < n_eff;
    const uint32_t src = act ? list[gid] : 0;
    const double x = act ? cl[3 * (size_t)src] : 0.0;
    const double y = act ? cl[3 * (size_t)src + 1] : 0.0;
    const double z = act ? cl[3 * (size_t)src + 2] : 0.0;
    const double R = cd.c.radius;
    const float qx0 = uniform_f(__double2float_rd(wave_min(act ? x : INFINITY) - R));
    const float qy0 = uniform_f(__double2float_rd(wave_min(act ? y : INFINITY) - R));
    const float qx1 = uniform_f(__double2float_ru(wave_max(act ? x : -INFINITY) + R));
    const float qy1 = uniform_f(__double2float_ru(wave_max(act ? y : -INFINITY) + R));
    DropLane L = fresh_lane(x, y, z);
    unsigned long long pairs = 0, brent = 0;
    // a triangle whose contacts all stay below the running height cannot
    // lift, and in the reference's fold a triangle that does not lift
    // changes nothing: skipped
    const bool comp = cd.c.kind == K_COMPOSITE;
    const bool capped = comp ? comp_has_cap(cd) : has_contact_cap(cd.c);
    walk_boxes(m.tree, qx0, qy0, qx1, qy1, [&](uint32_t b) {
        if (capped && m.bzmax) {
            // the same bound for the whole box: its highest vertex at the
            // distance to its (outward rounded) XY box
            const float4 bb = m.tree.box[m.tree.lvl[1] + b];
            const double dx = fmax(fmax((double)bb.x - x, x - (double)bb.z), 0.0);
            const double dy = fmax(fmax((double)bb.y - y, y - (double)bb.w), 0.0);
            const double d2 = dx * dx + dy * dy;
            const double cap = m.bzmax[b] - (comp ? comp_prof_lower64(cd, d2) : prof_lower64(cd.c, d2));
            if (!__any(act && cap >= L.z - 1e-9 * (1.0 + fabs(L.z)))) return;
        }
        const uint32_t e = min(m.ntri, (b + 1) * kLeaf);
        for (uint32_t k = b * kLeaf; k < e; ++k) {
            const uint32_t ku = uniform_u(k);
            const TriX t = load_tri(m.tri, ku);
            if (act) {
                L.cur_tri = ku;
                if (cutter_overlaps(R, L.x, L.y, t)) {
                    if (L.z < t.maxz) ++pairs;
                    const double cap = !capped ? INFINITY
                                       : t.maxz - (comp ? comp_prof_lower64(cd, tri_box_d2(t, L.x, L.y))
                                                        : prof_lower64(cd.c, tri_box_d2(t, L.x, L.y)));
                    if (cap >= L.z - 1e-9 * (1.0 + fabs(L.z))) brent += drop_triangle_exact(cd, L, t);
                }
            }
        }
    });
    if (act) write_drop(o, m.order, src, L);
    if (cnt) {
        wave_add(&cnt->pairs, pairs);
        wave_add(&cnt->brent, brent);
    }
}

// One wavefront per block, as the selection kernel: per-point work varies,
// and a finished wavefront frees its slot at once.  XCD runs of
// xcd_group x 4 wavefronts.
__global__ void __launch_bounds__(64) k_drop_exact(MeshView m, CutterDev cd, const double* __restrict__ cl,
                                                   const uint32_t* __restrict__ list,
                                                   const uint32_t* __restrict__ count, uint32_t n, DropOut o,
                                                   Counters* cnt) {
    exact_points(m, cd, cl, list, count ? *count : n, o, cnt, xcd_block(m.xcd_group * 4u) * 64u + threadIdx.x);
}

// The exact-path jobs of a batch (composites, or every job in EXACT64
// mode) in one launch: wavefront c takes 64-point chunk c of the
// concatenation of the jobs' chunks (chunk_off: prefix sums, largest jobs
// first), so small jobs run beside the large ones instead of one after
// another.
struct ExactJob {
    MeshView m;
    CutterDev cd;
    const double* cl;
    const uint32_t* list;
    uint32_t n;
    DropOut o;
    Counters* cnt;
};

__global__ void __launch_bounds__(64) k_drop_exact_jobs(const ExactJob* __restrict__ jobs,
                                                        const uint32_t* __restrict__ chunk_off, uint32_t njobs) {
    const uint32_t c = uniform_u(blockIdx.x);
    if (c >= chunk_off[njobs]) return;
    uint32_t lo = 0, hi = njobs;  // the job: last j with chunk_off[j] <= c
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (chunk_off[mid] <= c) lo = mid;
        else hi = mid;
    }
    const ExactJob& J = jobs[lo];
    exact_points(J.m, J.cd, J.cl, J.list, J.n, J.o, J.cnt, (c - chunk_off[lo]) * 64u + lane_id());
}


// ----------------------------------------------------- selection kernels
// Per-triangle selection records for one cutter, in T, box-local frames.
template <class T>
__global__ void k_sel_prep(const double* __restrict__ tri, uint32_t ntri, const BoxOrigin* __restrict__ org,
                           const uint32_t* __restrict__ fid, CutterDev cd, SelErr er, T* __restrict__ rec) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntri) return;
    const TriX t = load_tri(tri, i);
    const BoxOrigin o = org[i / kLeaf];
    build_record<T>(t, cd, o.ox, o.oy, o.oz, er, rec + (size_t)rec_len<T>() * i, fid + 6 * (size_t)i);
}

// The two sub-cutter records of each triangle for a composite cutter (FP32).
__global__ void k_sel_prep_comp(const double* __restrict__ tri, uint32_t ntri, const BoxOrigin* __restrict__ org,
                                const uint32_t* __restrict__ fid, CutterDev cd, SelErr er, float* __restrict__ rec) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntri) return;
    const TriX t = load_tri(tri, i);
    const BoxOrigin o = org[i / kLeaf];
    build_record_comp<float>(t, cd, o.ox, o.oy, o.oz, er, rec + (size_t)(2 * rec_len<float>()) * i,
                             fid + 6 * (size_t)i);
}

// Per level-0 box: the largest cull cap of its records (box frame), and the
// largest contact error bound of the mesh (atomic max of non-negative float
// bits, which order like unsigned integers).  nrec: records per triangle
// (2 for a composite cutter's sub-cutters).
template <class T>
__global__ void k_box_cap(const T* __restrict__ rec, uint32_t ntri, uint32_t nbox, uint32_t nrec,
                          T* __restrict__ cap, unsigned int* __restrict__ emax_bits) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    T c = -FM<T>::inf();
    float em = 0.0f;
    const uint32_t e = min(ntri, (b + 1) * kLeaf);
    for (uint32_t i = b * kLeaf; i < e; ++i)
        for (uint32_t s = 0; s < nrec; ++s) {
            const T* r = rec + (size_t)rec_len<T>() * (nrec * i + s);
            c = FM<T>::max_(c, r[R_ZCAP]);
            em = fmaxf(em, __double2float_ru((double)r[R_EMAX]));
        }
    cap[b] = c;
    atomicMax(emax_bits, __float_as_uint(em));
}

// Box origins: xy centre of the level-0 box, mid z of its triangles.
__global__ void k_box_frames(const double* __restrict__ tri, uint32_t ntri, const float4* __restrict__ box,
                             uint32_t nbox, BoxOrigin* __restrict__ org) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    const float4 q = box[b];
    double z0 = INFINITY, z1 = -INFINITY;
    const uint32_t e = min(ntri, (b + 1) * kLeaf);
    for (uint32_t i = b * kLeaf; i < e; ++i) {
        const double* r = tri + (size_t)kRec * i;
        for (int k = 0; k < 3; ++k) {
            z0 = fmin(z0, r[3 * k + 2]);
            z1 = fmax(z1, r[3 * k + 2]);
        }
    }
    BoxOrigin o;
    o.ox = 0.5 * ((double)q.x + (double)q.z);
    o.oy = 0.5 * ((double)q.y + (double)q.w);
    o.oz = 0.5 * (z0 + z1);
    o.hx = __double2float_ru(0.5 * ((double)q.z - (double)q.x));
    o.hy = __double2float_ru(0.5 * ((double)q.w - (double)q.y));
    org[b] = o;
}

constexpr int kNearTie = 8;  // near-tie set capacity per lane

// Per CL point (escalated ones): triangles (rank indices) whose contacts
// reach the band below the selection's best, their count (0xff = none
// usable) and the bottom of the band.
struct NearTieSets {
    uint32_t* set;   // kNearTie per point
    uint8_t* cnt;    // status: near-tie set size, kSelDecided or kSelNone
    double* zband;
    uint32_t* bt;    // selection winner (rank index, feature), its height and error bound
    uint8_t* bf;
    double* zsel;
    float* esel;
    float* zgap;     // FP32: the highest upper value of the other contacts, less zsel
                     // (rounded up; -inf: none, or an FP64 selection)
    uint32_t* q;     // queue of the points with a near-tie set (count *qn),
    uint32_t* qn;    // filled by the selection pass
};
constexpr uint8_t kSelDecided = 0xfe;  // winner decided, refine it
constexpr uint8_t kSelNone = 0xff;     // undecided, no usable near-tie set

template <class T>
struct WaveState {
    T zb, zt[4], za, zr;
};

// Arguments of one selection pass (one job of a batch).
template <class T>
struct SelArgs {
    MeshView m;
    const T* rec;
    const BoxOrigin* org;
    const T* boxcap;
    const float* capup;      // FP32: absolute cap per hierarchy box (all levels), or NULL
    const float4* topv;      // FP32: highest vertex per level-0 box (box frame), or NULL
    SelParams<T> P;
    double R;                // cutter radius
    const double* cl;
    const uint32_t* list;    // points in selection order
    const uint32_t* count;   // point count in device memory (or NULL: n)
    uint32_t n;
    DropOut o;
    double* zfloor;
    NearTieSets nts;
    int refine;
    Counters* cnt;
};

// Selection over a list of points, one wavefront of 64 neighbouring points
// (gid = position in the list).  Writes final results for lanes that need no
// escalation and the selection records of the others.  refine = 0 is the
// FP32-only mode: the selection result is written as is.
//   my:   the wavefront's LDS copy of the records of the level-0 box being
//         evaluated (kLeaf * kFastRec values)
//   nt_k, nt_h: per-lane near-tie sets: rank indices (and upper values, wave
//         frame) of the triangles whose contacts reach the band below the
//         lane's known height (in evaluation order, not rank order)
//   bl_b, bl_key: the wavefront's list of level-0 boxes and their sort keys
//         (kBoxList entries)
//   ptab: the wavefront's table of the (owner lane, triangle) pairs of the
//         batch being evaluated (64 entries)
// Level-0 boxes a wavefront sorts before evaluating them (a power of two:
// the bitonic sort runs over the list rounded up to one).  With the walk
// pruned by the hierarchy caps, 32 beat 64 and 128 on the 1M-point Ball
// workload (select 1.355 against 1.395 and 1.395 ms; BullCutter unchanged);
// before the pruning 64 had beaten 32, 128 and 256.
constexpr uint32_t kBoxList = 32;

// Pair packing: the (point, triangle) pairs of a level-0 box are evaluated
// 64 to a wavefront step, each lane for the pair's owner, and folded back by
// the owners.  It pays where a pair's contact code costs far more than the
// fold (the FP64 BullCutter edge solve); for the other selections the
// rank-order scan, one triangle per step for the lanes that need it, is
// faster.  Cones always scan: their fold depends on the order.
template <class T, int K>
constexpr bool pack_pairs() {
    return K == K_BULL && sizeof(T) == 8;
}

template <class T, int K>
__device__ __attribute__((always_inline)) inline void select_points(const SelArgs<T>& A, uint32_t gid, T* my,
                                                                   uint32_t (*nt_k)[64], float (*nt_h)[64],
                                                                   uint32_t* bl_b, float* bl_key, uint32_t* ptab) {
    typedef FM<T> F;
    const MeshView& m = A.m;
    const T* __restrict__ rec = A.rec;
    const BoxOrigin* __restrict__ org = A.org;
    const T* __restrict__ boxcap = A.boxcap;
    const SelParams<T>& P = A.P;
    const double* __restrict__ cl = A.cl;
    const uint32_t* __restrict__ list = A.list;
    const DropOut& o = A.o;
    double* __restrict__ zfloor = A.zfloor;
    const NearTieSets& nts = A.nts;
    const int refine = A.refine;
    Counters* cnt = A.cnt;
    const uint32_t ln = lane_id();
    uint32_t ntn = 0;
    bool ovf = false;
    const uint32_t n_eff = A.count ? *A.count : A.n;
    const uint32_t wave_first = gid & ~63u;
    if (wave_first >= n_eff) return;
    const bool act = gid < n_eff;
    const uint32_t src = act ? list[gid] : 0;
    const double x = act ? cl[3 * (size_t)src] : 0.0;
    const double y = act ? cl[3 * (size_t)src + 1] : 0.0;
    const double z0 = act ? cl[3 * (size_t)src + 2] : 0.0;
    const double R = A.R;
    const float qx0 = uniform_f(__double2float_rd(wave_min(act ? x : INFINITY) - R));
    const float qy0 = uniform_f(__double2float_rd(wave_min(act ? y : INFINITY) - R));
    const float qx1 = uniform_f(__double2float_ru(wave_max(act ? x : -INFINITY) + R));
    const float qy1 = uniform_f(__double2float_ru(wave_max(act ? y : -INFINITY) + R));
    Sel<T> S;
    WaveState<T> W;
    uint32_t evals = 0;
    double ozw = 0.0;
    bool framed = false;
    sel_init(S, (T)0);
    W.zb = (T)0;
    for (int i = 0; i < 4; ++i) W.zt[i] = -F::inf();
    W.za = -F::inf();
    W.zr = -F::inf();
    // XY hull of the wave's points
    const double hull_x0 = wave_min(act ? x : INFINITY), hull_x1 = wave_max(act ? x : -INFINITY);
    const double hull_y0 = wave_min(act ? y : INFINITY), hull_y1 = wave_max(act ? y : -INFINITY);
    // XY hull of the points of the lane's row of 16 (neighbours in the
    // selection order; +-inf for a row without points)
    const double row_x0 = row_min_t<double>(act ? x : INFINITY), row_x1 = row_max_t<double>(act ? x : -INFINITY);
    const double row_y0 = row_min_t<double>(act ? y : INFINITY), row_y1 = row_max_t<double>(act ? y : -INFINITY);
    // Floor of every lane's final height from the vertices of the triangles
    // right under the wavefront's points (boxes overlapping their XY hull).
    T zc = -F::inf();
    bool top_floor = false;
    if constexpr (sizeof(T) == 4) top_floor = K != K_CONE && P.cull && A.topv != nullptr;
    if constexpr (sizeof(T) == 4) if (top_floor) {
        // FP32: the highest vertex of each box within R/4 of the points'
        // hull (one vertex per box: on the 1M-point workloads as good a
        // floor as every vertex of the boxes under the hull, at a fraction
        // of its cost; R/4 beat 0 and R/2 on the BullCutter run)
        const double fe = 0.25 * R;
        const float tx0 = uniform_f(__double2float_rd(hull_x0 - fe));
        const float ty0 = uniform_f(__double2float_rd(hull_y0 - fe));
        const float tx1 = uniform_f(__double2float_ru(hull_x1 + fe));
        const float ty1 = uniform_f(__double2float_ru(hull_y1 + fe));
        // The walk only lists the boxes (up to kBoxList at a time); their
        // frames and vertices are then loaded together, one box per lane,
        // into LDS (the record staging area, free at this point), instead of
        // one dependent load after another.
        double* fo = reinterpret_cast<double*>(my);            // ox, oy per box
        float* fv = reinterpret_cast<float*>(my) + 4 * kBoxList;  // vx, vy, vz, dz per box
        static_assert(8 * kBoxList * sizeof(float) <= kLeaf * rec_len<float>() * sizeof(float), "floor staging");
        BoxWalk fw;
        for (;;) {
            uint32_t nb = 0;
            walk_boxes_resumable(m.tree, tx0, ty0, tx1, ty1, fw, [](uint32_t) {}, [&](uint32_t b, uint32_t) -> bool {
                if (lane_id() == 0) bl_b[nb] = b;
                ++nb;
                return nb < kBoxList;
            });
            if (nb) {
                wave_lds_sync();
                if (!framed) {
                    framed = true;
                    ozw = org[uniform_u(bl_b[0])].oz;
                    W.zb = (T)(z0 - ozw);
                }
                if (ln < nb) {
                    const uint32_t b = bl_b[ln];
                    const BoxOrigin bo = org[b];
                    const float4 tv = A.topv[b];
                    fo[2 * ln] = bo.ox;
                    fo[2 * ln + 1] = bo.oy;
                    fv[4 * ln] = tv.x;
                    fv[4 * ln + 1] = tv.y;
                    fv[4 * ln + 2] = tv.z;
                    fv[4 * ln + 3] = (float)(bo.oz - ozw);
                }
                wave_lds_sync();
                for (uint32_t i = 0; i < nb; ++i) {
                    const T v3t[3] = {(T)fv[4 * i], (T)fv[4 * i + 1], (T)fv[4 * i + 2]};
                    T mm;
                    const T zv = vertex_height_at<T, K>(v3t, (T)(x - fo[2 * i]), (T)(y - fo[2 * i + 1]), P, mm);
                    zc = F::max_(zc, (act & (mm > P.S2V)) ? zv - P.E + (T)fv[4 * i + 3] : -F::inf());
                }
                wave_lds_sync();
            }
            if (fw.done) break;
        }
    }
    if (!top_floor && K != K_CONE && P.cull) {
        const float tx0 = uniform_f(__double2float_rd(hull_x0));
        const float ty0 = uniform_f(__double2float_rd(hull_y0));
        const float tx1 = uniform_f(__double2float_ru(hull_x1));
        const float ty1 = uniform_f(__double2float_ru(hull_y1));
        walk_boxes(m.tree, tx0, ty0, tx1, ty1, [&](uint32_t b) {
            const BoxOrigin bo = org[b];
            if (!framed) {
                framed = true;
                ozw = bo.oz;
                W.zb = (T)(z0 - ozw);
            }
            const T dz = (T)(bo.oz - ozw);
            const T xl = (T)(x - bo.ox), yl = (T)(y - bo.oy);
            // the box's vertex coordinates (9 per triangle) through LDS, all
            // loads in flight together
            const uint32_t t0 = b * kLeaf;
            const uint32_t nt = min(m.ntri - t0, (uint32_t)kLeaf);
            const uint32_t nv = 9u * nt;
            for (uint32_t i = lane_id(); i < 9u * kLeaf; i += 64) {
                const uint32_t tr = i / 9u, c = i - 9u * tr;
                if (i < nv) my[i] = rec[(size_t)rec_len<T>() * (t0 + tr) + R_P0X + c];
            }
            wave_lds_sync();
            for (uint32_t k = 0; k < nt; ++k) zc = F::max_(zc, vertex_floor_at<T, K>(my + 9u * k, xl, yl, P, act) + dz);
            wave_lds_sync();
        });
    }
    // lane j holds the frame and cull cap of box base+j of the level-0 block
    // being walked, loaded together when the walk enters the block
    const uint32_t nbox0 = (m.ntri + kLeaf - 1) / kLeaf;
    double pox = 0.0, poy = 0.0, poz = 0.0;
    float phx = 0.0f, phy = 0.0f;
    T pcap = (T)0;
    auto block = [&](uint32_t base) {
        const uint32_t bj = base + lane_id();
        if (lane_id() < (uint32_t)kFan && bj < nbox0) {
            const BoxOrigin o = org[bj];
            pox = o.ox;
            poy = o.oy;
            poz = o.oz;
            phx = o.hx;
            phy = o.hy;
            pcap = boxcap[bj];
        }
    };
    // Evaluates one level-0 box for every lane.  Any box order gives the
    // same decisions (each lane keeps the strict-'>' maximum, and its
    // near-tie set holds every triangle within the band of its final best),
    // except for cones, whose facet-first rule makes the fold order
    // dependent: they keep the rank order.
    auto eval_box = [&](uint32_t b, const BoxOrigin& bo, T cap) {
        if (!framed) {
            framed = true;
            ozw = bo.oz;
            W.zb = (T)(z0 - ozw);
        }
        const T dz = (T)(bo.oz - ozw);
        const T xl = (T)(x - bo.ox), yl = (T)(y - bo.oy);
        if (P.cull) {
            // no lane can gain anything from this box (zc is -inf for cones)
            const T d2 = box_d2<T>(xl, yl, -(T)bo.hx, (T)bo.hx, -(T)bo.hy, (T)bo.hy);
            const bool rel = act & (cap + dz - prof_lower<T, K>(d2, P) >= F::max_(W.zb, zc) - P.CUT);
            if (!__any(rel)) return;
        }
        // stage the box's records in LDS with coalesced 16-byte loads: all
        // of them are in flight together instead of one dependent load chain
        // per triangle
        const uint32_t t0 = b * kLeaf;
        const uint32_t nt = min(m.ntri - t0, (uint32_t)kLeaf);
        {
            constexpr uint32_t kVec = kLeaf * rec_len<T>() * sizeof(T) / 16;
            constexpr uint32_t kIt = (kVec + 63) / 64;
            const uint32_t nv = nt * (uint32_t)(rec_len<T>() * sizeof(T) / 16);
            const float4* src = reinterpret_cast<const float4*>(rec + (size_t)rec_len<T>() * t0);
            float4* dst = reinterpret_cast<float4*>(my);
            // clamped (always in-bounds) loads first, then the stores, so the
            // loads of one box are all in flight together
#pragma unroll
            for (uint32_t it = 0; it < kIt; ++it) {
                const uint32_t i = lane_id() + 64 * it;
                const float4 v = src[min(i, nv - 1)];
                if (i < nv) dst[i] = v;
            }
            wave_lds_sync();
        }
        // enter the box frame
        const uint32_t bt_in = S.bt;
        S.zb = W.zb - dz;
        S.zr = W.zr - dz;
        for (int i = 0; i < 4; ++i) S.zt[i] = -F::inf();
        S.za = -F::inf();
        // Pre-test of the box's triangles, one lane per (row of 16 points,
        // triangle): the per-lane tests of the triangle loop below, made for
        // the XY hull of the row's points and the row's lowest threshold.
        // Each bound is at least the one of any point of the row (rounding
        // is monotone: the hull's coordinates in the box frame bound the
        // points' own), so a triangle no row keeps reaches no lane.
        uint32_t tmask;
        {
            const uint32_t jl = lane_id() & (uint32_t)(kLeaf - 1);
            static_assert(kLeaf == 16, "four rows of one triangle per lane");
            bool keep = false;
            T thr_row = F::inf();
            if (P.cull) thr_row = row_min_t<T>(act ? F::max_(S.zb, zc - dz) : F::inf()) - P.CUT;
            if (jl < nt) {
                const T* r = my + rec_len<T>() * jl;
                const T bx0 = r[R_BX0], bx1 = r[R_BX1], by0 = r[R_BY0], by1 = r[R_BY1];
                const T hx0 = (T)(row_x0 - bo.ox), hx1 = (T)(row_x1 - bo.ox);
                const T hy0 = (T)(row_y0 - bo.oy), hy1 = (T)(row_y1 - bo.oy);
                keep = !((bx1 < hx0 - P.R) | (bx0 > hx1 + P.R) | (by1 < hy0 - P.R) | (by0 > hy1 + P.R));
                if (P.cull) {
                    const T ddx = F::max_(F::max_(bx0 - hx1, hx0 - bx1), (T)0);
                    const T ddy = F::max_(F::max_(by0 - hy1, hy0 - by1), (T)0);
                    T cap = r[R_ZCAP] - prof_lower<T, K>(F::fma_(ddx, ddx, ddy * ddy), P);
                    if (K != K_CONE) {
                        // the facet plane's highest value over the hull (at
                        // the corner its gradient points to)
                        const T gx = r[R_GX], gy = r[R_GY];
                        const T pl = F::fma_(gy, gy >= (T)0 ? hy1 : hy0, F::fma_(gx, gx >= (T)0 ? hx1 : hx0, r[R_F0]));
                        cap = F::min_(cap, pl + r[R_PMAR]);
                    }
                    keep = keep & (cap >= thr_row);
                }
            }
            const uint64_t rowkeep = __ballot(keep);
            tmask = (uint32_t)((rowkeep | (rowkeep >> 16) | (rowkeep >> 32) | (rowkeep >> 48)) & 0xffffu);
        }
        // near-tie set: triangles reaching the band below the best height
        // the lane is known to reach (running best or floor)
        auto nt_add = [&](bool keep, T th, uint32_t ku) {
            if (__any(keep) && keep) {
                if (ntn == (uint32_t)kNearTie) {
                    // full: drop the entries that fell below the band since
                    const float lim = (float)(F::max_(S.zb + dz, zc) - P.CUT);
                    uint32_t w = 0;
                    for (uint32_t i = 0; i < (uint32_t)kNearTie; ++i) {
                        const float h = nt_h[i][ln];
                        const uint32_t kk = nt_k[i][ln];
                        if (h >= lim) {
                            nt_h[w][ln] = h;
                            nt_k[w][ln] = kk;
                            ++w;
                        }
                    }
                    ntn = w;
                }
                if (ntn < (uint32_t)kNearTie) {
                    // upper value in the wave frame, rounded up
                    const float hw = (float)(th + dz);
                    nt_h[ntn][ln] = hw + 1e-6f * (1.0f + fabsf(hw));
                    nt_k[ntn][ln] = ku;
                    ++ntn;
                } else {
                    ovf = true;
                }
            }
        };
        if constexpr (!pack_pairs<T, K>()) {
            // every lane folds its candidates among the box's triangles in
            // rank order, one triangle per wavefront step
            for (; tmask; tmask &= tmask - 1) {
                const uint32_t j = (uint32_t)__builtin_ctz(tmask);
                const uint32_t ku = uniform_u(t0 + j);
                const T* r = my + rec_len<T>() * j;
                const T bx0 = r[R_BX0], bx1 = r[R_BX1], by0 = r[R_BY0], by1 = r[R_BY1];
                // (the contact code skips a lane whose running best is at or
                // above the triangle's top: the reference's below() test)
                bool on = act & (S.zb < r[R_BZ1]) &
                          !((bx1 < xl - P.R) | (bx0 > xl + P.R) | (by1 < yl - P.R) | (by0 > yl + P.R));
                if (P.cull) {
                    const T d2 = box_d2<T>(xl, yl, bx0, bx1, by0, by1);
                    on = on & (tri_cap<T, K>(r, xl, yl, d2, P) >= F::max_(S.zb, zc - dz) - P.CUT);
                }
                if (!__any(on)) continue;
                ++evals;
                T th;
                sel_triangle<T, K>(r, ku, xl, yl, P, S, on, &th);
                if (K != K_CONE && refine) nt_add(on & (th >= F::max_(S.zb, zc - dz) - P.CUT), th, ku);
            }
        } else {
            // The lanes' (point, triangle) pairs of the box, packed 64 to a
            // batch: a lane's candidates are the box's triangles within R of
            // it whose cap reaches its threshold.  Lane p owns the pair
            // slots offp .. offp + cntp - 1, its candidates in rank order.
            uint32_t lmask = 0;
            for (; tmask; tmask &= tmask - 1) {
                const uint32_t j = (uint32_t)__builtin_ctz(tmask);
                const T* r = my + rec_len<T>() * j;
                const T bx0 = r[R_BX0], bx1 = r[R_BX1], by0 = r[R_BY0], by1 = r[R_BY1];
                bool on = act & !((bx1 < xl - P.R) | (bx0 > xl + P.R) | (by1 < yl - P.R) | (by0 > yl + P.R));
                if (P.cull) {
                    const T d2 = box_d2<T>(xl, yl, bx0, bx1, by0, by1);
                    on = on & (tri_cap<T, K>(r, xl, yl, d2, P) >= F::max_(S.zb, zc - dz) - P.CUT);
                }
                lmask |= on ? (1u << j) : 0u;
            }
            const uint32_t cntp = (uint32_t)__popc(lmask);
            uint32_t offp = 0, total = 0;
#pragma unroll
            for (int bit = 0; bit < 5; ++bit) {  // cntp <= 16
                const unsigned long long bal = __ballot((cntp >> bit) & 1u);
                offp += (uint32_t)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u))
                        << bit;
                total += (uint32_t)__popcll(bal) << bit;
            }
            total = uniform_u(total);
            for (uint32_t base = 0; base < total; base += 64) {
                // the owners write this batch's slots: (owner lane, triangle)
                {
                    uint32_t mm = lmask, g = offp;
                    while (mm && g < base) {
                        mm &= mm - 1;
                        ++g;
                    }
                    while (mm && g < base + 64) {
                        ptab[g - base] = (ln << 4) | (uint32_t)__builtin_ctz(mm);
                        mm &= mm - 1;
                        ++g;
                    }
                }
                wave_lds_sync();
                // lane i evaluates slot base + i for its owner, from the
                // owner's running best at the start of the batch
                const bool has = base + ln < total;
                const uint32_t ent = has ? ptab[ln] : 0u;
                wave_lds_sync();
                const uint32_t po = ent >> 4, j = ent & 15u;
                const T px = __shfl(xl, (int)po, 64), py = __shfl(yl, (int)po, 64), pz = __shfl(S.zb, (int)po, 64);
                Sel<T> Q;
                sel_init(Q, pz);
                T th;
                sel_triangle<T, K>(my + rec_len<T>() * j, t0 + j, px, py, P, Q, has, &th);
                ++evals;
                const uint32_t meta = (uint32_t)(Q.bf & 0xff) | (Q.bamb ? 0x100u : 0u) | (Q.pairs ? 0x200u : 0u) |
                                      (Q.bt != 0xffffffffu ? 0x400u : 0u) | (j << 16);
                // each owner folds its slots of the batch in rank order, as
                // the lane's own scan of the box would have taken them
                const uint32_t lo = max(offp, base), hi = min(offp + cntp, base + 64u);
                const uint32_t nf = hi > lo ? hi - lo : 0u;
                const uint32_t nmax = uniform_u(wave_max_u(nf));
                for (uint32_t it = 0; it < nmax; ++it) {
                    const bool v = it < nf;
                    const int sl = v ? (int)(lo - base + it) : (int)ln;
                    T qz[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) qz[i] = __shfl(Q.zt[i], sl, 64);
                    const T qa = __shfl(Q.za, sl, 64), qb = __shfl(Q.zb, sl, 64), qe = __shfl(Q.eb, sl, 64);
                    const T qth = __shfl(th, sl, 64);
                    const uint32_t qm = (uint32_t)__shfl((int)meta, sl, 64);
                    const uint32_t qk = t0 + (qm >> 16);
                    if (v) {
#pragma unroll
                        for (int i = 0; i < 4; ++i) S.zt[i] = F::max_(S.zt[i], qz[i]);
                        S.za = F::max_(S.za, qa);
                        S.pairs += (qm & 0x200u) ? 1u : 0u;
                        if ((qm & 0x400u) && qb > S.zb) {
                            S.zb = qb;
                            S.eb = qe;
                            S.bt = qk;
                            S.bf = (int)(qm & 0xffu);
                            S.bamb = (qm & 0x100u) != 0;
                        }
                    }
                    if (refine) nt_add(v & (qth >= F::max_(S.zb, zc - dz) - P.CUT), qth, qk);
                }
            }
        }
        // leave the box frame
        if (S.bt != bt_in) W.zb = S.zb + dz;
        W.zr = S.zr + dz;
        for (int i = 0; i < 4; ++i) W.zt[i] = F::max_(W.zt[i], S.zt[i] + dz);
        W.za = F::max_(W.za, S.za + dz);
    };
    // Best first: the walk collects the level-0 boxes that may matter to
    // some lane, keyed by the most any lane could reach from them, and each
    // full list (and the last) is evaluated highest first, so every lane's
    // running best rises early and the profile cull skips most of the rest.
    // Cones evaluate their lists unsorted, in rank order.
    uint32_t nlist = 0;
    const double hcx = 0.5 * (hull_x0 + hull_x1), hcy = 0.5 * (hull_y0 + hull_y1);
    const double hhx = 0.5 * (hull_x1 - hull_x0), hhy = 0.5 * (hull_y1 - hull_y0);
    // (FP64 passes keep the walk order: the list's LDS would cost them a
    // wavefront per SIMD)
    if constexpr (sizeof(T) != 4) {
        walk_boxes(m.tree, qx0, qy0, qx1, qy1, [&](uint32_t b) { eval_box(b, org[b], boxcap[b]); });
    } else {
    BoxWalk walk;
    const float* __restrict__ capup = A.capup;
    const bool wprune = K != K_CONE && P.cull && capup != nullptr;
    for (;;) {
        // The walk leaves out boxes, and whole subtrees, whose absolute cap
        // less the profile height at their distance from the hull stays
        // below every lane's threshold: the same test the sorted list's
        // early exit makes, with the thresholds known so far (they only
        // rise), so none of them would have been evaluated.
        double wthr = -INFINITY;
        if (wprune && framed)
            wthr = (double)wave_min_t<T>(act ? F::max_(W.zb, zc) : F::inf()) + ozw - (double)P.CUT;
        auto hit = [&](uint32_t i) -> bool {
            const float4 bx = m.tree.box[i];
            if (!box_hits(bx, qx0, qy0, qx1, qy1)) return false;
            if (!(wthr > -INFINITY)) return true;
            const double gx = fmax(fmax((double)bx.x - hull_x1, hull_x0 - (double)bx.z), 0.0);
            const double gy = fmax(fmax((double)bx.y - hull_y1, hull_y0 - (double)bx.w), 0.0);
            const double key =
                (double)capup[i] - (double)prof_lower<T, K>((T)__double2float_rd(gx * gx + gy * gy), P);
            return key + 1e-5 * (1.0 + fabs(key)) >= wthr;
        };
        walk_boxes_resumable_if(m.tree, hit, walk, [](uint32_t) {}, [&](uint32_t b, uint32_t) -> bool {
            if (lane_id() == 0) bl_b[nlist] = b;
            ++nlist;
            return nlist < kBoxList;
        });
        if (nlist) {
            if (K != K_CONE) {
                // keys, one box per lane: the most the box's triangles can
                // give any lane, its cap less the profile height at the
                // distance between the box and the wave's hull
                uint32_t n2 = 1;
                while (n2 < nlist) n2 <<= 1;
                wave_lds_sync();
                for (uint32_t i = lane_id(); i < n2; i += 64) {
                    float key = -INFINITY;
                    if (i < nlist) {
                        const uint32_t b = bl_b[i];
                        const BoxOrigin o = org[b];
                        const double gx = fmax(fabs(o.ox - hcx) - (double)o.hx - hhx, 0.0);
                        const double gy = fmax(fabs(o.oy - hcy) - (double)o.hy - hhy, 0.0);
                        key = (float)((double)(boxcap[b] - prof_lower<T, K>((T)(gx * gx + gy * gy), P)) + o.oz);
                    } else {
                        bl_b[i] = 0xffffffffu;
                    }
                    bl_key[i] = key;
                }
                wave_lds_sync();
                // sort by key, highest first (bitonic, in LDS)
                for (uint32_t k = 2; k <= n2; k <<= 1)
                    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                        for (uint32_t i = lane_id(); i < n2; i += 64) {
                            const uint32_t l = i ^ j;
                            if (l > i) {
                                const float ki = bl_key[i], kl = bl_key[l];
                                // descending inside blocks whose bit k is clear
                                const bool desc = (i & k) == 0;
                                if (desc ? (ki < kl) : (ki > kl)) {
                                    const uint32_t bi = bl_b[i];
                                    bl_key[i] = kl;
                                    bl_key[l] = ki;
                                    bl_b[i] = bl_b[l];
                                    bl_b[l] = bi;
                                }
                            }
                        }
                        wave_lds_sync();
                    }
            } else {
                wave_lds_sync();
            }
            for (uint32_t i = 0; i < nlist; ++i) {
                if (K != K_CONE && P.cull && framed) {
                    // keys only fall from here on: once the box's key (the
                    // most any lane can reach from it) is below every lane's
                    // threshold, no later box can matter either
                    const double thr = (double)wave_min_t<T>(act ? F::max_(W.zb, zc) : F::inf()) + ozw -
                                       (double)P.CUT;
                    const float kb = bl_key[i];
                    if ((double)kb + 1e-5 * (1.0 + fabs((double)kb)) < thr) break;
                }
                const uint32_t b = uniform_u(bl_b[i]);
                eval_box(b, org[b], boxcap[b]);
            }
            nlist = 0;
        }
        if (walk.done) break;
    }
    }
    // decision in the wave frame
    Sel<T> D = S;
    D.zb = W.zb;
    for (int i = 0; i < 4; ++i) D.zt[i] = W.zt[i];
    D.za = W.za;
    D.zr = W.zr;
    const T z0w = (T)(z0 - ozw);
    const bool up = refine && framed && sel_escalate(D, z0w, P);
    if (refine) {
        if (act) {
            // selection record for the resolve pass: the winner to refine,
            // or the near-tie set of an undecided point
            const bool usable = K != K_CONE && !ovf && framed && ntn > 0;
            const uint8_t stt = !up ? (uint8_t)kSelDecided : usable ? (uint8_t)ntn : (uint8_t)kSelNone;
            nts.cnt[src] = stt;
            nts.bt[src] = D.bt;
            nts.bf[src] = (uint8_t)D.bf;
            nts.zsel[src] = framed ? ozw + (double)D.zb : z0;
            nts.esel[src] = (float)D.eb;
            nts.zgap[src] = (sizeof(T) == 4 && framed) ? __double2float_ru((double)D.zr - (double)D.zb) : -INFINITY;
            if (up && usable) {
                for (uint32_t i = 0; i < ntn; ++i) nts.set[(size_t)src * kNearTie + i] = nt_k[i][ln];
                const double zs0 = ozw + (double)(D.zb - P.CUT);
                nts.zband[src] = fmax(z0, zs0 - 1e-9 * (1.0 + fabs(zs0)));
            }
            // Floor of the final height for the later passes: the vertex
            // floor, and the winner's value when its validity was unambiguous
            // (a decided point only reaches them when its refine fails).
            T fl = zc;
            if (K != K_CONE && up && !D.bamb && D.bt != 0xffffffffu) fl = F::max_(fl, D.zb - (T)2 * D.eb - P.E);
            zfloor[src] = (K == K_CONE || !framed) ? -INFINITY : ozw + (double)fl - 1e-9 * (1.0 + fabs(ozw));
        }
        wave_push(act && up && K != K_CONE && !ovf && framed && ntn > 0, src, nts.q, nts.qn);
    } else if (act) {
        // FP32-only mode: the selection result as is (no CC point)
        DropLane L = fresh_lane(x, y, z0);
        if (D.bt != 0xffffffffu) {
            L.z = ozw + (double)D.zb;
            L.cc = v3(NAN, NAN, NAN);
            const int sl = feat_slot(D.bf);
            L.type = (sl == 0) ? CC_VERTEX : (sl == 1) ? CC_EDGE : (K == K_CONE) ? ((sl == 2) ? CC_FACET_TIP : CC_FACET_CYL) : CC_FACET;
            L.tri = D.bt;
            L.feat = D.bf;
        }
        write_drop(o, m.order, src, L);
    }
    if (cnt) {  // the FP32/FP64 selection runs no Brent solves
        wave_add(&cnt->pairs, act ? S.pairs : 0);
        wave_add(&cnt->escalated, (act && up) ? 1 : 0);
        wave_add(&cnt->evals, lane_id() == 0 ? evals : 0);
    }
}



// One selection pass: a grid over the point list, one wavefront per block.
// Wavefront run times differ by up to 3x; a block's LDS and slot are only
// freed when all its wavefronts are done, so with 4-wavefront blocks the
// early finishers idled (13% of the wave slots on the 1M-point Ball
// workload); single-wavefront blocks free each slot for the next wavefront
// at once.  XCD-aware mapping in runs of xcd_group x kSelRun wavefronts
// (the points of xcd_group 256-point tiles).  4 wavefronts per SIMD for the
// FP32 pass, 3 for the FP64 one (register budget).
constexpr uint32_t kSelRun = 4;
template <class T, int K>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 3)))
k_drop_select(SelArgs<T> A) {
    __shared__ __align__(16) T srec[kLeaf * rec_len<T>()];
    __shared__ uint32_t nt_k[kNearTie][64];
    __shared__ float nt_h[kNearTie][64];
    __shared__ uint32_t bl_b[sizeof(T) == 4 ? kBoxList : 1];
    __shared__ float bl_key[sizeof(T) == 4 ? kBoxList : 1];
    __shared__ uint32_t ptab[64];
    const uint32_t gid = xcd_block(A.m.xcd_group * kSelRun) * 64u + threadIdx.x;
    select_points<T, K>(A, gid, srec, nt_k, nt_h, bl_b, bl_key, ptab);
}

// Grid of k_drop_select over n points.
static inline uint32_t select_grid(uint32_t n, uint32_t xcd_group) {
    return xcd_grid((n + 63) / 64, xcd_group * kSelRun);
}

// Arguments of a composite selection pass: the pass's own (P: the
// composite's constants, rec: two records per triangle) and the
// sub-cutters' constants.
struct CompSelArgs {
    SelArgs<float> A;
    SelParams<float> PS[2];
};

// FP32 selection for a two-part composite cutter (sel_triangle_comp), one
// wavefront of 64 neighbouring points (gid = position in the list).  The
// level-0 boxes are taken in walk (rank) order and each triangle's two
// records staged in LDS (my: kLeaf x 2 records).  No vertex floor and no
// near-tie sets: an undecided point goes on to the exact fold, a decided
// one to the refine of its winner (refine = 0: FP32 only, the result as is).
__device__ __attribute__((always_inline)) inline void select_points_comp(const CompSelArgs& C, uint32_t gid,
                                                                        float* my) {
    typedef FM<float> F;
    constexpr uint32_t kLen = 2 * rec_len<float>();
    const SelArgs<float>& A = C.A;
    const MeshView& m = A.m;
    const SelParams<float>& P = A.P;
    const uint32_t n_eff = A.count ? *A.count : A.n;
    if ((gid & ~63u) >= n_eff) return;
    const bool act = gid < n_eff;
    const uint32_t src = act ? A.list[gid] : 0;
    const double x = act ? A.cl[3 * (size_t)src] : 0.0;
    const double y = act ? A.cl[3 * (size_t)src + 1] : 0.0;
    const double z0 = act ? A.cl[3 * (size_t)src + 2] : 0.0;
    const double R = A.R;
    const float qx0 = uniform_f(__double2float_rd(wave_min(act ? x : INFINITY) - R));
    const float qy0 = uniform_f(__double2float_rd(wave_min(act ? y : INFINITY) - R));
    const float qx1 = uniform_f(__double2float_ru(wave_max(act ? x : -INFINITY) + R));
    const float qy1 = uniform_f(__double2float_ru(wave_max(act ? y : -INFINITY) + R));
    Sel<float> S;
    sel_init(S, 0.0f);
    float wzb = 0.0f, wza = -F::inf(), wzr = -F::inf(), wzt[4];
    for (int i = 0; i < 4; ++i) wzt[i] = -F::inf();
    uint32_t evals = 0;
    double ozw = 0.0;
    bool framed = false;
    walk_boxes(m.tree, qx0, qy0, qx1, qy1, [&](uint32_t b) {
        const BoxOrigin bo = A.org[b];
        if (!framed) {
            framed = true;
            ozw = bo.oz;
            wzb = (float)(z0 - ozw);
        }
        const float dz = (float)(bo.oz - ozw);
        const float xl = (float)(x - bo.ox), yl = (float)(y - bo.oy);
        // the lanes the box can matter to: its cap less the composite
        // profile's lower bound at the distance to the box
        const float bd2 = box_d2<float>(xl, yl, -bo.hx, bo.hx, -bo.hy, bo.hy);
        const bool rel = act & (A.boxcap[b] + dz - comp_prof_lower<float>(bd2, P) >= wzb - P.CUT);
        if (!__any(rel)) return;
        const float lx0 = (float)((double)qx0 - bo.ox), lx1 = (float)((double)qx1 - bo.ox);
        const float ly0 = (float)((double)qy0 - bo.oy), ly1 = (float)((double)qy1 - bo.oy);
        const uint32_t t0 = b * kLeaf;
        const uint32_t nt = min(m.ntri - t0, (uint32_t)kLeaf);
        {
            constexpr uint32_t kVec = kLeaf * kLen * sizeof(float) / 16;
            constexpr uint32_t kIt = (kVec + 63) / 64;
            const uint32_t nv = nt * (uint32_t)(kLen * sizeof(float) / 16);
            const float4* srcv = reinterpret_cast<const float4*>(A.rec + (size_t)kLen * t0);
            float4* dst = reinterpret_cast<float4*>(my);
#pragma unroll
            for (uint32_t it = 0; it < kIt; ++it) {
                const uint32_t i = lane_id() + 64 * it;
                const float4 v = srcv[min(i, nv - 1)];
                if (i < nv) dst[i] = v;
            }
            wave_lds_sync();
        }
        const uint32_t bt_in = S.bt;
        S.zb = wzb - dz;
        S.zr = wzr - dz;
        for (int i = 0; i < 4; ++i) S.zt[i] = -F::inf();
        S.za = -F::inf();
        // triangles whose XY box meets the wave's query box, one per lane
        uint32_t tmask;
        {
            const uint32_t jl = lane_id();
            bool keep = false;
            if (jl < nt) {
                const float* r = my + kLen * jl;
                keep = !((r[R_BX1] < lx0) | (r[R_BX0] > lx1) | (r[R_BY1] < ly0) | (r[R_BY0] > ly1));
            }
            tmask = (uint32_t)__ballot(keep);
        }
        for (; tmask; tmask &= tmask - 1) {
            const uint32_t j = (uint32_t)__builtin_ctz(tmask);
            const uint32_t ku = uniform_u(t0 + j);
            const float* r = my + kLen * j;
            const float bx0 = r[R_BX0], bx1 = r[R_BX1], by0 = r[R_BY0], by1 = r[R_BY1];
            bool on = rel & !((bx1 < xl - P.R) | (bx0 > xl + P.R) | (by1 < yl - P.R) | (by0 > yl + P.R));
            const float d2 = box_d2<float>(xl, yl, bx0, bx1, by0, by1);
            on = on & (comp_tri_cap<float>(r, xl, yl, d2, P) >= S.zb - P.CUT);
            if (!__any(on)) continue;
            ++evals;
            sel_triangle_comp<float>(r, ku, xl, yl, P, C.PS, S, on);
        }
        if (S.bt != bt_in) wzb = S.zb + dz;
        wzr = S.zr + dz;
        for (int i = 0; i < 4; ++i) wzt[i] = F::max_(wzt[i], S.zt[i] + dz);
        wza = F::max_(wza, S.za + dz);
    });
    Sel<float> D = S;
    D.zb = wzb;
    for (int i = 0; i < 4; ++i) D.zt[i] = wzt[i];
    D.za = wza;
    D.zr = wzr;
    const bool up = A.refine && framed && sel_escalate(D, (float)(z0 - ozw), P);
    const NearTieSets& nts = A.nts;
    if (A.refine) {
        if (act) {
            nts.cnt[src] = up ? kSelNone : kSelDecided;
            nts.bt[src] = D.bt;
            nts.bf[src] = (uint8_t)D.bf;
            nts.zsel[src] = framed ? ozw + (double)D.zb : z0;
            nts.esel[src] = D.eb;
            nts.zgap[src] = framed ? __double2float_ru((double)D.zr - (double)D.zb) : -INFINITY;
            A.zfloor[src] = -INFINITY;
        }
    } else if (act) {
        // FP32-only mode: the selection result as is (no CC point)
        DropLane L = fresh_lane(x, y, z0);
        if (D.bt != 0xffffffffu) {
            L.z = ozw + (double)D.zb;
            L.cc = v3(NAN, NAN, NAN);
            const int sl = feat_slot(D.bf);
            L.type = (sl == 0) ? CC_VERTEX : (sl == 1) ? CC_EDGE : CC_FACET;
            L.tri = D.bt;
            L.feat = D.bf;
        }
        write_drop(A.o, m.order, src, L);
    }
    if (A.cnt) {
        wave_add(&A.cnt->pairs, act ? S.pairs : 0);
        wave_add(&A.cnt->escalated, (act && up) ? 1 : 0);
        wave_add(&A.cnt->evals, lane_id() == 0 ? evals : 0);
    }
}

__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) k_drop_select_comp(CompSelArgs C) {
    __shared__ __align__(16) float srec[kLeaf * 2 * rec_len<float>()];
    const uint32_t gid = xcd_block(C.A.m.xcd_group * kSelRun) * 64u + threadIdx.x;
    select_points_comp(C, gid, srec);
}

// The selection passes of a batch of jobs in one persistent launch: the
// wavefronts pull 64-point chunks from a device work queue that runs over
// every job's chunks (chunk_off: prefix sums of the jobs' chunk counts,
// largest jobs first so the tail is short) and run the job's pass on them.
// Cutter kinds are dispatched per chunk; composite jobs are not queued.
// Three wavefronts per SIMD: with the four kinds' code in one body, a budget
// of four (128 VGPRs) spilled 17 VGPRs to scratch in the FP32 pass and took
// the C3 batch's selection from 20.5 to 21.2 ms.
template <class T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
k_drop_select_jobs(const SelArgs<T>* __restrict__ jobs, const uint32_t* __restrict__ chunk_off, uint32_t njobs,
                   uint32_t* __restrict__ queue) {
    __shared__ __align__(16) T srec[4][kLeaf * rec_len<T>()];
    __shared__ uint32_t nt_k[4][kNearTie][64];
    __shared__ float nt_h[4][kNearTie][64];
    __shared__ uint32_t bl_b[4][sizeof(T) == 4 ? kBoxList : 1];
    __shared__ float bl_key[4][sizeof(T) == 4 ? kBoxList : 1];
    __shared__ uint32_t ptab[4][64];
    const uint32_t wv = threadIdx.x >> 6;
    const uint32_t nchunks = chunk_off[njobs];
    for (;;) {
        uint32_t c = 0;
        if (lane_id() == 0) c = atomicAdd(queue, 1u);
        c = uniform_u(__shfl(c, 0, 64));
        if (c >= nchunks) break;
        uint32_t lo = 0, hi = njobs;  // the job: last j with chunk_off[j] <= c
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (chunk_off[mid] <= c) lo = mid;
            else hi = mid;
        }
        const SelArgs<T>& A = jobs[lo];
        const uint32_t gid = (c - chunk_off[lo]) * 64u + lane_id();
        switch (A.P.kind) {
        case K_CYL: select_points<T, K_CYL>(A, gid, srec[wv], nt_k[wv], nt_h[wv], bl_b[wv], bl_key[wv], ptab[wv]); break;
        case K_BALL: select_points<T, K_BALL>(A, gid, srec[wv], nt_k[wv], nt_h[wv], bl_b[wv], bl_key[wv], ptab[wv]); break;
        case K_BULL: select_points<T, K_BULL>(A, gid, srec[wv], nt_k[wv], nt_h[wv], bl_b[wv], bl_key[wv], ptab[wv]); break;
        default: select_points<T, K_CONE>(A, gid, srec[wv], nt_k[wv], nt_h[wv], bl_b[wv], bl_key[wv], ptab[wv]); break;
        }
    }
}

// Resolve pass, one lane per CL point (input order), after a selection pass:
//  - decided points: the winner's feature is re-evaluated by the exact FP64
//    code and kept if it is consistent with the selection (same contact
//    class, lifted, height within the selection's error bound);
//  - undecided points with a near-tie set: the reference's fold
//    (MillingCutter::dropCutter over triangles in rank order) restricted to
//    the set and started at the bottom of its band.  Every other triangle's
//    contacts stay below that start, and for Cyl/Ball/Bull the reference's
//    result is the maximum of the triangles' own drops (ties to the lowest
//    rank), so when a set triangle lifts above the start the fold's result
//    is the reference's.
// Everything else goes on to the cooperative passes.
__device__ __attribute__((always_inline)) inline void resolve_points(
    const MeshView& m, const CutterDev& cd, const double* __restrict__ cl, const uint32_t* __restrict__ perm,
    uint32_t n, const NearTieSets& nts, double etie, const DropOut& o, uint32_t* __restrict__ next,
    uint32_t* __restrict__ next_n, Counters* cnt, uint32_t g) {
    // points in the Hilbert order of the selection pass: neighbouring lanes
    // refine the same few triangles
    if ((g & ~63u) >= n) return;
    const bool act = g < n;
    const uint32_t i = act ? perm[g] : 0u;
    const uint8_t stt = act ? nts.cnt[i] : kSelNone;
    bool done = false;
    unsigned long long brent = 0;
    if (act && stt != kSelNone) {
        const double x = cl[3 * (size_t)i], y = cl[3 * (size_t)i + 1], z0 = cl[3 * (size_t)i + 2];
        if (stt == kSelDecided) {
            const uint32_t bt = nts.bt[i];
            if (bt == 0xffffffffu) {
                write_drop(o, m.order, i, fresh_lane(x, y, z0));
                done = true;
            } else {
                const int bf = nts.bf[i];
                const TriX t = load_tri(m.tri, bt);
                DropLane X = fresh_lane(x, y, -INFINITY);
                X.cur_tri = bt;
                int it = 0;
                if (cd.c.kind == K_COMPOSITE) exact_feature_comp(cd, X, t, bf, it);
                else exact_feature(cd.c, X, t, bf, it);
                brent += it;
                // the winner's FP64 height must stay consistent with its
                // selection value and above every other contact's upper value
                if (X.type != CC_NONE && feat_slot(X.feat) == feat_slot(bf) && X.z > z0 &&
                    fabs(X.z - nts.zsel[i]) <= 4.0 * ((double)nts.esel[i] + etie) &&
                    X.z > nts.zsel[i] + (double)nts.zgap[i]) {
                    X.tri = bt;
                    write_drop(o, m.order, i, X);
                    done = true;
                }
            }
        }
    }
    // near-tie points were queued by the selection pass for k_drop_neartie,
    // which folds them with full wavefronts
    const bool tie = act && stt != kSelNone && stt != kSelDecided;
    wave_push(act && !done && !tie, i, next, next_n);
    if (cnt) wave_add(&cnt->brent, brent);
}

__global__ void __launch_bounds__(256) k_drop_resolve(MeshView m, CutterDev cd, const double* __restrict__ cl,
                                                      const uint32_t* __restrict__ perm, uint32_t n, NearTieSets nts,
                                                      double etie, DropOut o, uint32_t* __restrict__ next,
                                                      uint32_t* __restrict__ next_n, Counters* cnt) {
    resolve_points(m, cd, cl, perm, n, nts, etie, o, next, next_n, cnt, blockIdx.x * blockDim.x + threadIdx.x);
}

// Near-tie pass over the queued points, 8 lanes per point (lane e takes
// entry e of the point's near-tie set): the reference's fold
// (MillingCutter::dropCutter over triangles in rank order) restricted to the
// set and started at the bottom of its band.  Every other triangle's
// contacts stay below that start.  For Cyl/Ball/Bull the fold's result is
// the highest of the set triangles' own drops, ties to the lowest rank (a
// triangle's own drop is its facet contact when that is valid, else its
// best vertex/edge contact), so each lane computes one triangle's drop and
// the 8 lanes take the arg-max.  Two rounding cases make the fold depend on
// the running height (a vertex/edge contact above a valid facet contact of
// the same triangle, a contact above its triangle's max z): such points, and
// points where no set triangle lifts above the band, go to `next`.
// g0: the thread's index in the grid, stride: the grid's thread count (a
// multiple of 64: the loop is wave-uniform).
__device__ __attribute__((always_inline)) inline void neartie_points(
    const MeshView& m, const CutterDev& cd, const double* __restrict__ cl, const uint32_t* __restrict__ ntq,
    uint32_t n_eff, const NearTieSets& nts, const DropOut& o, uint32_t* __restrict__ next,
    uint32_t* __restrict__ next_n, unsigned long long& brent, uint32_t g0, uint32_t stride) {
    static_assert(kNearTie == 8, "one lane per near-tie set entry");
    for (uint32_t g = g0; ((g & ~63u) >> 3) < n_eff; g += stride) {
        const uint32_t slot = g >> 3, e = g & 7u;
        const bool act = slot < n_eff;
        const uint32_t i = act ? ntq[slot] : 0u;
        const uint32_t stt = act ? nts.cnt[i] : 0u;
        const double x = act ? cl[3 * (size_t)i] : 0.0, y = act ? cl[3 * (size_t)i + 1] : 0.0;
        const double zb = act ? nts.zband[i] : 0.0;
        const bool has = act && e < stt;
        const uint32_t k = has ? nts.set[(size_t)i * kNearTie + e] : 0xffffffffu;
        DropLane A = fresh_lane(x, y, -INFINITY), V = fresh_lane(x, y, -INFINITY);
        double maxz = -INFINITY;
        if (has) {
            const TriX t = load_tri(m.tri, k);
            if (cutter_overlaps(cd.c.radius, x, y, t)) {
                maxz = t.maxz;
                A.cur_tri = V.cur_tri = k;
                int it = 0;
                exact_feature(cd.c, A, t, FT_FACET, it);
                V.cur_feat = 0;
                vertex_drop(cd, V, t);
                if (V.z < t.maxz) {
                    edge_simple(cd.c, V, t, it);
                    brent += it;
                }
            }
        }
        const bool fv = A.type != CC_NONE, vv = V.type != CC_NONE;
        const double val = fv ? A.z : (vv ? V.z : -INFINITY);
        bool odd = (fv && vv && V.z > A.z) || ((fv || vv) && val > maxz);
        double best = val;
        uint32_t bk = k;
        for (int off = 4; off > 0; off >>= 1) {
            const double ob = __shfl_xor(best, off, 64);
            const uint32_t ok = (uint32_t)__shfl_xor((int)bk, off, 64);
            const bool take = (ob > best) || (ob == best && ok < bk);
            best = take ? ob : best;
            bk = take ? ok : bk;
            odd = odd || (__shfl_xor((int)odd, off, 64) != 0);
        }
        const bool lift = !odd && best > zb;
        if (lift && has && k == bk) {
            DropLane L = fv ? A : V;
            L.tri = k;
            write_drop(o, m.order, i, L);
        }
        wave_push(act && e == 0 && !lift, i, next, next_n);
    }
}

__global__ void __launch_bounds__(256) k_drop_neartie(MeshView m, CutterDev cd, const double* __restrict__ cl,
                                                      const uint32_t* __restrict__ ntq,
                                                      const uint32_t* __restrict__ ntq_n, NearTieSets nts, DropOut o,
                                                      uint32_t* __restrict__ next, uint32_t* __restrict__ next_n,
                                                      Counters* cnt) {
    unsigned long long brent = 0;
    neartie_points(m, cd, cl, ntq, *ntq_n, nts, o, next, next_n, brent, blockIdx.x * blockDim.x + threadIdx.x,
                   gridDim.x * blockDim.x);
    if (cnt) wave_add(&cnt->brent, brent);
}

// ------------------------------------------------- cooperative tiers
// One wavefront per point (points of an escalation queue are scattered, so
// a lane-per-point wave would walk a huge union box).  The wave walks the
// hierarchy with the point's own cutter box, gathers the overlapping
// triangles in rank order into batches of 64 (one per lane), evaluates the
// batch in parallel and folds it into the point's state in rank order.
constexpr int kCoopWaves = 4;  // wavefronts per 256-thread block


// Gather driver: calls batch(list, cnt) with cnt <= 64 candidate triangle
// indices in rank order.  Candidates are triangles whose XY box overlaps
// [x-R, x+R] x [y-R, y+R] (the reference overlaps() test, exact FP64).
// BoxCull: optional per-level-0-box cull (cap = box cull cap in the box frame,
// frames in org); a box is skipped when cap + oz - h(distance) < thr().
struct BoxCull {
    const double* cap;
    const BoxOrigin* org;
    const float* capup;  // absolute caps of every hierarchy box (the cap array's records), or NULL
};

// The walk's level-0 boxes are tested four at a time, one triangle per lane
// (lane = 16 x slot + triangle), so each round of dependent loads covers
// 64 triangles; candidates are appended in rank order.  qbuf holds
// kCoopQueue entries.
constexpr uint32_t kCoopQueue = 128;
template <class Thr, class Keep, class Batch>
__device__ inline void coop_gather(const MeshView& m, uint32_t* qbuf, double x, double y, const CutterP& c,
                                   const BoxCull& bc, Thr&& thr, Keep&& keep, Batch&& batch) {
    static_assert(4 * kLeaf == 64, "four level-0 boxes per wavefront");
    const uint32_t lane = lane_id();
    const double R = c.radius;
    const float qx0 = __double2float_rd(x - R), qy0 = __double2float_rd(y - R);
    const float qx1 = __double2float_ru(x + R), qy1 = __double2float_ru(y + R);
    uint32_t qn = 0;
    const uint32_t nbox0 = (m.ntri + kLeaf - 1) / kLeaf;
    double pcap = INFINITY, pox = 0.0, poy = 0.0, poz = 0.0;
    float phx = 0.0f, phy = 0.0f;
    auto block = [&](uint32_t base) {
        const uint32_t bj = base + lane;
        if (bc.cap && lane < (uint32_t)kFan && bj < nbox0) {
            const BoxOrigin o = bc.org[bj];
            pcap = bc.cap[bj];
            pox = o.ox;
            poy = o.oy;
            poz = o.oz;
            phx = o.hx;
            phy = o.hy;
        }
    };
    // With the hierarchy's absolute caps the walk enters no box, at any
    // level, whose cap less the profile height at its distance from the
    // point stays below the threshold (the contacts of every triangle under
    // it do too; the threshold only rises while the point is folded).
    auto hit = [&](uint32_t i) -> bool {
        const float4 bx = m.tree.box[i];
        if (!box_hits(bx, qx0, qy0, qx1, qy1)) return false;
        if (!bc.capup) return true;
        const double gx = fmax(fmax((double)bx.x - x, x - (double)bx.z), 0.0);
        const double gy = fmax(fmax((double)bx.y - y, y - (double)bx.w), 0.0);
        const double d2 = fmax(fma(gx, gx, gy * gy) * (1.0 - 1e-12) - 1e-12, 0.0);
        return (double)bc.capup[i] - prof_lower64(c, d2) >= thr();
    };
    BoxWalk w;
    for (;;) {
        // up to four level-0 boxes that survive the box-level cull
        uint32_t pb0 = 0, pb1 = 0, pb2 = 0, pb3 = 0, npend = 0;  // wave-uniform
        if (!w.done)
            walk_boxes_resumable_if(m.tree, hit, w, block, [&](uint32_t b, uint32_t jb) -> bool {
                if (bc.cap) {
                    // every contact in the box is below cap + oz - h(d)
                    const double ox = readlane_d(pox, jb), oy = readlane_d(poy, jb), oz = readlane_d(poz, jb);
                    const double hx = readlane_f(phx, jb), hy = readlane_f(phy, jb);
                    const double dx = fmax(fabs(x - ox) - hx, 0.0), dy = fmax(fabs(y - oy) - hy, 0.0);
                            if (readlane_d(pcap, jb) + oz - prof_lower64(c, dx * dx + dy * dy) < thr()) return true;
                }
                pb0 = (npend == 0) ? b : pb0;
                pb1 = (npend == 1) ? b : pb1;
                pb2 = (npend == 2) ? b : pb2;
                pb3 = (npend == 3) ? b : pb3;
                return ++npend < 4;
            });
        if (npend) {
            const uint32_t slot = lane >> 4;
            const uint32_t b = (slot == 0) ? pb0 : (slot == 1) ? pb1 : (slot == 2) ? pb2 : pb3;
            const uint32_t k = b * kLeaf + (lane & 15u);
            bool cand = false;
            if (slot < npend && k < m.ntri) {
                const TriX t = load_tri(m.tri, k);
                cand = cutter_overlaps(R, x, y, t) && keep(k, t);
            }
            const unsigned long long bal = __ballot(cand);
            const uint32_t pos = qn + __popcll(bal & ((1ull << lane) - 1ull));
            if (cand) qbuf[pos] = k;
            qn += __popcll(bal);
            __builtin_amdgcn_wave_barrier();
        }
        if (qn >= 64 || (w.done && qn > 0)) {
            const uint32_t nb = qn < 64 ? qn : 64u;
            batch(qbuf, nb);
            const uint32_t rest = qn - nb;
            const uint32_t moved = (lane < rest) ? qbuf[64 + lane] : 0u;
            __builtin_amdgcn_wave_barrier();
            if (lane < rest) qbuf[lane] = moved;
            __builtin_amdgcn_wave_barrier();
            qn = rest;
        }
        if (w.done && qn == 0) break;
    }
}


// Exact FP64 tier: reproduces the reference's sequential dropCutter fold.
// Per triangle the lane computes, from z = -inf, the facet (cone: tip)
// contact, the cone rim contact and the best vertex/edge contact; the fold
// then applies them in rank order exactly as MillingCutter::dropCutter does
// against the running z (a lifting facet skips the triangle's vertices and
// edges; the cone rim is tried only when the tip did not lift).  Comp: a
// composite cutter, whose facet contact is the highest valid one of its
// sub-cutters' and whose edge contacts are its sub-cutters' window-tested
// ones (compositecutter.cpp:152-194); the fold is the same.
// w0: the wavefront's index in the grid, nwaves: the grid's wavefronts;
// qbuf: the wavefront's kCoopQueue-entry LDS queue.
template <bool Comp>
__device__ __attribute__((always_inline)) inline void coop_exact_points(
    const MeshView& m, const CutterDev& cd, const double* __restrict__ cl, const uint32_t* __restrict__ list,
    uint32_t n_eff, const double* __restrict__ zfloor, const double* __restrict__ boxcap,
    const BoxOrigin* __restrict__ org, const DropOut& o, uint32_t* qbuf, unsigned long long& brent, uint32_t w0,
    uint32_t nwaves, const float* __restrict__ capup = nullptr) {
    const uint32_t lane = lane_id();
    for (uint32_t w = w0; w < n_eff; w += nwaves) {
        const uint32_t src = uniform_u(list[w]);
        const double x = cl[3 * (size_t)src], y = cl[3 * (size_t)src + 1], z0 = cl[3 * (size_t)src + 2];
        const CutterP& c = cd.c;
        double zr = z0;
        uint32_t win_k = 0xffffffffu;
        int win_part = -1;
        V3 win_cc = v3(0, 0, 0);
        int win_type = CC_NONE, win_feat = FT_NONE;
        // Cyl/Ball/Bull: the reference's result is the maximum over the
        // triangles' own drops, so a triangle whose contacts all stay below
        // a known floor of that maximum cannot lift or tie.  Cones are
        // order dependent: they skip only triangles whose contacts all stay
        // below the running height (which at a triangle's turn is at least
        // the running height when it is gathered), as those cannot lift.
        // Composites are order dependent too (the edges' below() test): they
        // skip triangles whose contacts all stay below the running height.
        const bool cull = !Comp && zfloor && (c.kind == K_CYL || c.kind == K_BALL || c.kind == K_BULL);
        const bool crun = Comp ? comp_has_cap(cd) : (cull || (c.kind == K_CONE && has_contact_cap(c)));
        const double zf = cull ? zfloor[src] : -INFINITY;
        auto thr = [&]() {
            const double t = fmax(zf, zr);
            return t - 1e-9 * (1.0 + fabs(t));
        };
        auto keep = [&](uint32_t, const TriX& t) {
            if (!crun) return true;
            if constexpr (Comp) return t.maxz - comp_prof_lower64(cd, tri_box_d2(t, x, y)) >= thr();
            return contact_cap(c, t, x, y) >= thr();
        };
        const BoxCull bcl{(crun && !Comp) ? boxcap : nullptr, org, (crun && !Comp && boxcap) ? capup : nullptr};
        coop_gather(m, qbuf, x, y, c, bcl, thr, keep, [&](uint32_t* q, uint32_t nb) {
            const bool has = lane < nb;
            const uint32_t k = has ? q[lane] : 0u;
            DropLane A = fresh_lane(x, y, -INFINITY), B = fresh_lane(x, y, -INFINITY), V = fresh_lane(x, y, -INFINITY);
            double maxz = -INFINITY;
            if (has) {
                const TriX t = load_tri(m.tri, k);
                maxz = t.maxz;
                A.cur_tri = B.cur_tri = V.cur_tri = k;
                int it = 0;
                if constexpr (Comp) {
                    A.cur_feat = 0;
                    facet_composite(cd, A, t);
                } else {
                    exact_feature(c, A, t, FT_FACET, it);
                    if (c.kind == K_CONE) exact_feature(c, B, t, FT_FACET_RIM, it);
                }
                V.cur_feat = 0;
                vertex_drop(cd, V, t);
                if (V.z < t.maxz) {
                    int it2 = 0;
                    if constexpr (Comp) edge_composite(cd, V, t, it2);
                    else edge_simple(c, V, t, it2);
                    brent += it2;
                }
            }
            // Cyl/Ball/Bull: a triangle's own drop is its facet contact when
            // that is valid (a valid facet contact is the triangle's highest),
            // else its best vertex/edge contact, and the rank-ordered fold
            // keeps the highest of these, ties to the lowest rank: a wave
            // arg-max.  Batches where a contact rounds above a valid facet
            // contact of its triangle or above its triangle's max z (the
            // fold's outcome then depends on the running height) and cones
            // take the rank-ordered fold below.
            const bool fv = A.type != CC_NONE, vv = V.type != CC_NONE;
            if (c.kind != K_CONE && !__any((fv && vv && V.z > A.z) || ((fv || vv) && (fv ? A.z : V.z) > maxz))) {
                double best = fv ? A.z : (vv ? V.z : -INFINITY);
                if (!(maxz > zr)) best = -INFINITY;  // below()
                uint32_t bl = lane;
                for (int off = 32; off > 0; off >>= 1) {
                    const double ob = __shfl_xor(best, off, 64);
                    const uint32_t ol = (uint32_t)__shfl_xor((int)bl, off, 64);
                    const bool take = (ob > best) || (ob == best && ol < bl);
                    best = take ? ob : best;
                    bl = take ? ol : bl;
                }
                best = __shfl(best, 0, 64);
                bl = (uint32_t)__shfl((int)bl, 0, 64);
                if (best > zr) {
                    const uint32_t wl = uniform_u(bl);
                    zr = best;
                    win_k = __shfl(k, wl, 64);
                    win_part = 0;
                    const double cx = fv ? A.cc.x : V.cc.x, cy = fv ? A.cc.y : V.cc.y, cz = fv ? A.cc.z : V.cc.z;
                    const int ty = fv ? A.type : V.type, ft = fv ? A.feat : V.feat;
                    win_cc = v3(__shfl(cx, wl, 64), __shfl(cy, wl, 64), __shfl(cz, wl, 64));
                    win_type = __shfl(ty, wl, 64);
                    win_feat = __shfl(ft, wl, 64);
                }
                return;
            }
            // rank-ordered fold (uniform across the wave)
            for (uint32_t i = 0; i < nb; ++i) {
                const double zmax = __shfl(maxz, i, 64);
                if (!(zr < zmax)) continue;  // below()
                const double z1 = __shfl(A.z, i, 64), z2 = __shfl(B.z, i, 64), z3 = __shfl(V.z, i, 64);
                const bool v1 = __shfl(A.type, i, 64) != CC_NONE, v2 = __shfl(B.type, i, 64) != CC_NONE;
                const bool v3_ = __shfl(V.type, i, 64) != CC_NONE;
                // millingcutter.cpp:386-409 / conecutter.cpp:70-105
                int part = -1;
                if (v1 && z1 > zr) part = 0;
                else if (v2 && z2 > zr) part = 1;
                else if (v3_ && z3 > zr) part = 2;
                if (part >= 0) {
                    zr = (part == 0) ? z1 : (part == 1) ? z2 : z3;
                    win_k = __shfl(k, i, 64);
                    win_part = part;
                    // select by value (a reference select would spill the lanes to scratch)
                    const double cx = (part == 0) ? A.cc.x : (part == 1) ? B.cc.x : V.cc.x;
                    const double cy = (part == 0) ? A.cc.y : (part == 1) ? B.cc.y : V.cc.y;
                    const double cz = (part == 0) ? A.cc.z : (part == 1) ? B.cc.z : V.cc.z;
                    const int ty = (part == 0) ? A.type : (part == 1) ? B.type : V.type;
                    const int ft = (part == 0) ? A.feat : (part == 1) ? B.feat : V.feat;
                    win_cc = v3(__shfl(cx, i, 64), __shfl(cy, i, 64), __shfl(cz, i, 64));
                    win_type = __shfl(ty, i, 64);
                    win_feat = __shfl(ft, i, 64);
                }
            }
        });
        if (lane == 0) {
            DropLane L = fresh_lane(x, y, z0);
            if (win_part >= 0) {
                L.z = zr;
                L.cc = win_cc;
                L.type = win_type;
                L.tri = win_k;
                L.feat = win_feat;
            }
            write_drop(o, m.order, src, L);
        }
    }
}

__global__ void __launch_bounds__(256) k_drop_coop_exact(MeshView m, CutterDev cd, const double* __restrict__ cl,
                                                         const uint32_t* __restrict__ list,
                                                         const uint32_t* __restrict__ count,
                                                         const double* __restrict__ zfloor,
                                                         const double* __restrict__ boxcap,
                                                         const BoxOrigin* __restrict__ org, DropOut o,
                                                         Counters* cnt, const float* __restrict__ capup) {
    __shared__ uint32_t qmem[kCoopWaves][kCoopQueue];
    const uint32_t wv = threadIdx.x >> 6;
    unsigned long long brent = 0;
    coop_exact_points<false>(m, cd, cl, list, *count, zfloor, boxcap, org, o, qmem[wv], brent,
                             xcd_block(m.xcd_group) * kCoopWaves + wv, gridDim.x * kCoopWaves, capup);
    if (cnt) wave_add(&cnt->brent, brent);
}

// The exact tier for a composite cutter (no floors, no box caps).
__global__ void __launch_bounds__(256) k_drop_coop_exact_comp(MeshView m, CutterDev cd,
                                                              const double* __restrict__ cl,
                                                              const uint32_t* __restrict__ list,
                                                              const uint32_t* __restrict__ count,
                                                              const BoxOrigin* __restrict__ org, DropOut o,
                                                              Counters* cnt) {
    __shared__ uint32_t qmem[kCoopWaves][kCoopQueue];
    const uint32_t wv = threadIdx.x >> 6;
    unsigned long long brent = 0;
    coop_exact_points<true>(m, cd, cl, list, *count, nullptr, nullptr, org, o, qmem[wv], brent,
                            xcd_block(m.xcd_group) * kCoopWaves + wv, gridDim.x * kCoopWaves);
    if (cnt) wave_add(&cnt->brent, brent);
}

// The passes after the selection for every fast job of a batch, one launch
// each instead of one per job: the refine pass over the jobs' 64-point
// chunks (chunk_off over the jobs in table order), then persistent near-tie
// and exact passes whose wavefronts take every job's queue in turn (the
// queue lengths are only known on the device).
struct PostJob {
    MeshView m;
    CutterDev cd;
    const double* cl;
    const uint32_t* perm;
    uint32_t n;
    NearTieSets nts;
    double etie;
    DropOut o;
    uint32_t* q1;      // exact-tier queue; its count is qcount[0]
    uint32_t* qcount;  // queue counts: exact tier, (unused), near-tie
    const double* zfloor;
    const double* cap64;
    const BoxOrigin* org;
    Counters* cnt2;
    Counters* cnt3;
    const float* capup;  // absolute hierarchy caps of the job's FP64 records, or NULL
};

__global__ void __launch_bounds__(256) k_drop_resolve_jobs(const PostJob* __restrict__ jobs,
                                                           const uint32_t* __restrict__ chunk_off, uint32_t njobs) {
    const uint32_t c = uniform_u(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (c >= chunk_off[njobs]) return;
    uint32_t lo = 0, hi = njobs;  // the job: last j with chunk_off[j] <= c
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (chunk_off[mid] <= c) lo = mid;
        else hi = mid;
    }
    const PostJob& J = jobs[lo];
    resolve_points(J.m, J.cd, J.cl, J.perm, J.n, J.nts, J.etie, J.o, J.q1, J.qcount, J.cnt2,
                   (c - chunk_off[lo]) * 64u + lane_id());
}

__global__ void __launch_bounds__(256) k_drop_neartie_jobs(const PostJob* __restrict__ jobs, uint32_t njobs) {
    for (uint32_t j = 0; j < njobs; ++j) {
        const PostJob& J = jobs[j];
        unsigned long long brent = 0;
        neartie_points(J.m, J.cd, J.cl, J.nts.q, J.qcount[2], J.nts, J.o, J.q1, J.qcount, brent,
                       blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
        if (J.cnt2) wave_add(&J.cnt2->brent, brent);
    }
}

__global__ void __launch_bounds__(256) k_drop_coop_exact_jobs(const PostJob* __restrict__ jobs, uint32_t njobs) {
    __shared__ uint32_t qmem[kCoopWaves][kCoopQueue];
    const uint32_t wv = threadIdx.x >> 6;
    const uint32_t nwaves = gridDim.x * kCoopWaves;
    // consecutive jobs start on different wavefronts, so short queues of
    // many jobs spread over the grid
    uint32_t first = 0;
    for (uint32_t j = 0; j < njobs; ++j) {
        const PostJob& J = jobs[j];
        const uint32_t nq = J.qcount[0];
        const uint32_t w0 = (blockIdx.x * kCoopWaves + wv + nwaves - first % nwaves) % nwaves;
        unsigned long long brent = 0;
        coop_exact_points<false>(J.m, J.cd, J.cl, J.q1, nq, J.zfloor, J.cap64, J.org, J.o, qmem[wv], brent, w0,
                                 nwaves, J.capup);
        if (J.cnt3) wave_add(&J.cnt3->brent, brent);
        first += nq;
    }
}

// FP64 selection tier for Cyl/Ball/Bull: per-triangle selection summaries
// (from z = -inf) reduced in rank order; undecided points go to the exact
// tier queue, decided ones are refined by the exact path.
template <int K>
__global__ void __launch_bounds__(256) k_drop_coop_select(MeshView m, const double* __restrict__ rec,
                                                          const BoxOrigin* __restrict__ org, SelParams<double> P,
                                                          CutterDev cd, const double* __restrict__ cl,
                                                          const uint32_t* __restrict__ list,
                                                          const uint32_t* __restrict__ count,
                                                          double* __restrict__ zfloor,
                                                          const double* __restrict__ boxcap, DropOut o,
                                                          uint32_t* __restrict__ next, uint32_t* __restrict__ next_n,
                                                          Counters* cnt) {
    __shared__ uint32_t qmem[kCoopWaves][kCoopQueue];
    const uint32_t lane = lane_id();
    const uint32_t wv = threadIdx.x >> 6;
    uint32_t* qbuf = qmem[wv];
    const uint32_t n_eff = *count;
    const uint32_t nwaves = gridDim.x * kCoopWaves;
    unsigned long long brent = 0, nesc = 0;
    for (uint32_t w = xcd_block(m.xcd_group) * kCoopWaves + wv; w < n_eff; w += nwaves) {
        const uint32_t src = uniform_u(list[w]);
        const double x = cl[3 * (size_t)src], y = cl[3 * (size_t)src + 1], z0 = cl[3 * (size_t)src + 2];
        // wave state in the global z frame
        double zb = z0, eb = 0.0, zt[4] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY}, za = -INFINITY;
        uint32_t bt = 0xffffffffu;
        int bf = FT_NONE;
        bool bamb = false;
        // floor of the final height handed down by the first pass: triangles
        // whose cull cap stays below it cannot lift, tie or be ambiguous
        const double zf = zfloor[src];
        auto thr = [&]() { return fmax(zf, zb) - (double)P.CUT; };
        auto keep = [&](uint32_t k, const TriX& t) {
            const double cap = rec[(size_t)kFastRec * k + R_ZCAP] - t.maxz;  // slack above max z
            return contact_cap(cd.c, t, x, y) + cap + org[k / kLeaf].oz >= thr();
        };
        const BoxCull bcl{boxcap, org};
        coop_gather(m, qbuf, x, y, cd.c, bcl, thr, keep, [&](uint32_t* q, uint32_t nb) {
            const bool has = lane < nb;
            const uint32_t k = has ? q[lane] : 0u;
            Sel<double> S;
            sel_init<double>(S, -INFINITY);
            double oz = 0.0;
            if (has) {
                const BoxOrigin bo = org[k / kLeaf];
                oz = bo.oz;
                sel_triangle<double, K>(rec + (size_t)kFastRec * k, k, x - bo.ox, y - bo.oy, P, S, true);
            }
            const bool lifted = has && S.bt != 0xffffffffu;
            const double zl = lifted ? S.zb + oz : -INFINITY;
            for (int i = 0; i < 4; ++i) zt[i] = fmax(zt[i], wave_max(has ? S.zt[i] + oz : -INFINITY));
            za = fmax(za, wave_max(has ? S.za + oz : -INFINITY));
            // first lane (lowest rank) holding the batch maximum
            const double zmaxb = wave_max(zl);
            const unsigned long long hit = __ballot(lifted && zl == zmaxb);
            if (hit && zmaxb > zb) {
                const uint32_t i = (uint32_t)__ffsll((long long)hit) - 1u;
                zb = zmaxb;
                eb = __shfl(S.eb, i, 64);
                bt = __shfl(k, i, 64);
                bf = __shfl(S.bf, i, 64);
                bamb = __shfl((int)S.bamb, i, 64) != 0;
            }
        });
        Sel<double> D;
        sel_init(D, zb);
        D.eb = eb;
        D.bt = bt;
        D.bf = bf;
        D.bamb = bamb;
        for (int i = 0; i < 4; ++i) D.zt[i] = zt[i];
        D.za = za;
        bool up = sel_escalate(D, z0, P);
        bool rfail = false;
        DropLane L = fresh_lane(x, y, z0);
        if (!up && bt != 0xffffffffu) {
            // exact refine of the selected feature (lane 0's result is used)
            const TriX t = load_tri(m.tri, bt);
            DropLane X = fresh_lane(x, y, -INFINITY);
            X.cur_tri = bt;
            int it = 0;
            exact_feature(cd.c, X, t, bf, it);
            brent += (lane == 0) ? it : 0;
            if (X.type == CC_NONE || feat_slot(X.feat) != feat_slot(bf) || !(X.z > z0) ||
                fabs(X.z - zb) > 4.0 * (eb + (double)P.E)) {
                up = true;
                rfail = true;
            } else {
                L = X;
            }
        }
        if (up) {
            if (lane == 0) {
                // an unambiguously valid winner is a floor for the exact pass
                const double fl = (bamb || rfail || bt == 0xffffffffu) ? zf : fmax(zf, zb - 2.0 * eb - (double)P.E);
                zfloor[src] = fl;
                next[atomicAdd(next_n, 1u)] = src;
            }
            nesc += (lane == 0);
        } else if (lane == 0) {
            write_drop(o, m.order, src, L);
        }
    }
    if (cnt) {
        wave_add(&cnt->brent, brent);
        wave_add(&cnt->escalated, nesc);
    }
}

// ------------------------------------------------------ stable compaction
// list_out = [list_in[i] for i < n_eff if flag[i]] in order; *count_out = size.
constexpr int kCompactItems = 4096;  // per block of 256 threads

__global__ void __launch_bounds__(256) k_flag_count(const uint8_t* __restrict__ flag, const uint32_t* __restrict__ count,
                                                    uint32_t n, uint32_t* __restrict__ blocksum) {
    const uint32_t n_eff = count ? *count : n;
    __shared__ uint32_t red[256];
    const uint32_t base = blockIdx.x * kCompactItems;
    uint32_t c = 0;
    for (int j = 0; j < kCompactItems / 256; ++j) {
        const uint32_t i = base + j * 256 + threadIdx.x;
        if (i < n_eff) c += flag[i];
    }
    red[threadIdx.x] = c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) blocksum[blockIdx.x] = red[0];
}

// exclusive scan of nb block sums (one block); total -> *count_out
__global__ void __launch_bounds__(1024) k_scan_blocks(uint32_t* __restrict__ blocksum, uint32_t nb,
                                                      uint32_t* __restrict__ count_out) {
    __shared__ uint32_t part[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (nb + 1023) / 1024;
    uint32_t s = 0;
    for (uint32_t k = 0; k < per; ++k) {
        const uint32_t i = t * per + k;
        if (i < nb) s += blocksum[i];
    }
    part[t] = s;
    __syncthreads();
    for (uint32_t o = 1; o < 1024; o <<= 1) {
        const uint32_t v = (t >= o) ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - s;
    for (uint32_t k = 0; k < per; ++k) {
        const uint32_t i = t * per + k;
        if (i < nb) {
            const uint32_t v = blocksum[i];
            blocksum[i] = run;
            run += v;
        }
    }
    if (t == 1023) *count_out = part[1023];
}

__global__ void __launch_bounds__(256) k_compact(const uint8_t* __restrict__ flag, const uint32_t* __restrict__ count,
                                                 uint32_t n, const uint32_t* __restrict__ list_in,
                                                 const uint32_t* __restrict__ blockoff, uint32_t* __restrict__ list_out) {
    const uint32_t n_eff = count ? *count : n;
    __shared__ uint32_t wsum[4];
    uint32_t run = blockoff[blockIdx.x];
    const uint32_t base = blockIdx.x * kCompactItems;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int j = 0; j < kCompactItems / 256; ++j) {
        const uint32_t i = base + j * 256 + threadIdx.x;
        const bool f = (i < n_eff) && flag[i];
        const unsigned long long bal = __ballot(f);
        const uint32_t below = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[wv] = __popcll(bal);
        __syncthreads();
        uint32_t off = run;
        for (int w = 0; w < wv; ++w) off += wsum[w];
        if (f) list_out[off + below] = list_in[i];
        run += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
}

// Empty mesh: every CL point keeps its z and has no contact.
__global__ void k_drop_none(const double* __restrict__ cl, uint32_t n, DropOut o) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    write_drop(o, nullptr, i, fresh_lane(cl[3 * (size_t)i], cl[3 * (size_t)i + 1], cl[3 * (size_t)i + 2]));
}

}  // namespace ocl_amd

#include "push_kernels.hpp"

using namespace ocl_amd;

// ---------------------------------------------------------- fast launcher
template <class T>
static SelErr make_err(const CutterDev& cd, double scale_e) {
    SelErr er;
    er.u = FM<T>::eps() * scale_e;
    er.R = cd.c.radius;
    er.zmag = 4.0 * cd.c.radius + 10.0;
    return er;
}

template <class T>
static SelParams<T> make_params(const CutterDev& cd, const SelErr& er) {
    return sel_params<T>(cd, er);
}

// Makes slot hold the records of cutter cd: kept if they are already its,
// rebuilt in place if nothing else holds them, else built in a new set (a
// batch keeps replaying the old one).
template <class T>
static int ensure_records(ocl_mesh* m, const CutterDev& cd, std::shared_ptr<RecSet>& slot, const SelErr& er,
                          hipStream_t st) {
    if (slot && slot->ok && std::memcmp(&slot->key, &cd, sizeof(cd)) == 0) return OCL_OK;
    if (!slot || slot.use_count() > 1) {
        slot = std::make_shared<RecSet>();
        if (!slot) return fail(OCL_E_NOMEM, "out of host memory");
    }
    RecSet& rs = *slot;
    rs.ok = false;
    const uint32_t n = (uint32_t)m->ntri, nbox = m->cnt[0];
    const bool comp = cd.c.kind == K_COMPOSITE;
    if (comp && sizeof(T) != 4) return fail(OCL_E_UNSUPPORTED, "composite selection records are FP32");
    const uint32_t nrec = comp ? 2u : 1u;
    int rc;
    if ((rc = rs.rec.ensure(sizeof(T) * rec_len<T>() * nrec * m->ntri)) ||
        (rc = rs.cap.ensure(sizeof(T) * nbox + 16)) || (rc = rs.emax.ensure(16)))
        return rc;
    if constexpr (sizeof(T) == 4) {
        if (comp)
            k_sel_prep_comp<<<(n + 255) / 256, 256, 0, st>>>(m->tri.as<double>(), n, m->origins.as<BoxOrigin>(),
                                                             m->fid.as<uint32_t>(), cd, er, rs.rec.as<float>());
    }
    if (!comp)
        k_sel_prep<T><<<(n + 255) / 256, 256, 0, st>>>(m->tri.as<double>(), n, m->origins.as<BoxOrigin>(),
                                                       m->fid.as<uint32_t>(), cd, er, rs.rec.as<T>());
    HIP_TRY(hipMemsetAsync(rs.emax.p, 0, 4, st));
    k_box_cap<T><<<(nbox + 255) / 256, 256, 0, st>>>(rs.rec.as<T>(), n, nbox, nrec, rs.cap.as<T>(),
                                                     rs.emax.as<unsigned int>());
    HIP_TRY(hipGetLastError());
    {
        // absolute caps of every hierarchy box: the FP32 selection's walk
        // cull (FP32 records) and the exact tier's (FP64 records)
        const size_t total = m->box_off[m->top] + m->cnt[m->top];
        if ((rc = rs.capup.ensure(sizeof(float) * total + 16))) return rc;
        if constexpr (sizeof(T) == 4) {
            if (!comp) {
                if ((rc = rs.topv.ensure(sizeof(float4) * nbox + 16))) return rc;
                k_box_topv<<<(nbox + 255) / 256, 256, 0, st>>>(rs.rec.as<float>(), n, nbox, rs.topv.as<float4>());
            }
        }
        float* cu = rs.capup.as<float>();
        k_capup_leaf<T><<<(nbox + 255) / 256, 256, 0, st>>>(rs.cap.as<T>(), m->origins.as<BoxOrigin>(), nbox, cu);
        for (int l = 1; l <= m->top; ++l)
            k_capup_up<<<(m->cnt[l] + 255) / 256, 256, 0, st>>>(cu + m->box_off[l - 1], m->cnt[l - 1],
                                                                 cu + m->box_off[l], m->cnt[l]);
        HIP_TRY(hipGetLastError());
    }
    float em = 0.0f;
    HIP_TRY(hipMemcpyAsync(&em, rs.emax.p, 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    rs.emax_h = em;
    rs.key = cd;
    rs.ok = true;
    return OCL_OK;
}

template <class T>
static void set_cut(SelParams<T>& P, const RecSet& rs) {
    P.CUT = (T)(2.0 * rs.emax_h * (1.0 + 1e-6) + (double)P.HPAD);
}

template <class T>
static SelArgs<T> sel_args(ocl_mesh* m, const MeshView& mv, const T* rec, const T* cap, const SelParams<T>& P,
                           const CutterDev& cd, const double* d_cl, const uint32_t* list, uint32_t n,
                           const DropOut& o, double* zfloor, const NearTieSets& nts, int refine, Counters* cnt) {
    SelArgs<T> A;
    A.m = mv;
    A.rec = rec;
    A.org = m->origins.as<BoxOrigin>();
    A.boxcap = cap;
    A.capup = nullptr;
    A.topv = nullptr;
    A.P = P;
    A.R = cd.c.radius;
    A.cl = d_cl;
    A.list = list;
    A.count = nullptr;
    A.n = n;
    A.o = o;
    A.zfloor = zfloor;
    A.nts = nts;
    A.refine = refine;
    A.cnt = cnt;
    return A;
}

constexpr uint32_t kCoopBlocks = 2048;  // persistent grid of the cooperative tiers
constexpr uint32_t kNearTieBlocks = 512;  // persistent grid of the near-tie pass (8 waves per CU)

static double env_scale(const char* name, double def) {
    const char* v = std::getenv(name);
    return v ? std::atof(v) : def;
}

// Precision pipeline over the Hilbert-sorted points:
//   mixed: FP32 selection -> near-tie pass (Cyl/Ball/Bull) -> FP64 selection
//          of what is still undecided (one wavefront per point) -> exact pass;
//   fp64:  FP64 selection -> near-tie pass -> exact pass;
//   fp32:  FP32 selection only.
// Cones skip the near-tie and FP64 passes (their result depends on the
// evaluation order).  Undecided points carry a floor of their final height
// (zfloor) that lets the later passes skip triangles that cannot matter.
// One job of the fast pipeline: workspace, cached selection records and the
// arguments of its passes.
// Device pointers of one job's per-call workspace (a DropWork's buffers,
// or a carved-up scratch area).
struct WorkPtrs {
    uint32_t *esc_list, *esc_list2, *esc_list3;
    double* zfloor;
    uint32_t* ntset;
    uint8_t* ntcnt;
    double* zband;
    double* selrec;  // 20 bytes per point
    uint32_t* qcount;
    Counters *counters, *cnt2, *cnt3;
    int* exact_slot;  // host: the queue-count slot the exact pass reads
};

struct FastPlan {
    ocl_mesh* m;
    Counters *cnt2, *cnt3;            // counters of the passes after the selection
    int* exact_slot;
    std::shared_ptr<RecSet> r32, r64; // the job's selection records
    MeshView mv;
    CutterDev cd;
    const double* cl;
    const uint32_t* perm;  // selection order (Hilbert-sorted point indices)
    uint32_t n;
    int precision;
    DropOut o;
    Counters* cnt;
    NearTieSets nts;
    double* zfloor;
    uint32_t *q1, *q2, *qnt, *qcount;  // qcount: q1, q2, qnt entries
    SelParams<double> P64;
    SelParams<float> P32;
    SelParams<float> PS32[2];  // composite: the sub-cutters' FP32 constants
    bool comp;                 // composite FP32 selection (no FP64 records)
    double etie;
};

// Workspace and records of a job (ensure_records may synchronise the stream
// once per mesh and cutter: call it before any graph capture).
static WorkPtrs work_ptrs(DropWork& w) {
    WorkPtrs p;
    p.esc_list = w.esc_list.as<uint32_t>();
    p.esc_list2 = w.esc_list2.as<uint32_t>();
    p.esc_list3 = w.esc_list3.as<uint32_t>();
    p.zfloor = w.zfloor.as<double>();
    p.ntset = w.ntset.as<uint32_t>();
    p.ntcnt = w.ntcnt.as<uint8_t>();
    p.zband = w.zband.as<double>();
    p.selrec = w.selrec.as<double>();
    p.qcount = w.qcount.as<uint32_t>();
    p.counters = w.counters.as<Counters>();
    p.cnt2 = w.cnt2.as<Counters>();
    p.cnt3 = w.cnt3.as<Counters>();
    p.exact_slot = &w.exact_slot;
    return p;
}

// The plan's arguments from the workspace pointers and the records already
// in F.r32 / F.r64 (no allocation, no device work).
static void fast_bind(FastPlan& F, ocl_mesh* m, const WorkPtrs& w, const MeshView& mv, const CutterDev& cd,
                      const double* d_cl, const uint32_t* perm, uint32_t n, int precision, const DropOut& o,
                      Counters* cnt) {
    F.m = m;
    F.cnt2 = w.cnt2;
    F.cnt3 = w.cnt3;
    F.exact_slot = w.exact_slot;
    F.mv = mv;
    F.cd = cd;
    F.cl = d_cl;
    F.perm = perm;
    F.n = n;
    F.precision = precision;
    F.o = o;
    F.cnt = cnt;
    F.zfloor = w.zfloor;
    F.nts.set = w.ntset;
    F.nts.cnt = w.ntcnt;
    F.nts.bf = F.nts.cnt + n;
    F.nts.zband = w.zband;
    F.nts.zsel = w.selrec;
    F.nts.bt = reinterpret_cast<uint32_t*>(F.nts.zsel + n);
    F.nts.esel = reinterpret_cast<float*>(F.nts.bt + n);
    F.nts.zgap = F.nts.esel + n;  // the record's 20 bytes: zsel 8, bt 4, esel 4, zgap 4
    F.q1 = w.esc_list;
    F.q2 = w.esc_list2;
    F.qnt = w.esc_list3;
    F.qcount = w.qcount;
    F.nts.q = F.qnt;
    F.nts.qn = F.qcount + 2;
    const double se = env_scale("OCL_SEL_E_SCALE", 1.0);
    F.comp = cd.c.kind == K_COMPOSITE;
    if (!F.comp) {
        const SelErr er64 = make_err<double>(cd, se);
        F.P64 = make_params<double>(cd, er64);
        set_cut(F.P64, *F.r64);
        F.etie = (double)F.P64.E;
    }
    if (F.r32) {
        const SelErr er32 = make_err<float>(cd, se);
        F.P32 = F.comp ? comp_sel_params<float>(cd, er32, F.PS32) : make_params<float>(cd, er32);
        set_cut(F.P32, *F.r32);
        F.etie = (double)F.P32.E;
    }
}

// The precision modes whose pipeline starts with a selection pass: FP32 or
// FP64 for Cyl/Ball/Bull/Cone, FP32 for a composite cutter of two Cyl/Ball/
// Bull parts (comp_fast); every other composite takes the exact fold.
static bool selects(const CutterDev& cd, int precision, size_t ntri) {
    if (precision == OCL_PREC_EXACT64 || ntri == 0) return false;
    if (cd.c.kind != K_COMPOSITE) return true;
    return comp_fast(cd) && (precision == OCL_PREC_FP32 || precision == OCL_PREC_MIXED);
}

static bool uses_fp32_records(int precision) {
    return precision == OCL_PREC_FP32 || precision == OCL_PREC_MIXED;
}

// Workspace and records of a job (ensure_records may synchronise the stream
// once per mesh and cutter: call it before any graph capture).
static int fast_prepare(ocl_mesh* m, DropWork& w, const MeshView& mv, const CutterDev& cd, const double* d_cl,
                        const uint32_t* perm, uint32_t n, int precision, const DropOut& o, Counters* cnt,
                        hipStream_t st, FastPlan& F) {
    int rc;
    if ((rc = w.esc_list.ensure(4ull * n + 64)) || (rc = w.esc_list2.ensure(4ull * n + 64)) ||
        (rc = w.esc_list3.ensure(4ull * n + 64)) ||
        (rc = w.zfloor.ensure(8ull * n + 64)) || (rc = w.ntset.ensure(4ull * kNearTie * n + 64)) ||
        (rc = w.ntcnt.ensure(2ull * n + 64)) || (rc = w.zband.ensure(8ull * n + 64)) ||
        (rc = w.selrec.ensure(20ull * n + 64)) || (rc = w.ensure_counters()))
        return rc;
    const double se = env_scale("OCL_SEL_E_SCALE", 1.0);
    F.r64.reset();
    if (cd.c.kind != K_COMPOSITE) {
        if ((rc = ensure_records<double>(m, cd, m->rs64, make_err<double>(cd, se), st))) return rc;
        F.r64 = m->rs64;
    }
    F.r32.reset();
    if (uses_fp32_records(precision)) {
        if ((rc = ensure_records<float>(m, cd, m->rs32, make_err<float>(cd, se), st))) return rc;
        F.r32 = m->rs32;
    }
    fast_bind(F, m, work_ptrs(w), mv, cd, d_cl, perm, n, precision, o, cnt);
    return OCL_OK;
}

static SelArgs<float> fast_args32(const FastPlan& F) {
    SelArgs<float> A = sel_args<float>(F.m, F.mv, F.r32->rec.as<float>(), F.r32->cap.as<float>(), F.P32, F.cd,
                                       F.cl, F.perm, F.n, F.o, F.zfloor, F.nts,
                                       F.precision == OCL_PREC_MIXED ? 1 : 0, F.cnt);
    A.capup = F.r32->capup.as<const float>();
    A.topv = F.r32->topv.p ? F.r32->topv.as<const float4>() : nullptr;
    return A;
}

static SelArgs<double> fast_args64(const FastPlan& F) {
    return sel_args<double>(F.m, F.mv, F.r64->rec.as<double>(), F.r64->cap.as<double>(), F.P64, F.cd, F.cl,
                            F.perm, F.n, F.o, F.zfloor, F.nts, 1, F.cnt);
}

// Passes after the selection: resolve over all points -> q1 (count slot 0);
// mixed Cyl/Ball/Bull with OCL_COOP_SELECT: FP64 selection of what remains
// -> q2 (slot 1); the exact pass on the last list.  The queue counts must be
// zero when the selection starts.
static int fast_post(const FastPlan& F, hipStream_t st) {
    if (F.precision == OCL_PREC_FP32) return OCL_OK;
    ocl_mesh* m = F.m;
    const MeshView& mv = F.mv;
    const CutterDev& cd = F.cd;
    const uint32_t n = F.n;
    const bool mixed = F.precision == OCL_PREC_MIXED;
    k_drop_resolve<<<(n + 255) / 256, 256, 0, st>>>(mv, cd, F.cl, F.perm, n, F.nts, F.etie, F.o, F.q1, F.qcount,
                                                     F.cnt2);
    HIP_TRY(hipGetLastError());
    if (F.comp) {
        // composite: no near-tie sets, the exact fold of what is undecided
        k_drop_coop_exact_comp<<<kCoopBlocks, 256, 0, st>>>(mv, cd, F.cl, F.q1, F.qcount,
                                                            m->origins.as<BoxOrigin>(), F.o, F.cnt3);
        HIP_TRY(hipGetLastError());
        *F.exact_slot = 0;
        return OCL_OK;
    }
    k_drop_neartie<<<kNearTieBlocks, 256, 0, st>>>(
        mv, cd, F.cl, F.qnt, F.qcount + 2, F.nts, F.o, F.q1, F.qcount, F.cnt2);
    HIP_TRY(hipGetLastError());
    const uint32_t* exact_list = F.q1;
    const uint32_t* exact_n = F.qcount;
    const BoxOrigin* org = m->origins.as<BoxOrigin>();
    const double* cap64 = F.r64->cap.as<double>();
    if (mixed && cd.c.kind != K_CONE && env_scale("OCL_COOP_SELECT", 0.0) != 0.0) {
        Counters* c2 = F.cnt2;
        const double* r64 = F.r64->rec.as<double>();
        const SelParams<double>& P64 = F.P64;
        double* zfloor = F.zfloor;
        uint32_t *q1 = F.q1, *q2 = F.q2, *qcount = F.qcount;
        const double* d_cl = F.cl;
        const DropOut& o = F.o;
        switch (cd.c.kind) {
        case K_CYL: k_drop_coop_select<K_CYL><<<kCoopBlocks, 256, 0, st>>>(mv, r64, org, P64, cd, d_cl, q1, qcount, zfloor, cap64, o, q2, qcount + 1, c2); break;
        case K_BALL: k_drop_coop_select<K_BALL><<<kCoopBlocks, 256, 0, st>>>(mv, r64, org, P64, cd, d_cl, q1, qcount, zfloor, cap64, o, q2, qcount + 1, c2); break;
        default: k_drop_coop_select<K_BULL><<<kCoopBlocks, 256, 0, st>>>(mv, r64, org, P64, cd, d_cl, q1, qcount, zfloor, cap64, o, q2, qcount + 1, c2); break;
        }
        HIP_TRY(hipGetLastError());
        exact_list = F.q2;
        exact_n = F.qcount + 1;
    }
    const float* capup = F.r64->capup.p ? F.r64->capup.as<const float>() : nullptr;
    k_drop_coop_exact<<<kCoopBlocks, 256, 0, st>>>(mv, cd, F.cl, exact_list, exact_n, F.zfloor, cap64, org, F.o,
                                                   F.cnt3, capup);
    HIP_TRY(hipGetLastError());
    *F.exact_slot = (int)(exact_n - F.qcount);
    return OCL_OK;
}

// The selection and the passes after it for a prepared plan.
static int launch_drop_fast(const FastPlan& F, hipStream_t st, bool timed) {
    ocl_mesh* m = F.m;
    const MeshView& mv = F.mv;
    const CutterDev& cd = F.cd;
    const uint32_t n = F.n;
    const int precision = F.precision;
    if (timed) HIP_TRY(hipEventRecord(m->ev_s0, st));
    const uint32_t nb = select_grid(n, mv.xcd_group);
    const bool f32 = precision == OCL_PREC_FP32 || precision == OCL_PREC_MIXED;
    const SelArgs<float> A32 = f32 ? fast_args32(F) : SelArgs<float>{};
    const SelArgs<double> A64 = f32 ? SelArgs<double>{} : fast_args64(F);
    switch (cd.c.kind) {
    case K_CYL: if (f32) k_drop_select<float, K_CYL><<<nb, 64, 0, st>>>(A32); else k_drop_select<double, K_CYL><<<nb, 64, 0, st>>>(A64); break;
    case K_BALL: if (f32) k_drop_select<float, K_BALL><<<nb, 64, 0, st>>>(A32); else k_drop_select<double, K_BALL><<<nb, 64, 0, st>>>(A64); break;
    case K_BULL: if (f32) k_drop_select<float, K_BULL><<<nb, 64, 0, st>>>(A32); else k_drop_select<double, K_BULL><<<nb, 64, 0, st>>>(A64); break;
    case K_CONE: if (f32) k_drop_select<float, K_CONE><<<nb, 64, 0, st>>>(A32); else k_drop_select<double, K_CONE><<<nb, 64, 0, st>>>(A64); break;
    default: {
        if (!F.comp || !f32) return fail(OCL_E_UNSUPPORTED, "composite selection is FP32 only");
        CompSelArgs C;
        C.A = A32;
        C.PS[0] = F.PS32[0];
        C.PS[1] = F.PS32[1];
        k_drop_select_comp<<<nb, 64, 0, st>>>(C);
        break;
    }
    }
    HIP_TRY(hipGetLastError());
    if (timed) HIP_TRY(hipEventRecord(m->ev_s1, st));
    return fast_post(F, st);
}

static CutterP simple_from(const ocl_cutter_simple& s, bool& ok) {
    ok = true;
    switch (s.type) {
    case OCL_CUTTER_CYL: return CutterP::cyl(s.diameter, s.length);
    case OCL_CUTTER_BALL: return CutterP::ball(s.diameter, s.length);
    case OCL_CUTTER_BULL: return CutterP::bull(s.diameter, s.corner_radius, s.length);
    case OCL_CUTTER_CONE: return CutterP::cone(s.diameter, s.angle, s.length);
    default: ok = false; return CutterP{};
    }
}

static int cutter_from_desc(const ocl_cutter_desc* d, CutterDev& cd) {
    if (!d) return fail(OCL_E_ARG, "cutter descriptor is NULL");
    std::memset(&cd, 0, sizeof(cd));
    bool ok = true;
    if (d->c.type == OCL_CUTTER_COMPOSITE) {
        if (d->nsub < 1 || d->nsub > kMaxSub) return fail(OCL_E_ARG, "composite cutter needs 1..4 sub-cutters");
        cd.c.kind = K_COMPOSITE;
        cd.nsub = d->nsub;
        double r = 0;
        for (int i = 0; i < d->nsub; ++i) {
            cd.sub[i] = simple_from(d->sub[i], ok);
            if (!ok || d->sub[i].type == OCL_CUTTER_COMPOSITE)
                return fail(OCL_E_ARG, "composite sub-cutter must be Cyl/Ball/Bull/Cone");
            cd.rlim[i] = d->radius_limit[i];
            cd.hlim[i] = d->height_limit[i];
            cd.zoff[i] = d->zoffset[i];
            if (d->radius_limit[i] > r) r = d->radius_limit[i];
        }
        cd.c.radius = r;
        cd.c.diameter = 2 * r;
        cd.c.length = d->c.length;
    } else {
        cd.c = simple_from(d->c, ok);
        if (!ok) return fail(OCL_E_ARG, "unknown cutter type");
        if (!(cd.c.diameter > 0) || !(cd.c.length > 0)) return fail(OCL_E_ARG, "cutter diameter and length must be > 0");
        if (cd.c.kind == K_BULL && !(cd.c.r1 > 0 && cd.c.r2 > 0))
            return fail(OCL_E_ARG, "BullCutter needs 0 < corner radius < diameter/2");
        if (cd.c.kind == K_CONE && !(cd.c.tan_a > 0)) return fail(OCL_E_ARG, "ConeCutter angle must be in (0, pi/2)");
    }
    return OCL_OK;
}

// bbox per triangle, [minx maxx miny maxy minz maxz] (bbox.cpp:64-90)
static std::vector<double> tri_bboxes(const double* tri_xyz, size_t ntri) {
    std::vector<double> bb(6 * ntri);
    for (size_t i = 0; i < ntri; ++i) {
        const double* q = tri_xyz + 9 * i;
        for (int d = 0; d < 3; ++d) {
            double lo = q[d], hi = q[d];
            for (int k = 1; k < 3; ++k) {
                if (q[3 * k + d] > hi) hi = q[3 * k + d];
                if (q[3 * k + d] < lo) lo = q[3 * k + d];
            }
            bb[6 * i + 2 * d] = lo;
            bb[6 * i + 2 * d + 1] = hi;
        }
    }
    return bb;
}

namespace ocl_amd {
static int scratch_address(void** out);  // the allocation-free entry's workspace (below)
}

static const int kDimsXY[4] = {0, 1, 2, 3};
static const int kDimsYZ[4] = {2, 3, 4, 5};
static const int kDimsXZ[4] = {0, 1, 4, 5};

extern "C" {

const char* ocl_hip_last_error(void) { return g_err.c_str(); }

int ocl_hip_kd_order(const double* tri_xyz, size_t ntri, int dimset, unsigned bucket, uint32_t* order_out,
                     double* accept_out) {
    if (ntri > 0 && (!tri_xyz || !order_out)) return fail(OCL_E_ARG, "NULL pointer");
    if (dimset < 0 || dimset > 2) return fail(OCL_E_ARG, "dimset must be 0 (XY), 1 (YZ) or 2 (XZ)");
    const int* dims = dimset == 0 ? kDimsXY : dimset == 1 ? kDimsYZ : kDimsXZ;
    std::vector<double> bb = tri_bboxes(tri_xyz, ntri);
    std::vector<double> acc;
    std::vector<uint32_t> order = kd_rank_order(bb.data(), ntri, dims, bucket, accept_out ? &acc : nullptr);
    std::copy(order.begin(), order.end(), order_out);
    if (accept_out) std::copy(acc.begin(), acc.end(), accept_out);
    return OCL_OK;
}

size_t ocl_hip_mesh_size(const ocl_mesh* mesh) { return mesh ? mesh->ntri : 0; }

int ocl_hip_mesh_create(const double* tri_xyz, size_t ntri, unsigned bucket, ocl_mesh** out) {
    if (!out) return fail(OCL_E_ARG, "out is NULL");
    *out = nullptr;
    if (ntri > 0 && !tri_xyz) return fail(OCL_E_ARG, "tri_xyz is NULL");
    if (ntri >= 0x7fffffffu) return fail(OCL_E_ARG, "too many triangles");
    // the allocation-free entry's scratch address on this device (looked
    // up once per device, here outside any graph capture)
    void* scratch_at = nullptr;
    if (int rc0 = scratch_address(&scratch_at)) return rc0;
    ocl_mesh* m = new (std::nothrow) ocl_mesh();
    if (!m) return fail(OCL_E_NOMEM, "out of host memory");
    HIP_TRY(hipGetDevice(&m->device));
    m->host_tri.assign(tri_xyz, tri_xyz + 9 * ntri);
    if (const char* h = std::getenv("OCL_HILBERT_BITS")) {
        const int v = std::atoi(h);
        if (v >= 4 && v <= 12) m->hilbert_bits = (uint32_t)v;
    }
    if (const char* g = std::getenv("OCL_XCD_GROUP")) {
        const int v = std::atoi(g);
        if (v >= 1 && v <= 256 && (v & (v - 1)) == 0) m->xcd_group = (uint32_t)v;
    }
    m->ntri = ntri;
    m->bucket = bucket < 1 ? 1 : bucket;
    // host: bbox per triangle and the reference's KD leaf order
    std::vector<double> bb = tri_bboxes(tri_xyz, ntri);
    for (size_t i = 0; i < ntri; ++i) {
        for (int d = 0; d < 6; ++d) {
            const double v = bb[6 * i + d];
            if (i == 0) m->bbox[d] = v;
            else if (d % 2 == 0) m->bbox[d] = std::min(m->bbox[d], v);
            else m->bbox[d] = std::max(m->bbox[d], v);
        }
    }
    if (ntri == 0) {
        *out = m;
        return OCL_OK;
    }
    std::vector<uint32_t> order = kd_rank_order(bb.data(), ntri, kDimsXY, m->bucket);
    std::vector<double> sorted(9 * ntri);
    for (size_t r = 0; r < ntri; ++r) std::memcpy(&sorted[9 * r], tri_xyz + 9 * (size_t)order[r], 9 * sizeof(double));
    int rc;
    std::vector<uint32_t> rank_of(ntri);
    for (size_t r = 0; r < ntri; ++r) rank_of[order[r]] = (uint32_t)r;
    if ((rc = m->tri.ensure(sizeof(double) * kRec * ntri)) || (rc = m->order.ensure(sizeof(uint32_t) * ntri)) ||
        (rc = m->rank_of.ensure(sizeof(uint32_t) * ntri))) {
        delete m;
        return rc;
    }
    DevBuf staging;
    if ((rc = staging.ensure(sizeof(double) * 9 * ntri))) {
        delete m;
        return rc;
    }
    // hierarchy sizes
    uint32_t c = (uint32_t)((ntri + kLeaf - 1) / kLeaf);
    size_t total = 0;
    int lvl = 0;
    for (;;) {
        m->cnt[lvl] = c;
        m->box_off[lvl] = total;
        total += c;
        if (c <= (uint32_t)kFan || lvl + 1 == kMaxLevels) break;
        c = (c + kFan - 1) / kFan;
        ++lvl;
    }
    m->top = lvl;
    if ((rc = m->boxes.ensure(sizeof(float4) * total)) || (rc = m->levels.ensure(sizeof(uint32_t) * 2 * kMaxLevels))) {
        delete m;
        return rc;
    }
    {
        uint32_t lv[2 * kMaxLevels] = {};
        for (int l = 0; l <= m->top; ++l) {
            lv[2 * l] = m->cnt[l];
            lv[2 * l + 1] = (uint32_t)m->box_off[l];
        }
        if (hipMemcpy(m->levels.p, lv, sizeof(lv), hipMemcpyHostToDevice) != hipSuccess) {
            delete m;
            return fail(OCL_E_HIP, "level table upload failed");
        }
    }
    hipError_t e = hipMemcpy(staging.p, sorted.data(), sizeof(double) * 9 * ntri, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(m->order.p, order.data(), sizeof(uint32_t) * ntri, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(m->rank_of.p, rank_of.data(), sizeof(uint32_t) * ntri, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const std::vector<uint32_t> fid = feature_ids(tri_xyz, ntri, order.data());
        if ((rc = m->fid.ensure(sizeof(uint32_t) * fid.size()))) {
            delete m;
            return rc;
        }
        e = hipMemcpy(m->fid.p, fid.data(), sizeof(uint32_t) * fid.size(), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        delete m;
        return fail(OCL_E_HIP, std::string("mesh upload: ") + hipGetErrorString(e));
    }
    const uint32_t n32 = (uint32_t)ntri;
    k_tri_prep<<<(n32 + 255) / 256, 256>>>(staging.as<double>(), m->tri.as<double>(), n32);
    float4* boxes = m->boxes.as<float4>();
    k_box_leaf<<<(m->cnt[0] + 255) / 256, 256>>>(m->tri.as<double>(), n32, boxes, m->cnt[0]);
    for (int l = 1; l <= m->top; ++l)
        k_box_up<<<(m->cnt[l] + 255) / 256, 256>>>(boxes + m->box_off[l - 1], m->cnt[l - 1], boxes + m->box_off[l], m->cnt[l]);
    if ((rc = m->origins.ensure(sizeof(BoxOrigin) * m->cnt[0]))) {
        delete m;
        return rc;
    }
    k_box_frames<<<(m->cnt[0] + 255) / 256, 256>>>(m->tri.as<double>(), n32, boxes, m->cnt[0], m->origins.as<BoxOrigin>());
    if ((rc = m->boxzmax.ensure(sizeof(double) * m->cnt[0] + 16))) {
        delete m;
        return rc;
    }
    k_box_zmax<<<(m->cnt[0] + 255) / 256, 256>>>(m->tri.as<double>(), n32, m->boxzmax.as<double>(), m->cnt[0]);
    e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) {
        delete m;
        return fail(OCL_E_HIP, std::string("mesh prep: ") + hipGetErrorString(e));
    }
    *out = m;
    return OCL_OK;
}

int ocl_hip_stl_read(const char* path, double* tri_xyz_out, size_t cap, size_t* ntri_out) {
    if (!path || !ntri_out) return fail(OCL_E_ARG, "path / ntri_out is NULL");
    std::vector<double> v;
    if (!ocl_amd::stl_read(path, v)) return fail(OCL_E_ARG, std::string("cannot read STL file ") + path);
    const size_t n = v.size() / 9;
    *ntri_out = n;
    if (tri_xyz_out && cap >= n && n) std::memcpy(tri_xyz_out, v.data(), sizeof(double) * v.size());
    return OCL_OK;
}

int ocl_hip_mesh_create_stl(const char* path, unsigned bucket, ocl_mesh** out) {
    if (!out) return fail(OCL_E_ARG, "out is NULL");
    *out = nullptr;
    if (!path) return fail(OCL_E_ARG, "path is NULL");
    std::vector<double> v;
    if (!ocl_amd::stl_read(path, v)) return fail(OCL_E_ARG, std::string("cannot read STL file ") + path);
    return ocl_hip_mesh_create(v.data(), v.size() / 9, bucket, out);
}

int ocl_hip_mesh_destroy(ocl_mesh* mesh) {
    if (!mesh) return OCL_OK;
    if (mesh->batches > 0)
        return fail(OCL_E_ARG, "mesh is used by " + std::to_string(mesh->batches) +
                                   " live batch(es): destroy them first");
    if (mesh->ev0) (void)hipEventDestroy(mesh->ev0);
    if (mesh->ev1) (void)hipEventDestroy(mesh->ev1);
    if (mesh->ev_s0) (void)hipEventDestroy(mesh->ev_s0);
    if (mesh->ev_s1) (void)hipEventDestroy(mesh->ev_s1);
    delete mesh;
    return OCL_OK;
}

int ocl_hip_drop_device(ocl_mesh* m, const ocl_cutter_desc* desc, const double* d_cl, size_t n, int precision,
                        double* d_z, double* d_cc, uint8_t* d_type, uint32_t* d_tri, uint8_t* d_feat,
                        ocl_stats* stats, void* stream_) {
    if (!m) return fail(OCL_E_ARG, "mesh is NULL");
    if (n > 0 && (!d_cl || !d_z)) return fail(OCL_E_ARG, "cl/z pointers are NULL");
    if (n >= 0x7fffffffu) return fail(OCL_E_ARG, "too many CL points");
    if (precision < OCL_PREC_EXACT64 || precision > OCL_PREC_MIXED) return fail(OCL_E_ARG, "unknown precision mode");
    CutterDev cd;
    int rc = cutter_from_desc(desc, cd);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream_;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (n == 0) return OCL_OK;
    const uint32_t n32 = (uint32_t)n;
    DropOut o{d_z, d_cc, d_type, d_tri, d_feat};
    DropWork& w = m->work;
    if ((rc = w.ensure_counters())) return rc;
    Counters* cnt = w.counters.as<Counters>();
    // the counters and queue counts are cleared by k_cl_keys
    const ZeroSet zs{{reinterpret_cast<uint32_t*>(cnt), w.cnt2.as<uint32_t>(), w.cnt3.as<uint32_t>(),
                      w.qcount.as<uint32_t>()},
                     {(uint32_t)(sizeof(Counters) / 4), (uint32_t)(sizeof(Counters) / 4),
                      (uint32_t)(sizeof(Counters) / 4), 4u}};
    // bin CL points
    if ((rc = m->keys.ensure(8ull * n)) || (rc = m->hist.ensure(4ull * rs_scratch_words((uint32_t)n)))) return rc;
    const double R = cd.c.radius;
    BinFrame f;
    f.x0 = m->bbox[0] - R;
    f.y0 = m->bbox[2] - R;
    const double ex = std::max(m->bbox[1] + R - f.x0, 1e-9), ey = std::max(m->bbox[3] + R - f.y0, 1e-9);
    const uint32_t hb = m->hilbert_bits;
    f.sx = (double)(1u << hb) / ex;
    f.sy = (double)(1u << hb) / ey;
    const uint32_t nb = (n32 + 255) / 256;
    uint32_t* keys = m->keys.as<uint32_t>();
    k_cl_keys<<<nb, 256, 0, st>>>(d_cl, n32, f, hb, keys, keys + n32, zs);
    // stable: equal keys keep their input order, so the point -> wavefront
    // assignment, and with it every result, is the same on every run
    uint32_t *ko = nullptr, *perm = nullptr;
    rs_sort(keys, keys + n32, n32, 2 * hb, m->hist.as<uint32_t>(), &ko, &perm, st);
    HIP_TRY(hipGetLastError());
    if (stats) {
        if (!m->ev0) HIP_TRY(hipEventCreate(&m->ev0));
        if (!m->ev1) HIP_TRY(hipEventCreate(&m->ev1));
        if (!m->ev_s0) HIP_TRY(hipEventCreate(&m->ev_s0));
        if (!m->ev_s1) HIP_TRY(hipEventCreate(&m->ev_s1));
        HIP_TRY(hipEventRecord(m->ev0, st));
    }
    const MeshView mv = m->view();
    bool fast = false;
    if (!selects(cd, precision, m->ntri)) {
        if (m->ntri > 0)
            k_drop_exact<<<xcd_grid((n32 + 63) / 64, mv.xcd_group * 4u), 64, 0, st>>>(mv, cd, d_cl, perm, nullptr,
                                                                                    n32, o, cnt);
        else
            k_drop_none<<<nb, 256, 0, st>>>(d_cl, n32, o);
    } else {
        // the selection's pair/escalation counters feed only the statistics:
        // without them its wavefronts skip three same-address atomics each
        FastPlan F;
        if ((rc = fast_prepare(m, w, mv, cd, d_cl, perm, n32, precision, o, stats ? cnt : nullptr, st, F))) return rc;
        rc = launch_drop_fast(F, st, stats != nullptr);
        fast = true;
        if (rc) return rc;
    }
    HIP_TRY(hipGetLastError());
    if (stats) {
        HIP_TRY(hipEventRecord(m->ev1, st));
        HIP_TRY(hipEventSynchronize(m->ev1));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, m->ev0, m->ev1));
        Counters hc, h2, h3;
        HIP_TRY(hipMemcpy(&hc, cnt, sizeof(hc), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(&h2, w.cnt2.p, sizeof(h2), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(&h3, w.cnt3.p, sizeof(h3), hipMemcpyDeviceToHost));
        uint32_t qc[4] = {0, 0, 0, 0};
        if (fast && precision != OCL_PREC_FP32) HIP_TRY(hipMemcpy(qc, w.qcount.p, 16, hipMemcpyDeviceToHost));
        stats->points = n;
        stats->pairs = hc.pairs;
        stats->escalated = hc.escalated;
        stats->escalated_exact = fast ? qc[w.exact_slot] : 0;
        stats->brent_iters = hc.brent + h2.brent + h3.brent;
        stats->wave_evals = hc.evals;
        stats->kernel_ms = ms;
        stats->select_ms = ms;
        if (fast) {
            HIP_TRY(hipEventElapsedTime(&ms, m->ev_s0, m->ev_s1));
            stats->select_ms = ms;
        }
    }
    return OCL_OK;
}

}  // extern "C"

// ------------------------------------------------- allocation-free drop
namespace ocl_amd {

// Per-call workspace of ocl_hip_drop_device_noalloc, for up to
// kNoAllocPoints CL points: the Hilbert keys and the sort's scratch, the
// per-point selection records, near-tie sets, floors and escalation queues,
// and the call's counters.  One fixed array per GPU, held until the process
// ends; every call on a GPU shares it, so calls that overlap on one GPU
// race.  Its address on each device comes from hipGetSymbolAddress, never
// from the array's name in host code.
constexpr uint32_t kNoAllocPoints = 1u << 19;

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

constexpr size_t noalloc_bytes(uint32_t n) {
    return align256(8ull * n)                           // keys + values
           + align256(4ull * rs_scratch_words(n))       // radix sort scratch
           + 3 * align256(4ull * n + 64)                // escalation queues
           + align256(8ull * n + 64)                    // floors
           + align256(4ull * kNearTie * n + 64)         // near-tie sets
           + align256(2ull * n + 64)                    // set sizes, winner features
           + align256(8ull * n + 64)                    // band bottoms
           + align256(20ull * n + 64)                   // selection records
           + 3 * align256(sizeof(Counters)) + align256(16);  // counters, queue counts
}

static_assert(noalloc_bytes(kNoAllocPoints) <= (size_t)52 << 20, "the stated scratch size (52 MiB) is exceeded");
static __device__ __align__(256) unsigned char scratch[noalloc_bytes(kNoAllocPoints)];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_at[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (mesh creation does it first, outside any graph capture).
static int scratch_address(void** out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(OCL_E_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_at[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(OCL_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_at[dev] = p;
    }
    *out = g_scratch_at[dev];
    return OCL_OK;
}

// Carves the workspace of an n-point call out of the scratch area.
struct NoAllocWork {
    uint32_t* keys;   // 2n words: keys, then values
    uint32_t* hist;   // rs_scratch_words(n)
    WorkPtrs w;
};

static NoAllocWork carve_noalloc(unsigned char* base, uint32_t n, int* exact_slot) {
    NoAllocWork a;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        unsigned char* p = base + off;
        off += align256(bytes);
        return p;
    };
    a.keys = reinterpret_cast<uint32_t*>(take(8ull * n));
    a.hist = reinterpret_cast<uint32_t*>(take(4ull * rs_scratch_words(n)));
    a.w.esc_list = reinterpret_cast<uint32_t*>(take(4ull * n + 64));
    a.w.esc_list2 = reinterpret_cast<uint32_t*>(take(4ull * n + 64));
    a.w.esc_list3 = reinterpret_cast<uint32_t*>(take(4ull * n + 64));
    a.w.zfloor = reinterpret_cast<double*>(take(8ull * n + 64));
    a.w.ntset = reinterpret_cast<uint32_t*>(take(4ull * kNearTie * n + 64));
    a.w.ntcnt = take(2ull * n + 64);
    a.w.zband = reinterpret_cast<double*>(take(8ull * n + 64));
    a.w.selrec = reinterpret_cast<double*>(take(20ull * n + 64));
    a.w.counters = reinterpret_cast<Counters*>(take(sizeof(Counters)));
    a.w.cnt2 = reinterpret_cast<Counters*>(take(sizeof(Counters)));
    a.w.cnt3 = reinterpret_cast<Counters*>(take(sizeof(Counters)));
    a.w.qcount = reinterpret_cast<uint32_t*>(take(16));
    a.w.exact_slot = exact_slot;
    return a;
}

static bool records_hold(const std::shared_ptr<RecSet>& rs, const CutterDev& cd) {
    return rs && rs->ok && std::memcmp(&rs->key, &cd, sizeof(cd)) == 0;
}

}  // namespace ocl_amd

extern "C" int ocl_hip_drop_device_noalloc(ocl_mesh* m, const ocl_cutter_desc* desc, const double* d_cl, size_t n,
                                           int precision, double* d_z, double* d_cc, uint8_t* d_type,
                                           uint32_t* d_tri, uint8_t* d_feat, void* stream_) {
    using namespace ocl_amd;
    if (!m) return fail(OCL_E_ARG, "mesh is NULL");
    if (n > 0 && (!d_cl || !d_z)) return fail(OCL_E_ARG, "cl/z pointers are NULL");
    if (n > kNoAllocPoints)
        return fail(OCL_E_ARG, "at most " + std::to_string(kNoAllocPoints) +
                                   " CL points per allocation-free call (split the points into several calls)");
    if (precision < OCL_PREC_EXACT64 || precision > OCL_PREC_MIXED) return fail(OCL_E_ARG, "unknown precision mode");
    CutterDev cd;
    int rc = cutter_from_desc(desc, cd);
    if (rc) return rc;
    const bool exact = !selects(cd, precision, m->ntri);
    // the selection records are mesh state built by an allocating call
    const bool need64 = cd.c.kind != K_COMPOSITE;
    if (!exact && ((need64 && !records_hold(m->rs64, cd)) ||
                   (uses_fp32_records(precision) && !records_hold(m->rs32, cd))))
        return fail(OCL_E_ARG, "the mesh holds no selection records for this cutter and precision: call "
                               "ocl_hip_drop_device once with them first (that call builds them)");
    if (n == 0) return OCL_OK;
    void* base = nullptr;
    if ((rc = scratch_address(&base))) return rc;
    hipStream_t st = (hipStream_t)stream_;
    const uint32_t n32 = (uint32_t)n;
    int exact_slot = 0;
    const NoAllocWork a = carve_noalloc(static_cast<unsigned char*>(base), n32, &exact_slot);
    const DropOut o{d_z, d_cc, d_type, d_tri, d_feat};
    // the counters and queue counts are cleared by k_cl_keys (no memset)
    const ZeroSet zs{{reinterpret_cast<uint32_t*>(a.w.counters), reinterpret_cast<uint32_t*>(a.w.cnt2),
                      reinterpret_cast<uint32_t*>(a.w.cnt3), a.w.qcount},
                     {(uint32_t)(sizeof(Counters) / 4), (uint32_t)(sizeof(Counters) / 4),
                      (uint32_t)(sizeof(Counters) / 4), 4u}};
    const double R = cd.c.radius;
    BinFrame f;
    f.x0 = m->bbox[0] - R;
    f.y0 = m->bbox[2] - R;
    const double ex = std::max(m->bbox[1] + R - f.x0, 1e-9), ey = std::max(m->bbox[3] + R - f.y0, 1e-9);
    const uint32_t hb = m->hilbert_bits;
    f.sx = (double)(1u << hb) / ex;
    f.sy = (double)(1u << hb) / ey;
    const uint32_t nb = (n32 + 255) / 256;
    k_cl_keys<<<nb, 256, 0, st>>>(d_cl, n32, f, hb, a.keys, a.keys + n32, zs);
    uint32_t *ko = nullptr, *perm = nullptr;
    rs_sort(a.keys, a.keys + n32, n32, 2 * hb, a.hist, &ko, &perm, st);
    HIP_TRY(hipGetLastError());
    const MeshView mv = m->view();
    if (exact) {
        if (m->ntri > 0)
            k_drop_exact<<<xcd_grid((n32 + 63) / 64, mv.xcd_group * 4u), 64, 0, st>>>(mv, cd, d_cl, perm, nullptr,
                                                                                    n32, o, a.w.counters);
        else
            k_drop_none<<<nb, 256, 0, st>>>(d_cl, n32, o);
        HIP_TRY(hipGetLastError());
        return OCL_OK;
    }
    FastPlan F;
    if (need64) F.r64 = m->rs64;
    if (uses_fp32_records(precision)) F.r32 = m->rs32;
    fast_bind(F, m, a.w, mv, cd, d_cl, perm, n32, precision, o, nullptr);
    return launch_drop_fast(F, st, false);
}

// Builds a rank-ordered triangle copy with its box hierarchy over
// (coordinate cu, z), for the push trees (dims = the reference KD-tree's
// search dimensions).
static int build_tree(ocl_mesh* m, ocl_tree& T, const int dims[4], int cu) {
    const size_t ntri = m->ntri;
    std::vector<double> bb = tri_bboxes(m->host_tri.data(), ntri);
    std::vector<double> acc;
    std::vector<uint32_t> order = kd_rank_order(bb.data(), ntri, dims, m->bucket, &acc);
    std::vector<double> sorted(9 * ntri);
    for (size_t r = 0; r < ntri; ++r)
        std::memcpy(&sorted[9 * r], m->host_tri.data() + 9 * (size_t)order[r], 9 * sizeof(double));
    int rc;
    DevBuf staging;
    if ((rc = T.tri.ensure(sizeof(double) * kRec * ntri)) || (rc = T.order.ensure(sizeof(uint32_t) * ntri)) ||
        (rc = staging.ensure(sizeof(double) * 9 * ntri)) || (rc = T.acc.ensure(sizeof(double) * 4 * ntri)))
        return rc;
    HIP_TRY(hipMemcpy(T.acc.p, acc.data(), sizeof(double) * 4 * ntri, hipMemcpyHostToDevice));
    uint32_t c = (uint32_t)((ntri + kLeaf - 1) / kLeaf);
    size_t total = 0;
    int lvl = 0;
    for (;;) {
        T.cnt[lvl] = c;
        T.box_off[lvl] = total;
        total += c;
        if (c <= (uint32_t)kFan || lvl + 1 == kMaxLevels) break;
        c = (c + kFan - 1) / kFan;
        ++lvl;
    }
    T.top = lvl;
    if ((rc = T.boxes.ensure(sizeof(float4) * total)) || (rc = T.levels.ensure(sizeof(uint32_t) * 2 * kMaxLevels)))
        return rc;
    uint32_t lv[2 * kMaxLevels] = {};
    for (int l = 0; l <= T.top; ++l) {
        lv[2 * l] = T.cnt[l];
        lv[2 * l + 1] = (uint32_t)T.box_off[l];
    }
    HIP_TRY(hipMemcpy(T.levels.p, lv, sizeof(lv), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(staging.p, sorted.data(), sizeof(double) * 9 * ntri, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(T.order.p, order.data(), sizeof(uint32_t) * ntri, hipMemcpyHostToDevice));
    const uint32_t n32 = (uint32_t)ntri;
    k_tri_prep<<<(n32 + 255) / 256, 256>>>(staging.as<double>(), T.tri.as<double>(), n32);
    k_push_boundrec<<<(n32 + 255) / 256, 256>>>(T.tri.as<double>(), n32, cu ^ 1);
    float4* boxes = T.boxes.as<float4>();
    k_box_leaf_uz<<<(T.cnt[0] + 255) / 256, 256>>>(T.tri.as<double>(), n32, boxes, T.cnt[0], cu);
    for (int l = 1; l <= T.top; ++l)
        k_box_up<<<(T.cnt[l] + 255) / 256, 256>>>(boxes + T.box_off[l - 1], T.cnt[l - 1], boxes + T.box_off[l], T.cnt[l]);
    if ((rc = T.kdboxes.ensure(sizeof(float4) * total))) return rc;
    float4* kb = T.kdboxes.as<float4>();
    k_kdbox_leaf<<<(T.cnt[0] + 255) / 256, 256>>>(T.acc.as<double>(), n32, kb, T.cnt[0]);
    for (int l = 1; l <= T.top; ++l)
        k_box_up<<<(T.cnt[l] + 255) / 256, 256>>>(kb + T.box_off[l - 1], T.cnt[l - 1], kb + T.box_off[l], T.cnt[l]);
    if ((rc = T.reach.ensure(sizeof(float4) * total))) return rc;
    float4* rb = T.reach.as<float4>();
    k_box_leaf_reach<<<(T.cnt[0] + 255) / 256, 256>>>(T.tri.as<double>(), n32, rb, T.cnt[0], cu);
    for (int l = 1; l <= T.top; ++l)
        k_box_up<<<(T.cnt[l] + 255) / 256, 256>>>(rb + T.box_off[l - 1], T.cnt[l - 1], rb + T.box_off[l], T.cnt[l]);
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipGetLastError());
    T.built = true;
    return OCL_OK;
}

static MeshView tree_view(const ocl_tree& T, size_t ntri) {
    MeshView v;
    v.tri = T.tri.as<const double>();
    v.order = T.order.as<const uint32_t>();
    v.ntri = (uint32_t)ntri;
    v.xcd_group = 1;
    v.tree.box = T.boxes.as<const float4>();
    v.tree.lvl = T.levels.as<const uint32_t>();
    v.tree.top = T.top;
    v.bzmax = nullptr;
    return v;
}

namespace ocl_amd {

// Checks one direction's fibers and uploads them, sorted by (z, across) so
// that neighbouring wavefronts take neighbouring fibers.
static int push_side_prepare(ocl_mesh* m, int dir, const double* fib, size_t nf, hipStream_t st) {
    if (nf > 0 && !fib) return fail(OCL_E_ARG, "fiber pointer is NULL");
    if (nf >= 0x7fffffffu) return fail(OCL_E_ARG, "too many fibers");
    // Fiber::calcDir asserts a horizontal fiber; the facet push tests
    // p1.y == p2.y (X fiber) or p1.x == p2.x (Y fiber)
    for (size_t i = 0; i < nf; ++i) {
        const double* q = fib + 6 * i;
        if (q[2] != q[5]) return fail(OCL_E_ARG, "fiber " + std::to_string(i) + " is not horizontal");
        if (dir == 0 ? (q[1] != q[4]) : (q[0] != q[3]))
            return fail(OCL_E_ARG, "fiber " + std::to_string(i) + " does not run along the requested axis");
    }
    PushWork& W = m->pw[dir];
    // pool: 4 entries per fiber, or twice what the previous push in this
    // direction kept; fibers that do not fit go to the retry pass
    W.pool_cap = std::max<uint64_t>(4ull * nf, 2ull * W.total) + 1024;
    W.nf = (uint32_t)nf;
    W.total = 0;
    W.rcap = 0;
    W.nretry = 0;
    if (nf == 0 || m->ntri == 0) return OCL_OK;
    ocl_tree& T = m->push_tree[dir];
    if (!T.built) {
        static const int kYZ[4] = {2, 3, 4, 5}, kXZ[4] = {0, 1, 4, 5};
        int rc = build_tree(m, T, dir == 0 ? kYZ : kXZ, dir == 0 ? 1 : 0);
        if (rc) return rc;
    }
    std::vector<uint32_t> perm(nf);
    for (size_t i = 0; i < nf; ++i) perm[i] = (uint32_t)i;
    const int ca = dir == 0 ? 1 : 0;
    std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) {
        const double za = fib[6 * (size_t)a + 2], zb = fib[6 * (size_t)b + 2];
        if (za != zb) return za < zb;
        return fib[6 * (size_t)a + ca] < fib[6 * (size_t)b + ca];
    });
    int rc;
    if ((rc = W.fib.ensure(48ull * nf)) || (rc = W.perm.ensure(4ull * nf)) || (rc = W.n.ensure(4ull * nf + 64)) ||
        (rc = W.off.ensure(8ull * nf + 64)) || (rc = W.retry.ensure(4ull * nf + 64)) ||
        (rc = W.cnt.ensure(kPushCntBytes)) || (rc = W.pool.ensure(sizeof(PEntry) * W.pool_cap)))
        return rc;
    HIP_TRY(hipMemcpyAsync(W.fib.p, fib, 48ull * nf, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(W.perm.p, perm.data(), 4ull * nf, hipMemcpyHostToDevice, st));
    return OCL_OK;
}

static PushLists push_lists(PushWork& W) {
    unsigned char* c = W.cnt.as<unsigned char>();
    return PushLists{W.pool.as<PEntry>(), reinterpret_cast<unsigned long long*>(c + 8), W.pool_cap,
                     W.off.as<uint64_t>(), W.n.as<uint32_t>(), W.retry.as<uint32_t>(),
                     reinterpret_cast<uint32_t*>(c + (W.rcap ? 4 : 0)), W.big.as<PEntry>(), W.rcap};
}

// One side of a launch: the first pass over all of the direction's fibers,
// or (W.rcap > 0) the retry pass over the fibers listed in W.retry.
static PushSide push_side(ocl_mesh* m, int dir, int superset) {
    PushWork& W = m->pw[dir];
    const ocl_tree& T = m->push_tree[dir];
    PushSide S;
    S.pm = tree_view(T, m->ntri);
    if (superset) S.pm.tree.box = T.kdboxes.as<const float4>();  // walk the KD acceptance boxes
    S.acc = T.acc.as<const double>();
    S.reach = T.reach.as<const float4>();
    S.fib = W.fib.as<const double>();
    S.perm = W.rcap ? W.retry.as<const uint32_t>() : W.perm.as<const uint32_t>();
    S.nf = W.rcap ? W.nretry : W.nf;
    S.dir = dir;
    S.PL = push_lists(W);
    return S;
}

// Pushes the directions dirs[0..nd) in one launch (first pass), then the
// fibers whose lists did not fit in LDS or in the pool in retry passes over
// those fibers only, with list capacity 256, then 4x as long as some list
// outgrows it.  Adds the device time of every launch to *ms and the counters
// to *hc.
static int push_run(ocl_mesh* m, const CutterDev& cd, int superset, int skip, const int* dirs, int nd,
                    hipStream_t st, float* ms, PushCounters* hc) {
    int rc;
    if ((rc = m->push_cnt.ensure(sizeof(PushCounters))) || (rc = m->push_queue.ensure(64))) return rc;
    HIP_TRY(hipMemsetAsync(m->push_cnt.p, 0, sizeof(PushCounters), st));
    if (!m->ev0) HIP_TRY(hipEventCreate(&m->ev0));
    if (!m->ev1) HIP_TRY(hipEventCreate(&m->ev1));
    for (int i = 0; i < nd; ++i) {
        PushWork& W = m->pw[dirs[i]];
        W.rcap = 0;
        HIP_TRY(hipMemsetAsync(W.cnt.p, 0, kPushCntBytes, st));
    }
    int todo[2] = {dirs[0], nd > 1 ? dirs[1] : -1};
    int nt = nd;
    while (nt > 0) {
        PushSide S[2];
        uint32_t nfib = 0;
        for (int i = 0; i < nt; ++i) {
            PushWork& W = m->pw[todo[i]];
            if (W.rcap) {
                if ((rc = W.big.ensure(sizeof(PEntry) * (size_t)W.rcap * W.nretry))) return rc;
                HIP_TRY(hipMemsetAsync(W.cnt.as<unsigned char>() + 4, 0, 4, st));
            }
            S[i] = push_side(m, todo[i], superset);
            nfib += S[i].nf;
        }
        if ((rc = m->push_sides.ensure(sizeof(S)))) return rc;
        HIP_TRY(hipMemcpyAsync(m->push_sides.p, S, sizeof(PushSide) * nt, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemsetAsync(m->push_queue.p, 0, 4 * kXcds, st));
        HIP_TRY(hipEventRecord(m->ev0, st));
        const uint32_t nb = (uint32_t)std::min<size_t>((nfib + kPushWaves - 1) / kPushWaves, 2048);
        uint32_t* q = m->push_queue.as<uint32_t>();
        PushCounters* pc = m->push_cnt.as<PushCounters>();
        const uint32_t ns = (uint32_t)nt;
        const PushSide* sd = m->push_sides.as<const PushSide>();
        switch (cd.c.kind) {
        case K_CYL: k_push_fibers<K_CYL><<<nb, 256, 0, st>>>(sd, ns, q, superset, skip, cd, pc); break;
        case K_BALL: k_push_fibers<K_BALL><<<nb, 256, 0, st>>>(sd, ns, q, superset, skip, cd, pc); break;
        case K_BULL: k_push_fibers<K_BULL><<<nb, 256, 0, st>>>(sd, ns, q, superset, skip, cd, pc); break;
        case K_CONE: k_push_fibers<K_CONE><<<nb, 256, 0, st>>>(sd, ns, q, superset, skip, cd, pc); break;
        default: k_push_fibers<K_COMPOSITE><<<nb, 256, 0, st>>>(sd, ns, q, superset, skip, cd, pc); break;
        }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(m->ev1, st));
        uint32_t cnt[2][2] = {{0, 0}, {0, 0}};
        for (int i = 0; i < nt; ++i)
            HIP_TRY(hipMemcpyAsync(cnt[i], m->pw[todo[i]].cnt.p, 8, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        float t = 0;
        HIP_TRY(hipEventElapsedTime(&t, m->ev0, m->ev1));
        *ms += t;
        int again = 0;
        for (int i = 0; i < nt; ++i) {
            PushWork& W = m->pw[todo[i]];
            if (W.rcap == 0) {
                W.nretry = cnt[i][0];
                if (W.nretry == 0) continue;
                W.rcap = 256;
            } else {
                if (cnt[i][1] == 0) continue;
                if (W.rcap >= kMaxFiberList)
                    return fail(OCL_E_UNSUPPORTED,
                                "push: more than " + std::to_string(kMaxFiberList) + " intervals on one fiber");
                W.rcap *= 4;
            }
            todo[again++] = todo[i];
        }
        nt = again;
    }
    HIP_TRY(hipMemcpy(hc, m->push_cnt.p, sizeof(*hc), hipMemcpyDeviceToHost));
    return OCL_OK;
}

// Row offsets of one direction's last push (host and device) and its total
// (the next push in this direction sizes its pool from it).
static int push_side_finish(ocl_mesh* m, int dir, uint64_t* offsets_out) {
    PushWork& W = m->pw[dir];
    const size_t nf = W.nf;
    if (nf == 0 || m->ntri == 0) {
        for (size_t i = 0; i <= nf; ++i) offsets_out[i] = 0;
        W.total = 0;
        return OCL_OK;
    }
    std::vector<uint32_t> cnt(nf);
    HIP_TRY(hipMemcpy(cnt.data(), W.n.p, 4ull * nf, hipMemcpyDeviceToHost));
    uint64_t run = 0;
    for (size_t i = 0; i < nf; ++i) {
        offsets_out[i] = run;
        run += cnt[i];
    }
    offsets_out[nf] = run;
    W.total = run;
    int rc;
    if ((rc = W.offs.ensure(8ull * (nf + 1)))) return rc;
    HIP_TRY(hipMemcpy(W.offs.p, offsets_out, 8ull * (nf + 1), hipMemcpyHostToDevice));
    return OCL_OK;
}

// EXACT64 pushes every candidate, as the reference does; the other modes
// leave out the candidates that cannot change their fiber (same results).
static int push_precision(int& precision, int& superset, int& skip) {
    superset = (precision & OCL_PUSH_EXACT_FILTER) ? 0 : 1;
    precision &= ~OCL_PUSH_EXACT_FILTER;
    if (precision < OCL_PREC_EXACT64 || precision > OCL_PREC_MIXED) return fail(OCL_E_ARG, "unknown precision mode");
    skip = precision != OCL_PREC_EXACT64;
    return OCL_OK;
}

static void push_stats(ocl_stats* stats, uint64_t fibers, uint64_t rows, const PushCounters& hc, float ms) {
    if (!stats) return;
    stats->points = fibers;
    stats->pairs = hc.pairs;
    stats->intervals = rows;
    stats->brent_iters = hc.brent;
    stats->skipped = hc.skipped;
    stats->kernel_ms = ms;
    stats->select_ms = ms;
}

static int push_fetch_side(ocl_mesh* m, int dir, double* lohi, double* cc, hipStream_t st) {
    PushWork& W = m->pw[dir];
    const uint64_t total = W.total;
    if (total == 0) return OCL_OK;
    if (!lohi || !cc) return fail(OCL_E_ARG, "output pointers are NULL");
    int rc;
    if ((rc = W.lohi.ensure(16ull * total)) || (rc = W.cc.ensure(64ull * total))) return rc;
    const PushLists PL = push_lists(W);
    const uint32_t nf = W.nf;
    k_push_gather<<<(nf + 255) / 256, 256, 0, st>>>(PL, W.offs.as<uint64_t>(), nf, W.lohi.as<double>(),
                                                    W.cc.as<double>());
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(lohi, W.lohi.p, 16ull * total, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(cc, W.cc.p, 64ull * total, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return OCL_OK;
}

}  // namespace ocl_amd

extern "C" {

int ocl_hip_push_batch(ocl_mesh* m, const ocl_cutter_desc* desc, int dir, const double* fib, size_t nf,
                       int precision, uint64_t* offsets_out, uint64_t* total_out, ocl_stats* stats, void* stream_) {
    using namespace ocl_amd;
    if (!m) return fail(OCL_E_ARG, "mesh is NULL");
    if (dir != 0 && dir != 1) return fail(OCL_E_ARG, "dir must be 0 (X fibers) or 1 (Y fibers)");
    if (nf > 0 && !offsets_out) return fail(OCL_E_ARG, "offset pointer is NULL");
    int superset, skip, rc;
    if ((rc = push_precision(precision, superset, skip))) return rc;
    CutterDev cd;
    if ((rc = cutter_from_desc(desc, cd))) return rc;
    hipStream_t st = (hipStream_t)stream_;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (total_out) *total_out = 0;
    m->push_last = dir;
    if ((rc = push_side_prepare(m, dir, fib, nf, st))) return rc;
    if (nf == 0) {
        if (offsets_out) offsets_out[0] = 0;
        return OCL_OK;
    }
    PushCounters hc{};
    float ms = 0;
    if (m->ntri > 0 && (rc = push_run(m, cd, superset, skip, &dir, 1, st, &ms, &hc))) return rc;
    if ((rc = push_side_finish(m, dir, offsets_out))) return rc;
    if (total_out) *total_out = m->pw[dir].total;
    push_stats(stats, nf, m->pw[dir].total, hc, ms);
    return OCL_OK;
}

int ocl_hip_push_xy(ocl_mesh* m, const ocl_cutter_desc* desc, const double* xfib, size_t nx, const double* yfib,
                    size_t ny, int precision, uint64_t* x_offsets_out, uint64_t* y_offsets_out, uint64_t* totals_out,
                    ocl_stats* stats, void* stream_) {
    using namespace ocl_amd;
    if (!m) return fail(OCL_E_ARG, "mesh is NULL");
    if ((nx > 0 && !x_offsets_out) || (ny > 0 && !y_offsets_out)) return fail(OCL_E_ARG, "offset pointers are NULL");
    int superset, skip, rc;
    if ((rc = push_precision(precision, superset, skip))) return rc;
    CutterDev cd;
    if ((rc = cutter_from_desc(desc, cd))) return rc;
    hipStream_t st = (hipStream_t)stream_;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (totals_out) totals_out[0] = totals_out[1] = 0;
    if ((rc = push_side_prepare(m, 0, xfib, nx, st)) || (rc = push_side_prepare(m, 1, yfib, ny, st))) return rc;
    int dirs[2], nd = 0;
    if (nx > 0) dirs[nd++] = 0;
    if (ny > 0) dirs[nd++] = 1;
    PushCounters hc{};
    float ms = 0;
    if (nd > 0 && m->ntri > 0 && (rc = push_run(m, cd, superset, skip, dirs, nd, st, &ms, &hc))) return rc;
    if (x_offsets_out && (rc = push_side_finish(m, 0, x_offsets_out))) return rc;
    if (y_offsets_out && (rc = push_side_finish(m, 1, y_offsets_out))) return rc;
    if (totals_out) {
        totals_out[0] = m->pw[0].total;
        totals_out[1] = m->pw[1].total;
    }
    push_stats(stats, nx + ny, m->pw[0].total + m->pw[1].total, hc, ms);
    return OCL_OK;
}

int ocl_hip_push_fetch(ocl_mesh* m, double* lohi, double* cc, void* stream_) {
    if (!m) return ocl_amd::fail(OCL_E_ARG, "mesh is NULL");
    return ocl_amd::push_fetch_side(m, m->push_last, lohi, cc, (hipStream_t)stream_);
}

int ocl_hip_push_fetch_dir(ocl_mesh* m, int dir, double* lohi, double* cc, void* stream_) {
    if (!m) return ocl_amd::fail(OCL_E_ARG, "mesh is NULL");
    if (dir != 0 && dir != 1) return ocl_amd::fail(OCL_E_ARG, "dir must be 0 (X fibers) or 1 (Y fibers)");
    return ocl_amd::push_fetch_side(m, dir, lohi, cc, (hipStream_t)stream_);
}

int ocl_hip_drop_batch(ocl_mesh* m, const ocl_cutter_desc* desc, const double* cl, size_t n, int precision,
                       double* z, double* cc, uint8_t* type, uint32_t* tri, uint8_t* feat, ocl_stats* stats,
                       void* stream_) {
    if (!m) return fail(OCL_E_ARG, "mesh is NULL");
    if (n > 0 && (!cl || !z)) return fail(OCL_E_ARG, "cl/z pointers are NULL");
    hipStream_t st = (hipStream_t)stream_;
    int rc;
    if ((rc = m->cl.ensure(24ull * n + 8)) || (rc = m->out_z.ensure(8ull * n + 8)) ||
        (rc = m->out_cc.ensure(24ull * n + 8)) || (rc = m->out_ty.ensure(n + 8)) ||
        (rc = m->out_tri.ensure(4ull * n + 8)) || (rc = m->out_ft.ensure(n + 8)))
        return rc;
    if (n) HIP_TRY(hipMemcpyAsync(m->cl.p, cl, 24 * n, hipMemcpyHostToDevice, st));
    rc = ocl_hip_drop_device(m, desc, m->cl.as<double>(), n, precision, m->out_z.as<double>(),
                             cc ? m->out_cc.as<double>() : nullptr, type ? m->out_ty.as<uint8_t>() : nullptr,
                             tri ? m->out_tri.as<uint32_t>() : nullptr, feat ? m->out_ft.as<uint8_t>() : nullptr,
                             stats, stream_);
    if (rc) return rc;
    if (n) {
        HIP_TRY(hipMemcpyAsync(z, m->out_z.p, 8 * n, hipMemcpyDeviceToHost, st));
        if (cc) HIP_TRY(hipMemcpyAsync(cc, m->out_cc.p, 24 * n, hipMemcpyDeviceToHost, st));
        if (type) HIP_TRY(hipMemcpyAsync(type, m->out_ty.p, n, hipMemcpyDeviceToHost, st));
        if (tri) HIP_TRY(hipMemcpyAsync(tri, m->out_tri.p, 4 * n, hipMemcpyDeviceToHost, st));
        if (feat) HIP_TRY(hipMemcpyAsync(feat, m->out_ft.p, n, hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    return OCL_OK;
}

}  // extern "C"

// ============================================================= backward
namespace ocl_amd {

// Sort keys of the backward's per-vertex sums: the point's contact triangle
// (input index), or ntri when the point has no contact.
__global__ void __launch_bounds__(256) k_grad_keys(const uint8_t* __restrict__ type, const uint32_t* __restrict__ tri,
                                                   uint32_t n, uint32_t ntri, uint32_t* __restrict__ gkey,
                                                   uint32_t* __restrict__ gidx) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t t = tri[i];
    gkey[i] = (t != 0xffffffffu && t < ntri && type[i] != CC_NONE) ? t : ntri;
    gidx[i] = i;
}

constexpr uint32_t kSegBlock = 256;  // positions per block of the per-triangle sums

// One lane per CL point: derivatives of its dropped z from the recorded
// active contact (drop_grad.hpp).
// Per-vertex terms (gtri non-NULL): the lanes take the points in the order
// of a stable sort by contact triangle (skey, sidx: ascending triangle,
// equal triangles in point order), put their 9 weighted terms in LDS, and
// the first lane of each run of equal triangles in the block adds the run's
// terms in order.  A run that is a whole segment (it starts the triangle's
// segment and the segment ends inside the block) goes straight to gtri; a
// run that continues past the block end, or started before the block,
// leaves its sum in seg_part (slot 2b for the run at the block's first
// position, slot 2b+1 for the block's last run) and k_grad_carry adds the
// pieces of each such segment in block order.  Every sum has one fixed
// order (FP64 atomics made the last bits vary run to run), and the terms
// never leave the chip.
// dz/dr: each block writes its partial sum to gr_part[block]; k_sum_into
// adds them up (one same-address atomic per wavefront serialised ~15 k
// atomics per million points).
template <class T>
__global__ void __launch_bounds__(256) k_drop_grad(MeshView m, const uint32_t* __restrict__ rank_of, CutterDev cd,
                                                   const double* __restrict__ cl, const double* __restrict__ cc,
                                                   const uint8_t* __restrict__ type, const uint32_t* __restrict__ tri,
                                                   const uint8_t* __restrict__ feat, const double* __restrict__ gz,
                                                   uint32_t n, double* __restrict__ dz,
                                                   const uint32_t* __restrict__ skey, const uint32_t* __restrict__ sidx,
                                                   double* __restrict__ gtri, double* __restrict__ seg_part,
                                                   double* __restrict__ gr_part) {
    __shared__ double wsum[4];
    __shared__ double gsh[9 * kSegBlock];
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t i = (p < n) ? (sidx ? sidx[p] : p) : 0u;
    double acc_r = 0.0;
    GradOut o{};
    double g = 0.0;
    if (p < n) {
        const uint32_t t = tri[i];
        const int ty = type[i];
        if (t != 0xffffffffu && t < m.ntri && ty != CC_NONE) {
            const double* r = m.tri + (size_t)kRec * rank_of[t];
            const V3 V[3] = {v3(r[0], r[1], r[2]), v3(r[3], r[4], r[5]), v3(r[6], r[7], r[8])};
            const double x = cl[3 * (size_t)i], y = cl[3 * (size_t)i + 1];
            const V3 P = v3(cc[3 * (size_t)i], cc[3 * (size_t)i + 1], cc[3 * (size_t)i + 2]);
            const int f = feat[i];
            if (cd.c.kind == K_COMPOSITE) {
                const int sn = grad_sub_index(cd, f, sqrt(sqr(P.x - x) + sqr(P.y - y)));
                // constant indices only: a dynamic one would copy the cutter to scratch
                if (sn == 0) o = contact_grad<T>(cd.sub[0], V, f, ty, x, y, P);
                else if (sn == 1) o = contact_grad<T>(cd.sub[1], V, f, ty, x, y, P);
                else if (sn == 2) o = contact_grad<T>(cd.sub[2], V, f, ty, x, y, P);
                else o = contact_grad<T>(cd.sub[3], V, f, ty, x, y, P);
                o.gr = NAN;
            } else {
                o = contact_grad<T>(cd.c, V, f, ty, x, y, P);
            }
            g = gz ? gz[i] : 1.0;
            acc_r = g * o.gr;
        }
        if (dz) {
            dz[3 * (size_t)i] = o.gx;
            dz[3 * (size_t)i + 1] = o.gy;
            dz[3 * (size_t)i + 2] = o.gr;
        }
    }
    if (gtri) {
#pragma unroll
        for (int j = 0; j < 9; ++j) gsh[9 * threadIdx.x + j] = g * o.gv[j];
        __syncthreads();
        const uint32_t b0 = blockIdx.x * kSegBlock, bend = min(n, b0 + kSegBlock);
        if (p < n) {
            const uint32_t k = skey[p];
            const bool first = threadIdx.x == 0;
            if (k < m.ntri && (first || skey[p - 1] != k)) {
                double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                uint32_t q = p;
                for (; q < bend && skey[q] == k; ++q)
#pragma unroll
                    for (int j = 0; j < 9; ++j) a[j] += gsh[9 * (q - b0) + j];
                const bool starts = p == 0 || skey[p - 1] != k;
                const bool ends = q == n || skey[q] != k;
                if (starts && ends) {
#pragma unroll
                    for (int j = 0; j < 9; ++j) gtri[9 * (size_t)k + j] += a[j];
                } else {
                    double* d = seg_part + 9 * (2 * (size_t)blockIdx.x + (first ? 0 : 1));
#pragma unroll
                    for (int j = 0; j < 9; ++j) d[j] = a[j];
                }
            }
        }
    }
    if (gr_part) {
        for (int off = 32; off > 0; off >>= 1) acc_r += __shfl_xor(acc_r, off, 64);
        if (lane_id() == 0) wsum[threadIdx.x >> 6] = acc_r;
        __syncthreads();
        if (threadIdx.x == 0) gr_part[blockIdx.x] = ((wsum[0] + wsum[1]) + wsum[2]) + wsum[3];
    }
}

// One thread per block b: the segments that start in block b and cross its
// end (the run at the block's first position when it starts its segment, or
// the block's last run) collect their pieces from the following blocks'
// first-position slots, in order.
__global__ void __launch_bounds__(256) k_grad_carry(const uint32_t* __restrict__ key, uint32_t n, uint32_t ntri,
                                                    uint32_t nblk, double* __restrict__ gtri,
                                                    const double* __restrict__ part) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblk) return;
    const uint32_t bend = min(n, (b + 1) * kSegBlock);
    if (bend >= n) return;  // nothing crosses the last block's end
    const uint32_t k = key[bend - 1];
    if (k >= ntri || key[bend] != k) return;  // the block's last run does not cross
    // the crossing run's piece: slot 2b if it runs from the block's first
    // position (then it must also start its segment there), else 2b+1
    const uint32_t b0 = b * kSegBlock;
    const bool from_first = key[b0] == k;
    if (from_first && b0 > 0 && key[b0 - 1] == k) return;  // owned by an earlier block
    double a[9];
    const double* s0 = part + 9 * (2 * (size_t)b + (from_first ? 0 : 1));
#pragma unroll
    for (int j = 0; j < 9; ++j) a[j] = s0[j];
    for (uint32_t c = b + 1; c < nblk && key[c * kSegBlock] == k; ++c) {
        const double* s = part + 9 * (2 * (size_t)c);
#pragma unroll
        for (int j = 0; j < 9; ++j) a[j] += s[j];
    }
#pragma unroll
    for (int j = 0; j < 9; ++j) gtri[9 * (size_t)k + j] += a[j];
}

// *out += the sum of part[0..n), one 256-thread block, fixed order.
__global__ void __launch_bounds__(256) k_sum_into(const double* __restrict__ part, uint32_t n, double* out) {
    __shared__ double s[256];
    double a = 0.0;
    for (uint32_t j = threadIdx.x; j < n; j += 256) a += part[j];
    s[threadIdx.x] = a;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out += s[0];
}

}  // namespace ocl_amd

extern "C" int ocl_hip_drop_backward(ocl_mesh* m, const ocl_cutter_desc* desc, const double* d_cl, size_t n,
                                     const double* d_cc, const uint8_t* d_type, const uint32_t* d_tri,
                                     const uint8_t* d_feat, const double* d_grad_z, int precision, double* d_dz,
                                     double* d_grad_tri, double* d_grad_r, void* stream_) {
    using namespace ocl_amd;
    if (!m) return fail(OCL_E_ARG, "mesh is NULL");
    if (precision < OCL_PREC_EXACT64 || precision > OCL_PREC_MIXED) return fail(OCL_E_ARG, "unknown precision mode");
    if (n >= 0x7fffffffu) return fail(OCL_E_ARG, "too many points");
    if (n == 0) return OCL_OK;
    if (!d_cl || !d_cc || !d_type || !d_tri || !d_feat) return fail(OCL_E_ARG, "forward-pass arrays are NULL");
    CutterDev cd;
    int rc = cutter_from_desc(desc, cd);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream_;
    const uint32_t n32 = (uint32_t)n;
    const MeshView mv = m->view();
    const uint32_t* rk = m->rank_of.as<const uint32_t>();
    const uint32_t nb = (n32 + 255) / 256;
    if (m->ntri == 0) {
        // no contact anywhere: zero derivatives
        if (d_dz) HIP_TRY(hipMemsetAsync(d_dz, 0, 24 * n, st));
        return OCL_OK;
    }
    double* part = nullptr;
    if (d_grad_r) {
        if ((rc = m->gr_part.ensure(8ull * nb + 64))) return rc;
        part = m->gr_part.as<double>();
    }
    const uint32_t nseg = (n32 + kSegBlock - 1) / kSegBlock;
    const uint32_t *ks = nullptr, *is = nullptr;
    double* gp = nullptr;
    if (d_grad_tri) {
        // stable sort of the points by contact triangle (keys up to ntri);
        // the gradient kernel then takes them in that order
        if ((rc = m->gt_kv.ensure(8ull * n + 64)) || (rc = m->gt_hist.ensure(4ull * rs_scratch_words(n32))) ||
            (rc = m->gt_part.ensure(144ull * nseg + 64)))
            return rc;
        uint32_t* gkey = m->gt_kv.as<uint32_t>();
        uint32_t* gidx = gkey + n32;
        const uint32_t nt32 = (uint32_t)m->ntri;
        k_grad_keys<<<nb, 256, 0, st>>>(d_type, d_tri, n32, nt32, gkey, gidx);
        HIP_TRY(hipGetLastError());
        uint32_t bits = 1;
        while (bits < 32 && (nt32 >> bits) != 0) ++bits;
        uint32_t *ks_, *is_;
        rs_sort(gkey, gidx, n32, bits, m->gt_hist.as<uint32_t>(), &ks_, &is_, st);
        ks = ks_;
        is = is_;
        gp = m->gt_part.as<double>();
    }
    static_assert(kSegBlock == 256, "one gradient block per sum block");
    if (precision == OCL_PREC_FP32 || precision == OCL_PREC_MIXED)
        k_drop_grad<float><<<nb, 256, 0, st>>>(mv, rk, cd, d_cl, d_cc, d_type, d_tri, d_feat, d_grad_z, n32, d_dz,
                                               ks, is, d_grad_tri, gp, part);
    else
        k_drop_grad<double><<<nb, 256, 0, st>>>(mv, rk, cd, d_cl, d_cc, d_type, d_tri, d_feat, d_grad_z, n32, d_dz,
                                                ks, is, d_grad_tri, gp, part);
    HIP_TRY(hipGetLastError());
    if (d_grad_tri) {
        k_grad_carry<<<(nseg + 255) / 256, 256, 0, st>>>(ks, n32, (uint32_t)m->ntri, nseg, d_grad_tri, gp);
        HIP_TRY(hipGetLastError());
    }
    if (d_grad_r) {
        k_sum_into<<<1, 256, 0, st>>>(part, nb, d_grad_r);
        HIP_TRY(hipGetLastError());
    }
    return OCL_OK;
}

extern "C" int ocl_hip_drop_slopes(ocl_mesh* m, const ocl_cutter_desc* desc, const double* cl, size_t n,
                                   int precision, double* z, double* dz_dxy, double* cc, uint8_t* type,
                                   ocl_stats* stats, void* stream_) {
    using namespace ocl_amd;
    if (!m) return fail(OCL_E_ARG, "mesh is NULL");
    if (n > 0 && (!cl || !z || !dz_dxy)) return fail(OCL_E_ARG, "cl/z/dz_dxy pointers are NULL");
    hipStream_t st = (hipStream_t)stream_;
    int rc;
    if ((rc = m->cl.ensure(24ull * n + 8)) || (rc = m->out_z.ensure(8ull * n + 8)) ||
        (rc = m->out_cc.ensure(24ull * n + 8)) || (rc = m->out_ty.ensure(n + 8)) ||
        (rc = m->out_tri.ensure(4ull * n + 8)) || (rc = m->out_ft.ensure(n + 8)) ||
        (rc = m->out_dz.ensure(24ull * n + 8)))
        return rc;
    if (n == 0) {
        if (stats) std::memset(stats, 0, sizeof(*stats));
        return OCL_OK;
    }
    HIP_TRY(hipMemcpyAsync(m->cl.p, cl, 24 * n, hipMemcpyHostToDevice, st));
    rc = ocl_hip_drop_device(m, desc, m->cl.as<double>(), n, precision, m->out_z.as<double>(), m->out_cc.as<double>(),
                             m->out_ty.as<uint8_t>(), m->out_tri.as<uint32_t>(), m->out_ft.as<uint8_t>(), stats, stream_);
    if (rc) return rc;
    rc = ocl_hip_drop_backward(m, desc, m->cl.as<double>(), n, m->out_cc.as<double>(), m->out_ty.as<uint8_t>(),
                               m->out_tri.as<uint32_t>(), m->out_ft.as<uint8_t>(), nullptr, precision,
                               m->out_dz.as<double>(), nullptr, nullptr, stream_);
    if (rc) return rc;
    std::vector<double> d3(3 * n);
    HIP_TRY(hipMemcpyAsync(z, m->out_z.p, 8 * n, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(d3.data(), m->out_dz.p, 24 * n, hipMemcpyDeviceToHost, st));
    if (cc) HIP_TRY(hipMemcpyAsync(cc, m->out_cc.p, 24 * n, hipMemcpyDeviceToHost, st));
    if (type) HIP_TRY(hipMemcpyAsync(type, m->out_ty.p, n, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (size_t i = 0; i < n; ++i) {
        dz_dxy[2 * i] = d3[3 * i];
        dz_dxy[2 * i + 1] = d3[3 * i + 1];
    }
    return OCL_OK;
}

// ============================================================= batches
// Many drop-cutter jobs (each a mesh, a cutter and a CL array) in one
// submission: one key pass and one deterministic sort over every job's
// points, the selection passes of all fast jobs in one persistent launch fed
// by a device work queue, then each job's resolve and exact passes.  Nothing
// in the sequence synchronises with the host or records a memset, so it is
// captured once in a hipGraph and replayed.

namespace ocl_amd {

struct KeyJob {
    const double* cl;
    uint32_t n, base;
    BinFrame f;
};

// keys = (job << 2 hb) | Hilbert cell, vals = index inside the job
__global__ void k_cl_keys_jobs(const KeyJob* __restrict__ kj, uint32_t njobs, uint32_t total, uint32_t hb,
                               uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (kj[mid].base <= g) lo = mid;
        else hi = mid;
    }
    const KeyJob& J = kj[lo];
    const uint32_t i = g - J.base;
    const double cmax = (double)((1u << hb) - 1u);
    const double fx = (J.cl[3 * (size_t)i] - J.f.x0) * J.f.sx, fy = (J.cl[3 * (size_t)i + 1] - J.f.y0) * J.f.sy;
    const uint32_t ix = (uint32_t)fmin(fmax(fx, 0.0), cmax), iy = (uint32_t)fmin(fmax(fy, 0.0), cmax);
    keys[g] = (lo << (2 * hb)) | hilbert_key(ix, iy, hb);
    vals[g] = i;
}

// zeroes the device ranges of a table (one block per range)
struct ZeroRange {
    uint32_t* p;
    uint32_t words;
};
__global__ void k_zero_ranges(const ZeroRange* __restrict__ r) {
    const ZeroRange z = r[blockIdx.x];
    for (uint32_t i = threadIdx.x; i < z.words; i += blockDim.x) z.p[i] = 0u;
}

struct BatchJob {
    ocl_mesh* m;
    std::unique_ptr<DropWork> w;  // the job's own workspace and counters
    CutterDev cd;
    const double* cl;
    uint32_t n, base;
    DropOut o;
    bool fast;
    FastPlan F;                   // holds the job's selection records
    MeshView mv;
};

}  // namespace ocl_amd

struct ocl_batch {
    std::vector<ocl_amd::BatchJob> jobs;
    int precision = OCL_PREC_MIXED;
    ocl_tune tune{};
    uint32_t total = 0, hb = 8, key_bits = 16;
    ocl_amd::DevBuf keys, vals, scratch, keyjobs, zeros, seljobs, chunk_off, queue, exjobs, exoff, postjobs, postoff;
    uint32_t nzero = 0, nsel = 0, grid = 0, nexact = 0, exchunks = 0, npost = 0, postchunks = 0;
    uint32_t* sorted_vals = nullptr;
    std::vector<uint32_t> sel_order;  // job indices of the queue, largest first
    std::vector<uint32_t> comp_sel;   // composite selection jobs (launched one by one)
    hipStream_t cap = nullptr;
    hipStream_t side = nullptr;  // the composite jobs' passes, beside the others
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr, s0 = nullptr, s1 = nullptr, fork = nullptr, join = nullptr;
    ~ocl_batch() {
        for (const ocl_amd::BatchJob& J : jobs) --J.m->batches;
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (cap) (void)hipStreamDestroy(cap);
        if (side) (void)hipStreamDestroy(side);
        for (hipEvent_t e : {e0, e1, s0, s1, fork, join})
            if (e) (void)hipEventDestroy(e);
    }
};

namespace ocl_amd {

// Issues the whole batch on st: no host synchronisation, no memset.
static int batch_record(ocl_batch* b, hipStream_t st, bool timed) {
    const bool f32 = b->precision == OCL_PREC_FP32 || b->precision == OCL_PREC_MIXED;
    if (timed) HIP_TRY(hipEventRecord(b->e0, st));
    k_zero_ranges<<<b->nzero, 256, 0, st>>>(b->zeros.as<ZeroRange>());
    k_cl_keys_jobs<<<(b->total + 255) / 256, 256, 0, st>>>(b->keyjobs.as<KeyJob>(), (uint32_t)b->jobs.size(),
                                                           b->total, b->hb, b->keys.as<uint32_t>(),
                                                           b->vals.as<uint32_t>());
    uint32_t *ko = nullptr, *vo = nullptr;
    rs_sort(b->keys.as<uint32_t>(), b->vals.as<uint32_t>(), b->total, b->key_bits, b->scratch.as<uint32_t>(), &ko,
            &vo, st);
    if (vo != b->sorted_vals) return fail(OCL_E_HIP, "batch: sort buffer mismatch");
    HIP_TRY(hipGetLastError());
    for (const BatchJob& J : b->jobs) {
        if (J.fast || J.n == 0 || J.m->ntri > 0) continue;
        k_drop_none<<<(J.n + 255) / 256, 256, 0, st>>>(J.cl, J.n, J.o);
    }
    if (b->nexact)
        k_drop_exact_jobs<<<b->exchunks, 64, 0, st>>>(b->exjobs.as<ExactJob>(), b->exoff.as<uint32_t>(),
                                                                 b->nexact);
    // Composite jobs: their selection and later passes on the side stream,
    // where they overlap the other jobs' (a composite job with few points
    // on a large mesh is a long chain of few wavefronts).
    const bool forked = !b->comp_sel.empty();
    if (forked) {
        HIP_TRY(hipEventRecord(b->fork, st));
        HIP_TRY(hipStreamWaitEvent(b->side, b->fork, 0));
        for (uint32_t j : b->comp_sel) {
            const BatchJob& J = b->jobs[j];
            CompSelArgs C;
            C.A = fast_args32(J.F);
            C.PS[0] = J.F.PS32[0];
            C.PS[1] = J.F.PS32[1];
            k_drop_select_comp<<<select_grid(J.n, J.mv.xcd_group), 64, 0, b->side>>>(C);
            HIP_TRY(hipGetLastError());
        }
        for (uint32_t j : b->comp_sel) {
            int rc = fast_post(b->jobs[j].F, b->side);
            if (rc) return rc;
        }
        HIP_TRY(hipEventRecord(b->join, b->side));
    }
    if (timed) HIP_TRY(hipEventRecord(b->s0, st));
    if (b->nsel) {
        if (b->tune.per_job_launch) {
            for (uint32_t j : b->sel_order) {
                const BatchJob& J = b->jobs[j];
                const uint32_t nb = select_grid(J.n, J.mv.xcd_group);
                if (f32) {
                    const SelArgs<float> A = fast_args32(J.F);
                    switch (J.cd.c.kind) {
                    case K_CYL: k_drop_select<float, K_CYL><<<nb, 64, 0, st>>>(A); break;
                    case K_BALL: k_drop_select<float, K_BALL><<<nb, 64, 0, st>>>(A); break;
                    case K_BULL: k_drop_select<float, K_BULL><<<nb, 64, 0, st>>>(A); break;
                    default: k_drop_select<float, K_CONE><<<nb, 64, 0, st>>>(A); break;
                    }
                } else {
                    const SelArgs<double> A = fast_args64(J.F);
                    switch (J.cd.c.kind) {
                    case K_CYL: k_drop_select<double, K_CYL><<<nb, 64, 0, st>>>(A); break;
                    case K_BALL: k_drop_select<double, K_BALL><<<nb, 64, 0, st>>>(A); break;
                    case K_BULL: k_drop_select<double, K_BULL><<<nb, 64, 0, st>>>(A); break;
                    default: k_drop_select<double, K_CONE><<<nb, 64, 0, st>>>(A); break;
                    }
                }
            }
        } else if (f32) {
            k_drop_select_jobs<float><<<b->grid, 256, 0, st>>>(b->seljobs.as<SelArgs<float>>(),
                                                               b->chunk_off.as<uint32_t>(), b->nsel,
                                                               b->queue.as<uint32_t>());
        } else {
            k_drop_select_jobs<double><<<b->grid, 256, 0, st>>>(b->seljobs.as<SelArgs<double>>(),
                                                                b->chunk_off.as<uint32_t>(), b->nsel,
                                                                b->queue.as<uint32_t>());
        }
        HIP_TRY(hipGetLastError());
    }
    if (timed) HIP_TRY(hipEventRecord(b->s1, st));
    if (b->npost) {
        // the passes after the selection, every job in one launch each
        k_drop_resolve_jobs<<<(b->postchunks + 3) / 4, 256, 0, st>>>(b->postjobs.as<PostJob>(),
                                                                      b->postoff.as<uint32_t>(), b->npost);
        k_drop_neartie_jobs<<<kNearTieBlocks, 256, 0, st>>>(b->postjobs.as<PostJob>(), b->npost);
        k_drop_coop_exact_jobs<<<kCoopBlocks, 256, 0, st>>>(b->postjobs.as<PostJob>(), b->npost);
        HIP_TRY(hipGetLastError());
    } else {
        for (const BatchJob& J : b->jobs) {
            if (!J.fast || J.n == 0 || J.F.comp) continue;
            int rc = fast_post(J.F, st);
            if (rc) return rc;
        }
    }
    if (forked) HIP_TRY(hipStreamWaitEvent(st, b->join, 0));
    if (timed) HIP_TRY(hipEventRecord(b->e1, st));
    return OCL_OK;
}

}  // namespace ocl_amd

extern "C" {

int ocl_hip_batch_destroy(ocl_batch* b) {
    delete b;
    return OCL_OK;
}

int ocl_hip_batch_create(const ocl_drop_job* jobs, size_t njobs, int precision, const ocl_tune* tune,
                         ocl_batch** out) {
    using namespace ocl_amd;
    if (!out) return fail(OCL_E_ARG, "out is NULL");
    *out = nullptr;
    if (njobs == 0 || !jobs) return fail(OCL_E_ARG, "a batch needs at least one job");
    if (njobs > 4096) return fail(OCL_E_ARG, "at most 4096 jobs per batch");
    if (precision < OCL_PREC_EXACT64 || precision > OCL_PREC_MIXED) return fail(OCL_E_ARG, "unknown precision mode");
    std::unique_ptr<ocl_batch> b(new (std::nothrow) ocl_batch());
    if (!b) return fail(OCL_E_NOMEM, "out of host memory");
    b->precision = precision;
    if (tune) b->tune = *tune;
    b->hb = b->tune.hilbert_bits ? (uint32_t)b->tune.hilbert_bits : 9u;  // the single-call default
    if (b->hb < 4 || b->hb > 10) return fail(OCL_E_ARG, "hilbert_bits must be in 4..10");
    const uint32_t bpc = b->tune.blocks_per_cu > 0 ? (uint32_t)b->tune.blocks_per_cu : 4u;
    uint32_t jbits = 0;
    while ((1ull << jbits) < njobs) ++jbits;
    b->key_bits = 2 * b->hb + jbits;
    if (b->key_bits > 32) return fail(OCL_E_ARG, "too many jobs for this hilbert_bits");
    std::vector<KeyJob> kj(njobs);
    std::vector<ZeroRange> zr;
    uint64_t total = 0;
    hipStream_t st = nullptr;
    for (size_t j = 0; j < njobs; ++j) {
        const ocl_drop_job& q = jobs[j];
        if (!q.mesh) return fail(OCL_E_ARG, "job " + std::to_string(j) + ": mesh is NULL");
        if (q.n > 0 && (!q.d_cl_xyz || !q.d_z_out)) return fail(OCL_E_ARG, "job " + std::to_string(j) + ": cl/z are NULL");
        BatchJob J;
        J.m = q.mesh;
        J.w.reset(new (std::nothrow) DropWork());
        if (!J.w) return fail(OCL_E_NOMEM, "out of host memory");
        int rc = cutter_from_desc(&q.cutter, J.cd);
        if (rc) return rc;
        J.cl = q.d_cl_xyz;
        J.n = (uint32_t)q.n;
        J.base = (uint32_t)total;
        J.o = DropOut{q.d_z_out, q.d_cc_xyz_out, q.d_cc_type_out, q.d_tri_out, q.d_feat_out};
        J.mv = q.mesh->view();
        J.fast = selects(J.cd, precision, q.mesh->ntri);
        ocl_mesh* m = q.mesh;
        DropWork& w = *J.w;
        if ((rc = w.ensure_counters())) return rc;
        zr.push_back({w.counters.as<uint32_t>(), (uint32_t)(sizeof(Counters) / 4)});
        zr.push_back({w.cnt2.as<uint32_t>(), (uint32_t)(sizeof(Counters) / 4)});
        zr.push_back({w.cnt3.as<uint32_t>(), (uint32_t)(sizeof(Counters) / 4)});
        zr.push_back({w.qcount.as<uint32_t>(), 4u});
        const double R = J.cd.c.radius;
        KeyJob& k = kj[j];
        k.cl = q.d_cl_xyz;
        k.n = J.n;
        k.base = J.base;
        k.f.x0 = m->bbox[0] - R;
        k.f.y0 = m->bbox[2] - R;
        const double side = (double)(1u << b->hb);
        k.f.sx = side / std::max(m->bbox[1] + R - k.f.x0, 1e-9);
        k.f.sy = side / std::max(m->bbox[3] + R - k.f.y0, 1e-9);
        total += q.n;
        ++m->batches;  // released by ~ocl_batch
        b->jobs.push_back(std::move(J));
    }
    if (total >= 0x7fffffffu) return fail(OCL_E_ARG, "too many CL points in one batch");
    b->total = (uint32_t)total;
    int rc;
    const uint32_t tot = std::max<uint32_t>(b->total, 1u);
    if ((rc = b->keys.ensure(4ull * tot)) || (rc = b->vals.ensure(4ull * tot)) ||
        (rc = b->scratch.ensure(4ull * rs_scratch_words(tot))) || (rc = b->queue.ensure(64)))
        return rc;
    zr.push_back({b->queue.as<uint32_t>(), 1u});
    // where the sort leaves the values: ping-pong parity of its passes
    {
        const uint32_t passes = (b->key_bits + 7) / 8;
        const uint32_t ntiles = (tot + kRsTile - 1) / kRsTile;
        uint32_t* v2 = b->scratch.as<uint32_t>() + (size_t)256 * ntiles + tot;
        b->sorted_vals = (passes % 2) ? v2 : b->vals.as<uint32_t>();
    }
    // per-job plans (records are built and cached here, synchronously)
    std::vector<uint32_t> sel;
    for (BatchJob& J : b->jobs) {
        if (!J.fast || J.n == 0) continue;
        if ((rc = fast_prepare(J.m, *J.w, J.mv, J.cd, J.cl, b->sorted_vals + J.base, J.n, precision, J.o,
                               J.w->counters.as<Counters>(), st, J.F)))
            return rc;
        // composite selections have a kernel of their own, outside the queue
        (J.F.comp ? b->comp_sel : sel).push_back((uint32_t)(&J - b->jobs.data()));
    }
    std::stable_sort(sel.begin(), sel.end(), [&](uint32_t a, uint32_t c) { return b->jobs[a].n > b->jobs[c].n; });
    b->sel_order = sel;
    b->nsel = (uint32_t)sel.size();
    std::vector<uint32_t> coff(sel.size() + 1, 0);
    for (size_t i = 0; i < sel.size(); ++i) coff[i + 1] = coff[i] + (b->jobs[sel[i]].n + 63) / 64;
    const bool f32 = precision == OCL_PREC_FP32 || precision == OCL_PREC_MIXED;
    const size_t argsz = f32 ? sizeof(SelArgs<float>) : sizeof(SelArgs<double>);
    if ((rc = b->seljobs.ensure(argsz * (sel.size() + 1))) || (rc = b->chunk_off.ensure(4 * coff.size())) ||
        (rc = b->keyjobs.ensure(sizeof(KeyJob) * njobs)) || (rc = b->zeros.ensure(sizeof(ZeroRange) * zr.size())))
        return rc;
    if (f32) {
        std::vector<SelArgs<float>> a;
        for (uint32_t j : sel) a.push_back(fast_args32(b->jobs[j].F));
        if (!a.empty()) HIP_TRY(hipMemcpy(b->seljobs.p, a.data(), argsz * a.size(), hipMemcpyHostToDevice));
    } else {
        std::vector<SelArgs<double>> a;
        for (uint32_t j : sel) a.push_back(fast_args64(b->jobs[j].F));
        if (!a.empty()) HIP_TRY(hipMemcpy(b->seljobs.p, a.data(), argsz * a.size(), hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMemcpy(b->chunk_off.p, coff.data(), 4 * coff.size(), hipMemcpyHostToDevice));
    if (precision != OCL_PREC_FP32 && env_scale("OCL_COOP_SELECT", 0.0) == 0.0) {
        // the fast jobs' passes after the selection, in queue order
        std::vector<PostJob> pj;
        std::vector<uint32_t> poff(1, 0);
        for (uint32_t j : sel) {
            BatchJob& J = b->jobs[j];
            const FastPlan& F = J.F;
            ocl_mesh* m = F.m;
            pj.push_back(PostJob{F.mv, F.cd, F.cl, F.perm, F.n, F.nts, F.etie, F.o, F.q1, F.qcount, F.zfloor,
                                 F.r64->cap.as<double>(), m->origins.as<BoxOrigin>(), F.cnt2, F.cnt3,
                                 F.r64->capup.p ? F.r64->capup.as<const float>() : nullptr});
            poff.push_back(poff.back() + (F.n + 63) / 64);
            J.w->exact_slot = 0;
        }
        b->npost = (uint32_t)pj.size();
        b->postchunks = poff.back();
        if ((rc = b->postjobs.ensure(sizeof(PostJob) * (pj.size() + 1))) || (rc = b->postoff.ensure(4 * poff.size())))
            return rc;
        if (!pj.empty()) HIP_TRY(hipMemcpy(b->postjobs.p, pj.data(), sizeof(PostJob) * pj.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(b->postoff.p, poff.data(), 4 * poff.size(), hipMemcpyHostToDevice));
    }
    {
        // exact-path jobs, largest first
        std::vector<uint32_t> ex;
        for (const BatchJob& J : b->jobs)
            if (!J.fast && J.n > 0 && J.m->ntri > 0) ex.push_back((uint32_t)(&J - b->jobs.data()));
        std::stable_sort(ex.begin(), ex.end(), [&](uint32_t a, uint32_t c) { return b->jobs[a].n > b->jobs[c].n; });
        std::vector<ExactJob> ej;
        std::vector<uint32_t> eoff(1, 0);
        for (uint32_t j : ex) {
            const BatchJob& J = b->jobs[j];
            ej.push_back(ExactJob{J.mv, J.cd, J.cl, b->sorted_vals + J.base, J.n, J.o, J.w->counters.as<Counters>()});
            eoff.push_back(eoff.back() + (J.n + 63) / 64);
        }
        b->nexact = (uint32_t)ej.size();
        b->exchunks = eoff.back();
        if ((rc = b->exjobs.ensure(sizeof(ExactJob) * (ej.size() + 1))) || (rc = b->exoff.ensure(4 * eoff.size())))
            return rc;
        if (!ej.empty()) HIP_TRY(hipMemcpy(b->exjobs.p, ej.data(), sizeof(ExactJob) * ej.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(b->exoff.p, eoff.data(), 4 * eoff.size(), hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMemcpy(b->keyjobs.p, kj.data(), sizeof(KeyJob) * njobs, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b->zeros.p, zr.data(), sizeof(ZeroRange) * zr.size(), hipMemcpyHostToDevice));
    b->nzero = (uint32_t)zr.size();
    int dev = 0, cus = 256;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    b->grid = std::max(1u, std::min((uint32_t)cus * bpc, coff.back()));
    for (hipEvent_t* e : {&b->e0, &b->e1, &b->s0, &b->s1}) HIP_TRY(hipEventCreate(e));
    for (hipEvent_t* e : {&b->fork, &b->join}) HIP_TRY(hipEventCreateWithFlags(e, hipEventDisableTiming));
    if (!b->comp_sel.empty()) HIP_TRY(hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking));
    HIP_TRY(hipDeviceSynchronize());
    if (b->tune.use_graph && b->total > 0) {
        HIP_TRY(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
        HIP_TRY(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
        rc = batch_record(b.get(), b->cap, false);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(b->cap, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ec != hipSuccess) return fail(OCL_E_HIP, std::string("batch capture: ") + hipGetErrorString(ec));
        b->graph = g;
        HIP_TRY(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    }
    *out = b.release();
    return OCL_OK;
}

int ocl_hip_batch_run(ocl_batch* b, ocl_stats* stats, void* stream_) {
    using namespace ocl_amd;
    if (!b) return fail(OCL_E_ARG, "batch is NULL");
    hipStream_t st = (hipStream_t)stream_;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (b->total == 0) return OCL_OK;
    if (b->exec) {
        if (stats) HIP_TRY(hipEventRecord(b->e0, st));
        HIP_TRY(hipGraphLaunch(b->exec, st));
        if (stats) HIP_TRY(hipEventRecord(b->e1, st));
    } else {
        int rc = batch_record(b, st, stats != nullptr);
        if (rc) return rc;
    }
    if (stats) {
        HIP_TRY(hipEventSynchronize(b->e1));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, b->e0, b->e1));
        stats->kernel_ms = ms;
        stats->select_ms = ms;
        if (!b->exec) {
            HIP_TRY(hipEventElapsedTime(&ms, b->s0, b->s1));
            stats->select_ms = ms;
        }
        stats->points = b->total;
        for (const BatchJob& J : b->jobs) {
            Counters hc;
            HIP_TRY(hipMemcpy(&hc, J.w->counters.p, sizeof(hc), hipMemcpyDeviceToHost));
            stats->pairs += hc.pairs;
            stats->escalated += hc.escalated;
            stats->wave_evals += hc.evals;
        }
    }
    return OCL_OK;
}

int ocl_hip_batch_job_stats(ocl_batch* b, size_t job, ocl_stats* stats) {
    using namespace ocl_amd;
    if (!b || !stats) return fail(OCL_E_ARG, "batch / stats is NULL");
    if (job >= b->jobs.size()) return fail(OCL_E_ARG, "job index out of range");
    std::memset(stats, 0, sizeof(*stats));
    const BatchJob& J = b->jobs[job];
    stats->points = J.n;
    if (J.n == 0) return OCL_OK;
    Counters hc;
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(&hc, J.w->counters.p, sizeof(hc), hipMemcpyDeviceToHost));
    stats->pairs = hc.pairs;
    stats->escalated = hc.escalated;
    stats->wave_evals = hc.evals;
    return OCL_OK;
}

int ocl_hip_batch_autotune(const ocl_drop_job* jobs, size_t njobs, int precision, int trials, ocl_tune* best,
                           double* best_ms) {
    using namespace ocl_amd;
    if (!best) return fail(OCL_E_ARG, "best is NULL");
    if (trials < 1) trials = 3;
    const int hbs[3] = {7, 8, 9}, bpcs[3] = {2, 4, 8};
    double bt = 1e300;
    ocl_tune pick{};
    for (int hb : hbs)
        for (int bpc : bpcs) {
            ocl_tune t{};
            t.hilbert_bits = hb;
            t.blocks_per_cu = bpc;
            t.use_graph = 1;
            ocl_batch* b = nullptr;
            int rc = ocl_hip_batch_create(jobs, njobs, precision, &t, &b);
            if (rc) return rc;
            double tmin = 1e300;
            for (int r = 0; r <= trials && rc == OCL_OK; ++r) {
                ocl_stats s;
                rc = ocl_hip_batch_run(b, &s, nullptr);
                if (r > 0) tmin = std::min(tmin, s.kernel_ms);  // the first run warms up
            }
            ocl_hip_batch_destroy(b);
            if (rc) return rc;
            if (tmin < bt) {
                bt = tmin;
                pick = t;
            }
        }
    *best = pick;
    if (best_ms) *best_ms = bt;
    return OCL_OK;
}

}  // extern "C"
