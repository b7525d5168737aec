// MI355X (gfx950) batch drop-cutter: mesh preparation, CL binning, drop
// kernels and the C ABI declared in include/ocl_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/ocl_hip.h"
#include "cutter.hpp"
#include "drop_exact.hpp"
#include "kd_rank.hpp"
#include "mesh.hpp"
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

}  // namespace ocl_amd

struct ocl_mesh {
    int device = 0;
    size_t ntri = 0;
    double bbox[6] = {0, 0, 0, 0, 0, 0};
    unsigned bucket = 1;
    ocl_amd::DevBuf tri;     // kRec doubles per triangle, rank order
    ocl_amd::DevBuf order;   // rank -> input index
    ocl_amd::DevBuf boxes;   // all hierarchy levels
    uint32_t cnt[ocl_amd::kMaxLevels] = {};
    size_t box_off[ocl_amd::kMaxLevels] = {};
    int top = 0;
    // per-call workspace
    ocl_amd::DevBuf keys, perm, hist, cursor, cl, out_z, out_cc, out_ty, out_tri, out_ft, counters, fastrec;
    ocl_amd::DevBuf esc_list;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    ocl_amd::MeshView view() const {
        ocl_amd::MeshView v;
        v.tri = tri.as<const double>();
        v.order = order.as<const uint32_t>();
        v.ntri = (uint32_t)ntri;
        for (int l = 0; l < ocl_amd::kMaxLevels; ++l) {
            v.tree.box[l] = boxes.as<const float4>() ? boxes.as<const float4>() + box_off[l] : nullptr;
            v.tree.cnt[l] = cnt[l];
        }
        v.tree.top = top;
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

// Level-0 boxes: XY box of kLeaf consecutive triangles, rounded outward.
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

// ------------------------------------------------------------ CL binning
// CL points are grouped so that each wavefront holds 64 points that are close
// in XY: 16-bit Morton key over the (mesh box grown by the cutter radius),
// counting sort by key.  Order inside a key is arbitrary; per-point results
// do not depend on it.
struct BinFrame {
    double x0, y0, sx, sy;  // key = morton(clamp((x - x0) * sx), clamp((y - y0) * sy))
};

__device__ inline uint32_t spread8(uint32_t v) {
    v &= 0xff;
    v = (v | (v << 4)) & 0x0f0f;
    v = (v | (v << 2)) & 0x3333;
    v = (v | (v << 1)) & 0x5555;
    return v;
}

__global__ void k_cl_keys(const double* __restrict__ cl, uint32_t n, BinFrame f, uint32_t* __restrict__ keys,
                          uint32_t* __restrict__ hist) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double fx = (cl[3 * (size_t)i] - f.x0) * f.sx, fy = (cl[3 * (size_t)i + 1] - f.y0) * f.sy;
    const uint32_t ix = (uint32_t)fmin(fmax(fx, 0.0), 255.0), iy = (uint32_t)fmin(fmax(fy, 0.0), 255.0);
    const uint32_t k = spread8(ix) | (spread8(iy) << 1);
    keys[i] = k;
    atomicAdd(&hist[k], 1u);
}

// exclusive scan of 65536 counters by one 1024-thread block
__global__ void __launch_bounds__(1024) k_scan64k(uint32_t* __restrict__ hist) {
    __shared__ uint32_t part[1024];
    const int t = threadIdx.x;
    uint32_t local[64];
    uint32_t s = 0;
    for (int k = 0; k < 64; ++k) {
        local[k] = s;
        s += hist[t * 64 + k];
    }
    part[t] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        uint32_t v = (t >= o) ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    const uint32_t base = part[t] - s;
    for (int k = 0; k < 64; ++k) hist[t * 64 + k] = base + local[k];
}

__global__ void k_cl_scatter(const uint32_t* __restrict__ keys, uint32_t n, uint32_t* __restrict__ cursor,
                             uint32_t* __restrict__ perm) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t slot = atomicAdd(&cursor[keys[i]], 1u);
    perm[slot] = i;
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
    unsigned long long pairs, escalated, brent, nesc;
};

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

// Exact FP64 drop over a list of points (identity list or escalation list).
// One lane per point; the wavefront walks the hierarchy with the union box.
__global__ void __launch_bounds__(256) k_drop_exact(MeshView m, CutterDev cd, const double* __restrict__ cl,
                                                    const uint32_t* __restrict__ list, uint32_t n, DropOut o,
                                                    Counters* cnt) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t wave_first = gid & ~63u;
    if (wave_first >= n) return;
    const bool act = gid < n;
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
    walk_boxes(m.tree, qx0, qy0, qx1, qy1, [&](uint32_t b) {
        const uint32_t e = min(m.ntri, (b + 1) * kLeaf);
        for (uint32_t k = b * kLeaf; k < e; ++k) {
            const uint32_t ku = uniform_u(k);
            const TriX t = load_tri(m.tri, ku);
            if (act) {
                L.cur_tri = ku;
                if (cutter_overlaps(R, L.x, L.y, t)) {
                    if (L.z < t.maxz) ++pairs;
                    brent += drop_triangle_exact(cd, L, t);
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

// Empty mesh: every CL point keeps its z and has no contact.
__global__ void k_drop_none(const double* __restrict__ cl, uint32_t n, DropOut o) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    write_drop(o, nullptr, i, fresh_lane(cl[3 * (size_t)i], cl[3 * (size_t)i + 1], cl[3 * (size_t)i + 2]));
}

}  // namespace ocl_amd

using namespace ocl_amd;

static int launch_drop_fast(ocl_mesh* m, const MeshView& mv, const CutterDev& cd, const double* d_cl, uint32_t n,
                            int precision, const DropOut& o, Counters* cnt, hipStream_t st);

// ------------------------------------------------------------- C ABI impl
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

extern "C" {

const char* ocl_hip_last_error(void) { return g_err.c_str(); }

size_t ocl_hip_mesh_size(const ocl_mesh* mesh) { return mesh ? mesh->ntri : 0; }

int ocl_hip_mesh_create(const double* tri_xyz, size_t ntri, unsigned bucket, ocl_mesh** out) {
    if (!out) return fail(OCL_E_ARG, "out is NULL");
    *out = nullptr;
    if (ntri > 0 && !tri_xyz) return fail(OCL_E_ARG, "tri_xyz is NULL");
    if (ntri >= 0x7fffffffu) return fail(OCL_E_ARG, "too many triangles");
    ocl_mesh* m = new (std::nothrow) ocl_mesh();
    if (!m) return fail(OCL_E_NOMEM, "out of host memory");
    HIP_TRY(hipGetDevice(&m->device));
    m->ntri = ntri;
    m->bucket = bucket < 1 ? 1 : bucket;
    // host: bbox per triangle and the reference's KD leaf order
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
    static const int kXY[4] = {0, 1, 2, 3};
    std::vector<uint32_t> order = kd_rank_order(bb.data(), ntri, kXY, m->bucket);
    std::vector<double> sorted(9 * ntri);
    for (size_t r = 0; r < ntri; ++r) std::memcpy(&sorted[9 * r], tri_xyz + 9 * (size_t)order[r], 9 * sizeof(double));
    int rc;
    if ((rc = m->tri.ensure(sizeof(double) * kRec * ntri)) || (rc = m->order.ensure(sizeof(uint32_t) * ntri))) {
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
    if ((rc = m->boxes.ensure(sizeof(float4) * total))) {
        delete m;
        return rc;
    }
    hipError_t e = hipMemcpy(staging.p, sorted.data(), sizeof(double) * 9 * ntri, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(m->order.p, order.data(), sizeof(uint32_t) * ntri, hipMemcpyHostToDevice);
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
    e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) {
        delete m;
        return fail(OCL_E_HIP, std::string("mesh prep: ") + hipGetErrorString(e));
    }
    *out = m;
    return OCL_OK;
}

int ocl_hip_mesh_destroy(ocl_mesh* mesh) {
    if (!mesh) return OCL_OK;
    if (mesh->ev0) (void)hipEventDestroy(mesh->ev0);
    if (mesh->ev1) (void)hipEventDestroy(mesh->ev1);
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
    if ((rc = m->counters.ensure(sizeof(Counters)))) return rc;
    Counters* cnt = m->counters.as<Counters>();
    HIP_TRY(hipMemsetAsync(cnt, 0, sizeof(Counters), st));
    // bin CL points
    if ((rc = m->keys.ensure(4ull * n)) || (rc = m->perm.ensure(4ull * n)) || (rc = m->hist.ensure(4ull * 65536)))
        return rc;
    const double R = cd.c.radius;
    BinFrame f;
    f.x0 = m->bbox[0] - R;
    f.y0 = m->bbox[2] - R;
    const double ex = std::max(m->bbox[1] + R - f.x0, 1e-9), ey = std::max(m->bbox[3] + R - f.y0, 1e-9);
    f.sx = 256.0 / ex;
    f.sy = 256.0 / ey;
    uint32_t* hist = m->hist.as<uint32_t>();
    HIP_TRY(hipMemsetAsync(hist, 0, 4 * 65536, st));
    const uint32_t nb = (n32 + 255) / 256;
    k_cl_keys<<<nb, 256, 0, st>>>(d_cl, n32, f, m->keys.as<uint32_t>(), hist);
    k_scan64k<<<1, 1024, 0, st>>>(hist);
    k_cl_scatter<<<nb, 256, 0, st>>>(m->keys.as<uint32_t>(), n32, hist, m->perm.as<uint32_t>());
    if (stats) {
        if (!m->ev0) HIP_TRY(hipEventCreate(&m->ev0));
        if (!m->ev1) HIP_TRY(hipEventCreate(&m->ev1));
        HIP_TRY(hipEventRecord(m->ev0, st));
    }
    const MeshView mv = m->view();
    if (precision == OCL_PREC_EXACT64 || m->ntri == 0) {
        if (m->ntri > 0)
            k_drop_exact<<<nb, 256, 0, st>>>(mv, cd, d_cl, m->perm.as<uint32_t>(), n32, o, cnt);
        else
            k_drop_none<<<nb, 256, 0, st>>>(d_cl, n32, o);
    } else {
        rc = launch_drop_fast(m, mv, cd, d_cl, n32, precision, o, cnt, st);
        if (rc) return rc;
    }
    HIP_TRY(hipGetLastError());
    if (stats) {
        HIP_TRY(hipEventRecord(m->ev1, st));
        HIP_TRY(hipEventSynchronize(m->ev1));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, m->ev0, m->ev1));
        Counters hc;
        HIP_TRY(hipMemcpy(&hc, cnt, sizeof(hc), hipMemcpyDeviceToHost));
        stats->points = n;
        stats->pairs = hc.pairs;
        stats->escalated = hc.escalated;
        stats->brent_iters = hc.brent;
        stats->kernel_ms = ms;
    }
    return OCL_OK;
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

static int launch_drop_fast(ocl_mesh*, const MeshView&, const CutterDev&, const double*, uint32_t, int,
                            const DropOut&, Counters*, hipStream_t) {
    return fail(OCL_E_UNSUPPORTED, "precision mode not built yet");
}
