// Operation classes of the drop-in host API (see ocl/operations.hpp).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <sstream>

#include "ocl/operations.hpp"

namespace ocl {

static void check(int rc) {
    if (rc != OCL_OK) throw Error(std::string("ocl_hip: ") + ocl_hip_last_error());
}

// ------------------------------------------------------------ Interval / Fiber

Interval::Interval(double l, double u) : upper(u), lower(l) {
    if (l > u) throw Error("Interval: lower > upper");
}

void Interval::updateUpper(double t, const CCPoint& p) {
    if (upper_cc.type == NONE) {
        upper = lower = t;
        upper_cc = lower_cc = p;
    }
    if (t > upper) {
        upper = t;
        upper_cc = p;
    }
}

void Interval::updateLower(double t, const CCPoint& p) {
    if (lower_cc.type == NONE) {
        upper = lower = t;
        upper_cc = lower_cc = p;
    }
    if (t < lower) {
        lower = t;
        lower_cc = p;
    }
}

std::string Interval::str() const {
    std::ostringstream o;
    o << "[" << lower << ", " << upper << "]";
    return o.str();
}

Fiber::Fiber(const Point& a, const Point& b) : p1(a), p2(b) {
    if (a.z != b.z) throw Error("Fiber: p1 and p2 must have the same z");
    dir = p2 - p1;
    dir.normalize();
}

bool Fiber::contains(const Interval& i) const {
    for (const Interval& f : ints)
        if (i.inside(f)) return true;
    return false;
}

bool Fiber::missing(const Interval& i) const {
    for (const Interval& f : ints)
        if (!i.outside(f)) return false;
    return true;
}

// fiber.cpp:58-93: drop empty intervals and intervals strictly inside one
// already held, append disjoint ones, otherwise replace every overlapping
// interval and i by their hull (earlier intervals win equal extremes)
void Fiber::addInterval(const Interval& i) {
    if (i.empty()) return;
    if (ints.empty() || missing(i)) {
        ints.push_back(i);
        return;
    }
    if (contains(i)) return;
    std::vector<Interval> keep, merged;
    for (const Interval& f : ints) (f.outside(i) ? keep : merged).push_back(f);
    merged.push_back(i);
    Interval hull;
    for (const Interval& m : merged) {
        hull.updateLower(m.lower, m.lower_cc);
        hull.updateUpper(m.upper, m.upper_cc);
    }
    keep.push_back(hull);
    ints.swap(keep);
}

double Fiber::tval(const Point& p) const {
    const Point d = p2 - p1;
    return (p - p1).dot(d) / d.dot(d);
}

std::string Fiber::str() const {
    std::ostringstream o;
    o << "Fiber " << p1.str() << " - " << p2.str() << " with " << ints.size() << " intervals";
    return o.str();
}

// ------------------------------------------------------------ device surface

DeviceSurface::~DeviceSurface() {
    if (mesh) ocl_hip_mesh_destroy(mesh);
}

void DeviceSurface::build(const STLSurf& s, unsigned bucket) {
    if (mesh) {
        ocl_hip_mesh_destroy(mesh);
        mesh = nullptr;
    }
    tris.assign(s.tris.begin(), s.tris.end());
    for (KdOrder& k : kd) k.built = false;
    const std::vector<double> v = s.flat();
    check(ocl_hip_mesh_create(v.data(), tris.size(), bucket, &mesh));
}

void Operation::setSTL(const STLSurf& s) {
    surf = &s;
    dev.build(s, bucketSize);
}

void Operation::require_ready() const {
    if (!dev.handle()) throw Error("setSTL() must be called before run()");
    if (!cutter) throw Error("setCutter() must be called before run()");
}

static std::vector<CLPoint> drop_points(ocl_mesh* m, const MillingCutter& c, std::vector<CLPoint> pts, int precision,
                                        ocl_stats* st) {
    const size_t n = pts.size();
    std::vector<double> xyz(3 * n), z(n), cc(3 * n);
    std::vector<uint8_t> ty(n);
    for (size_t i = 0; i < n; ++i) {
        xyz[3 * i] = pts[i].x;
        xyz[3 * i + 1] = pts[i].y;
        xyz[3 * i + 2] = pts[i].z;
    }
    const ocl_cutter_desc d = c.desc();
    check(ocl_hip_drop_batch(m, &d, xyz.data(), n, precision, z.data(), cc.data(), ty.data(), nullptr, nullptr, st,
                             nullptr));
    for (size_t i = 0; i < n; ++i) {
        pts[i].z = z[i];
        if (ty[i] != NONE) pts[i].cc = CCPoint(cc[3 * i], cc[3 * i + 1], cc[3 * i + 2], (CCType)ty[i]);
    }
    return pts;
}

// ------------------------------------------------------------ BatchDropCutter

void BatchDropCutter::run() {
    require_ready();
    clpoints = drop_points(dev.handle(), *cutter, std::move(clpoints), precision, &stats);
    nCalls = stats.pairs;
}

// The search reaches rank k iff the box reaches the rank's acceptance bounds
// on both search axes (include/ocl_hip.h, ocl_hip_kd_order).
std::vector<Triangle> DeviceSurface::kd_query(int dimset, unsigned bucket, double a0, double a1, double b0,
                                              double b1) {
    KdOrder& k = kd[dimset];
    if (!k.built || k.bucket != bucket) {
        std::vector<double> v;
        v.reserve(9 * tris.size());
        for (const Triangle& t : tris)
            for (const Point& q : t.p) {
                v.push_back(q.x);
                v.push_back(q.y);
                v.push_back(q.z);
            }
        k.order.resize(tris.size());
        k.acc.resize(4 * tris.size());
        check(ocl_hip_kd_order(v.data(), tris.size(), dimset, bucket, k.order.data(), k.acc.data()));
        k.bucket = bucket;
        k.built = true;
    }
    std::vector<Triangle> out;
    for (size_t r = 0; r < tris.size(); ++r) {
        const double* q = &k.acc[4 * r];
        if (a1 >= q[0] && a0 <= q[1] && b1 >= q[2] && b0 <= q[3]) out.push_back(tris[k.order[r]]);
    }
    return out;
}

std::vector<Triangle> BatchDropCutter::getTrianglesUnderCutter(const CLPoint& cl, const MillingCutter& c) {
    if (!dev.handle()) throw Error("setSTL() must be called first");
    const double r = c.getRadius();
    return dev.kd_query(0, bucketSize, cl.x - r, cl.x + r, cl.y - r, cl.y + r);
}

void PointDropCutter::run(CLPoint& cl) {
    require_ready();
    cl = drop_points(dev.handle(), *cutter, {cl}, precision, &stats)[0];
    nCalls = stats.pairs;
}

// ------------------------------------------------------------ BatchPushCutter

// Pushes fibers (all along one axis: dir 0 = X, 1 = Y) on the device and
// stores each fiber's intervals, merged into any it already held.
static std::vector<double> fiber_array(const std::vector<Fiber>& fibers) {
    std::vector<double> fib(6 * fibers.size());
    for (size_t i = 0; i < fibers.size(); ++i) {
        const Fiber& f = fibers[i];
        const double v[6] = {f.p1.x, f.p1.y, f.p1.z, f.p2.x, f.p2.y, f.p2.z};
        std::memcpy(&fib[6 * i], v, sizeof(v));
    }
    return fib;
}

// Adds the device rows of the last push in direction dir to the fibers'
// interval lists (as they are, on a fiber without intervals; merged by
// Fiber::addInterval otherwise).
static void attach_rows(ocl_mesh* m, int dir, std::vector<Fiber>& fibers, const std::vector<uint64_t>& offs,
                        uint64_t total) {
    std::vector<double> lohi(2 * total), cc(8 * total);
    if (total) check(ocl_hip_push_fetch_dir(m, dir, lohi.data(), cc.data(), nullptr));
    for (size_t i = 0; i < fibers.size(); ++i) {
        Fiber& f = fibers[i];
        const bool fresh = f.ints.empty();
        for (uint64_t k = offs[i]; k < offs[i + 1]; ++k) {
            Interval iv;
            iv.lower = lohi[2 * k];
            iv.upper = lohi[2 * k + 1];
            const double* q = &cc[8 * k];
            iv.lower_cc = CCPoint(q[0], q[1], q[2], (CCType)(int)q[3]);
            iv.upper_cc = CCPoint(q[4], q[5], q[6], (CCType)(int)q[7]);
            if (fresh)
                f.ints.push_back(iv);
            else
                f.addInterval(iv);
        }
    }
}

static void push_fibers(ocl_mesh* m, const MillingCutter& c, int dir, std::vector<Fiber>& fibers, int precision,
                        bool exact_filter, ocl_stats* st) {
    const size_t nf = fibers.size();
    const std::vector<double> fib = fiber_array(fibers);
    const ocl_cutter_desc d = c.desc();
    std::vector<uint64_t> offs(nf + 1, 0);
    uint64_t total = 0;
    const int prec = precision | (exact_filter ? OCL_PUSH_EXACT_FILTER : 0);
    check(ocl_hip_push_batch(m, &d, dir, fib.data(), nf, prec, offs.data(), &total, st, nullptr));
    attach_rows(m, dir, fibers, offs, total);
}

void BatchPushCutter::run() {
    require_ready();
    if (dir < 0) throw Error("BatchPushCutter: call setXDirection() or setYDirection() first");
    push_fibers(dev.handle(), *cutter, dir, fibers, precision, exact_filter, &stats);
    nCalls = stats.pairs;
}

std::vector<CLPoint> BatchPushCutter::getCLPoints() const {
    std::vector<CLPoint> out;
    for (const Fiber& f : fibers)
        for (const Interval& i : f.ints) {
            if (i.empty()) continue;
            out.emplace_back(f.point(i.lower));
            out.back().cc = i.lower_cc;
            out.emplace_back(f.point(i.upper));
            out.back().cc = i.upper_cc;
        }
    return out;
}

std::vector<Triangle> BatchPushCutter::getOverlapTriangles(const Fiber& f) {
    if (!dev.handle() || !cutter) throw Error("setSTL() and setCutter() must be called first");
    if (dir < 0) throw Error("BatchPushCutter: no fiber direction set");
    const double r = cutter->getRadius(), L = cutter->getLength();
    // X fibers: the YZ tree at (y, z); Y fibers: the XZ tree at (x, z)
    const double across = dir == 0 ? f.p1.y : f.p1.x;
    return dev.kd_query(dir == 0 ? 1 : 2, bucketSize, across - r, across + r, f.p1.z, f.p1.z + L);
}

// ------------------------------------------------------------ paths

void PathDropCutter::run() {
    require_ready();
    if (!path) throw Error("PathDropCutter: setPath() first");
    if (!(sampling > 0)) throw Error("PathDropCutter: sampling must be > 0");
    std::vector<CLPoint> pts;
    for (const Span* s : path->span_list) {
        // pathdropcutter.cpp:70-80: num_steps + 1 samples, z = minimumZ
        const unsigned steps = (unsigned)(s->length2d() / sampling + 1);
        for (unsigned i = 0; i <= steps; ++i) {
            const Point p = s->getPoint((double)i / steps);
            pts.emplace_back(p.x, p.y, minimumZ);
        }
    }
    clpoints = drop_points(dev.handle(), *cutter, std::move(pts), precision, &stats);
    nCalls = stats.pairs;
}

namespace {
struct Seg {
    size_t span;
    double t0, t1;
    CLPoint c0, c1;
};
struct Leaf {
    size_t span;
    double t1;
    CLPoint c1;
};
}  // namespace

static bool flat3(const CLPoint& a, const CLPoint& b, const CLPoint& c, double cos_limit) {
    Point v1 = b - a, v2 = c - b;
    v1.normalize();
    v2.normalize();
    return v1.dot(v2) > cos_limit;
}

void AdaptivePathDropCutter::run() {
    require_ready();
    if (!path) throw Error("AdaptivePathDropCutter: setPath() first");
    if (chord_tol > 0)
        slope_run();
    else
        flatness_run();
}

void AdaptivePathDropCutter::flatness_run() {
    std::vector<const Span*> spans(path->span_list.begin(), path->span_list.end());
    // span ends, all in one batch (adaptivepathdropcutter.cpp:59-69)
    std::vector<CLPoint> ends;
    for (const Span* s : spans) {
        ends.emplace_back(s->getPoint(0.0));
        ends.emplace_back(s->getPoint(1.0));
    }
    ends = drop_points(dev.handle(), *cutter, std::move(ends), precision, &stats);
    unsigned long long calls = stats.pairs;
    std::vector<Seg> level;
    for (size_t k = 0; k < spans.size(); ++k) level.push_back({k, 0.0, 1.0, ends[2 * k], ends[2 * k + 1]});
    std::vector<Leaf> leaves;
    // adaptivepathdropcutter.cpp:72-86, one recursion level per batch
    levels = 0;
    while (!level.empty()) {
        ++levels;
        std::vector<CLPoint> mids;
        mids.reserve(level.size());
        for (const Seg& g : level) mids.emplace_back(spans[g.span]->getPoint(g.t0 + (g.t1 - g.t0) / 2.0));
        mids = drop_points(dev.handle(), *cutter, std::move(mids), precision, &stats);
        calls += stats.pairs;
        std::vector<Seg> next;
        for (size_t i = 0; i < level.size(); ++i) {
            const Seg& g = level[i];
            const double tm = g.t0 + (g.t1 - g.t0) / 2.0;
            const double step = (g.c1 - g.c0).xyNorm();
            if (step > sampling || (!flat3(g.c0, mids[i], g.c1, cosLimit) && step > min_sampling)) {
                next.push_back({g.span, g.t0, tm, g.c0, mids[i]});
                next.push_back({g.span, tm, g.t1, mids[i], g.c1});
            } else {
                leaves.push_back({g.span, g.t1, g.c1});
            }
        }
        level.swap(next);
    }
    // depth-first order = by span, then by segment end
    std::stable_sort(leaves.begin(), leaves.end(), [](const Leaf& a, const Leaf& b) {
        return a.span != b.span ? a.span < b.span : a.t1 < b.t1;
    });
    clpoints.clear();
    size_t li = 0;
    for (size_t k = 0; k < spans.size(); ++k) {
        clpoints.push_back(ends[2 * k]);
        for (; li < leaves.size() && leaves[li].span == k; ++li) clpoints.push_back(leaves[li].c1);
    }
    nCalls = calls;
}

// ------------------------------------------------ slope-driven refinement

namespace {
struct SPt {
    CLPoint p;
    double m;  // slope of z along the path direction
};
// An open interval of the slope-driven refinement: its ends, and, once
// dropped, its middle (mid) and its quarter points (q1, q3).
struct SSeg {
    size_t span;
    double t0, t1;
    SPt a, b;
    int stage = 0;  // 0: its middle is dropped next; 1: its quarter points are
    SPt m, q1, q3;
};
}  // namespace

// unit XY tangent of a span at parameter t (direction of increasing t)
static Point span_tangent(const Span* s, double t) {
    Point d;
    if (s->type() == LineSpanType) {
        const Line& l = static_cast<const LineSpan*>(s)->line;
        d = l.p2 - l.p1;
    } else {
        const Arc& a = static_cast<const ArcSpan*>(s)->arc;
        const Point r = a.getPoint(t) - a.c;
        d = a.dir ? Point(-r.y, r.x, 0) : Point(r.y, -r.x, 0);
    }
    d.z = 0;
    const double l = d.xyNorm();
    return l > 0 ? d * (1.0 / l) : Point(0, 0, 0);
}

// Largest departure from the chord (0, z0) - (D, z1) of the piecewise cubic
// Hermite curve through (0, z0, m0), (D/2, zm, mm), (D, z1, m1), sampled at
// 1/16 steps: the interval's three heights and three slopes model z on it.
static double hermite_chord_error(double D, double z0, double zm, double z1, double m0, double mm, double m1) {
    const double h = D / 2;
    double e = std::fabs(zm - (z0 + z1) / 2);
    for (int half = 0; half < 2; ++half) {
        const double za = half ? zm : z0, zb = half ? z1 : zm;
        const double ma = half ? mm : m0, mb = half ? m1 : mm;
        for (int k = 1; k < 8; ++k) {
            const double u = k / 8.0;
            const double h00 = (1 + 2 * u) * (1 - u) * (1 - u), h10 = u * (1 - u) * (1 - u);
            const double h01 = u * u * (3 - 2 * u), h11 = u * u * (u - 1);
            const double zh = h00 * za + h10 * h * ma + h01 * zb + h11 * h * mb;
            const double pos = (half + u) / 2;  // fraction of D
            e = std::max(e, std::fabs(zh - (z0 + (z1 - z0) * pos)));
        }
    }
    return e;
}

static std::vector<SPt> drop_with_slopes(ocl_mesh* m, const MillingCutter& c, const std::vector<CLPoint>& pts,
                                         const std::vector<Point>& dirs, int precision, ocl_stats* st) {
    const size_t n = pts.size();
    std::vector<double> xyz(3 * n), z(n), g(2 * n), cc(3 * n);
    std::vector<uint8_t> ty(n);
    for (size_t i = 0; i < n; ++i) {
        xyz[3 * i] = pts[i].x;
        xyz[3 * i + 1] = pts[i].y;
        xyz[3 * i + 2] = pts[i].z;
    }
    const ocl_cutter_desc d = c.desc();
    check(ocl_hip_drop_slopes(m, &d, xyz.data(), n, precision, z.data(), g.data(), cc.data(), ty.data(), st,
                              nullptr));
    std::vector<SPt> out(n);
    for (size_t i = 0; i < n; ++i) {
        out[i].p = pts[i];
        out[i].p.z = z[i];
        if (ty[i] != NONE) out[i].p.cc = CCPoint(cc[3 * i], cc[3 * i + 1], cc[3 * i + 2], (CCType)ty[i]);
        out[i].m = g[2 * i] * dirs[i].x + g[2 * i + 1] * dirs[i].y;
    }
    return out;
}

void AdaptivePathDropCutter::slope_run() {
    if (!(sampling > 0) || !(min_sampling > 0)) throw Error("AdaptivePathDropCutter: sampling steps must be > 0");
    std::vector<const Span*> spans(path->span_list.begin(), path->span_list.end());
    // first samples: at most `sampling` apart
    std::vector<CLPoint> pts;
    std::vector<Point> dirs;
    std::vector<std::pair<size_t, double>> where;
    for (size_t k = 0; k < spans.size(); ++k) {
        const unsigned steps = std::max(1u, (unsigned)std::ceil(spans[k]->length2d() / sampling));
        for (unsigned i = 0; i <= steps; ++i) {
            const double t = (double)i / steps;
            pts.emplace_back(spans[k]->getPoint(t));
            dirs.push_back(span_tangent(spans[k], t));
            where.emplace_back(k, t);
        }
    }
    std::vector<SPt> first = drop_with_slopes(dev.handle(), *cutter, pts, dirs, precision, &stats);
    unsigned long long calls = stats.pairs;
    std::vector<SSeg> level;
    for (size_t i = 0; i + 1 < first.size(); ++i)
        if (where[i].first == where[i + 1].first)
            level.push_back({where[i].first, where[i].second, where[i + 1].second, first[i], first[i + 1]});
    // Each level drops, in one batch, the middle of every interval that
    // needs one and the quarter points of every interval under check.  An
    // interval whose Hermite model (ends and middle, heights and slopes)
    // stays within half the tolerance of its chord is checked: it closes if
    // its quarter points also lie within half the tolerance of the chord,
    // else it splits and the quarter points are its halves' middles.  An
    // interval at the minimum step closes.
    std::vector<SSeg> done;
    levels = 1;
    auto at = [&](const SSeg& g, double f) { return g.t0 + (g.t1 - g.t0) * f; };
    std::vector<SSeg> next;
    // an interval whose middle is known: close it at the minimum step, check
    // its quarter points next if its Hermite model passes, else split it
    auto decide = [&](SSeg g) {
        const double D = spans[g.span]->length2d() * (g.t1 - g.t0);
        if (D <= min_sampling) {
            done.push_back(g);
        } else if (hermite_chord_error(D, g.a.p.z, g.m.p.z, g.b.p.z, g.a.m, g.m.m, g.b.m) <= 0.5 * chord_tol) {
            g.stage = 1;
            next.push_back(g);
        } else {
            next.push_back({g.span, g.t0, at(g, 0.5), g.a, g.m});
            next.push_back({g.span, at(g, 0.5), g.t1, g.m, g.b});
        }
    };
    while (!level.empty()) {
        pts.clear();
        dirs.clear();
        auto sample = [&](const SSeg& g, double f) {
            pts.emplace_back(spans[g.span]->getPoint(at(g, f)));
            dirs.push_back(span_tangent(spans[g.span], at(g, f)));
        };
        for (const SSeg& g : level) {
            if (g.stage == 0) {
                sample(g, 0.5);
            } else {
                sample(g, 0.25);
                sample(g, 0.75);
            }
        }
        ++levels;
        std::vector<SPt> got = drop_with_slopes(dev.handle(), *cutter, pts, dirs, precision, &stats);
        calls += stats.pairs;
        size_t gi = 0;
        next.clear();
        for (SSeg g : level) {
            if (g.stage == 0) {
                g.m = got[gi++];
                decide(g);
                continue;
            }
            g.q1 = got[gi++];
            g.q3 = got[gi++];
            const double c1 = g.a.p.z + 0.25 * (g.b.p.z - g.a.p.z), c3 = g.a.p.z + 0.75 * (g.b.p.z - g.a.p.z);
            if (std::fabs(g.q1.p.z - c1) <= 0.5 * chord_tol && std::fabs(g.q3.p.z - c3) <= 0.5 * chord_tol) {
                done.push_back(g);
            } else {
                // the quarter points are the halves' middles
                SSeg lo{g.span, g.t0, at(g, 0.5), g.a, g.m};
                SSeg hi{g.span, at(g, 0.5), g.t1, g.m, g.b};
                lo.m = g.q1;
                hi.m = g.q3;
                decide(lo);
                decide(hi);
            }
        }
        level.swap(next);
    }
    std::sort(done.begin(), done.end(), [](const SSeg& a, const SSeg& b) {
        return a.span != b.span ? a.span < b.span : a.t0 < b.t0;
    });
    clpoints.clear();
    for (size_t i = 0; i < done.size(); ++i) {
        if (i == 0 || done[i].span != done[i - 1].span) clpoints.push_back(done[i].a.p);
        clpoints.push_back(done[i].b.p);
    }
    nCalls = calls;
}

// ------------------------------------------------------------ waterline

// n + 1 values from start to end, accumulated (waterline.cpp:169-178)
static std::vector<double> generate_range(double start, double end, int n) {
    std::vector<double> out;
    const double d = (end - start) / (double)n;
    double v = start;
    for (int i = 0; i <= n; ++i) {
        out.push_back(v);
        v = v + d;
    }
    return out;
}

void Waterline::run() {
    require_ready();
    if (!(sampling > 0)) throw Error("Waterline: sampling must be > 0");
    if (!surf || !surf->bb.initialized) throw Error("Waterline: empty surface");
    const double r = cutter->getRadius();
    const double minx = surf->bb.minpt.x - 2 * r, maxx = surf->bb.maxpt.x + 2 * r;
    const double miny = surf->bb.minpt.y - 2 * r, maxy = surf->bb.maxpt.y + 2 * r;
    const int nx = (int)((maxx - minx) / sampling), ny = (int)((maxy - miny) / sampling);
    xfibers.clear();
    yfibers.clear();
    for (double y : generate_range(miny, maxy, ny)) xfibers.emplace_back(Point(minx, y, zh), Point(maxx, y, zh));
    for (double x : generate_range(minx, maxx, nx)) yfibers.emplace_back(Point(x, miny, zh), Point(x, maxy, zh));
    // both directions in one device launch (the reference runs its two
    // BatchPushCutters one after the other, waterline.cpp:80-81)
    const std::vector<double> fx = fiber_array(xfibers), fy = fiber_array(yfibers);
    std::vector<uint64_t> ox(xfibers.size() + 1, 0), oy(yfibers.size() + 1, 0);
    uint64_t tot[2] = {0, 0};
    const ocl_cutter_desc d = cutter->desc();
    const int prec = precision | (exact_filter ? OCL_PUSH_EXACT_FILTER : 0);
    check(ocl_hip_push_xy(dev.handle(), &d, fx.data(), xfibers.size(), fy.data(), yfibers.size(), prec, ox.data(),
                          oy.data(), tot, &stats, nullptr));
    attach_rows(dev.handle(), 0, xfibers, ox, tot[0]);
    attach_rows(dev.handle(), 1, yfibers, oy, tot[1]);
    nCalls = stats.pairs;
}

std::vector<std::vector<Point>> Waterline::getLoops() const {
    throw Error("Waterline::getLoops: the weave (Boost Graph) that builds loops from the fibers is not part of "
                "this library; use getXFibers()/getYFibers()");
}

}  // namespace ocl
