// Geometry value types of the drop-in host API (see ocl/geometry.hpp).
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

#include "ocl/geometry.hpp"
#include "ocl_hip.h"

namespace ocl {

static const char* const kCCNames[] = {"NONE", "VERTEX", "VERTEX_CYL", "EDGE", "EDGE_HORIZ", "EDGE_SHAFT",
                                       "EDGE_HORIZ_CYL", "EDGE_HORIZ_TOR", "EDGE_BALL", "EDGE_POS", "EDGE_NEG",
                                       "EDGE_CYL", "EDGE_CONE", "EDGE_CONE_BASE", "FACET", "FACET_TIP",
                                       "FACET_CYL", "ERROR"};

const char* cctype_name(CCType t) {
    const int i = (int)t;
    return (i >= 0 && i <= (int)ERROR) ? kCCNames[i] : "?";
}

void Point::normalize() {
    const double l = norm();
    if (l != 0.0) *this *= (1 / l);
}

void Point::xyRotate(double angle) {
    const double c = std::cos(angle), s = std::sin(angle);
    const double nx = -y * s + x * c;  // point.cpp:95-99
    y = x * s + c * y;
    x = nx;
}

// rotations about the coordinate axes (point.cpp:105-134)
static void rot3(Point& p, const double m[9]) {
    const double a = m[0] * p.x + m[1] * p.y + m[2] * p.z;
    const double b = m[3] * p.x + m[4] * p.y + m[5] * p.z;
    const double c = m[6] * p.x + m[7] * p.y + m[8] * p.z;
    p.x = a;
    p.y = b;
    p.z = c;
}
void Point::xRotate(double t) {
    const double m[9] = {1, 0, 0, 0, std::cos(t), -std::sin(t), 0, std::sin(t), std::cos(t)};
    rot3(*this, m);
}
void Point::yRotate(double t) {
    const double m[9] = {std::cos(t), 0, std::sin(t), 0, 1, 0, -std::sin(t), 0, std::cos(t)};
    rot3(*this, m);
}
void Point::zRotate(double t) {
    const double m[9] = {std::cos(t), -std::sin(t), 0, std::sin(t), std::cos(t), 0, 0, 0, 1};
    rot3(*this, m);
}

// right of the directed XY line p1 -> p2 (point.cpp:204-222)
bool Point::isRight(const Point& p1, const Point& p2) const {
    const double ex = p2.x - p1.x, ey = p2.y - p1.y;
    return ey * (x - p1.x) - ex * (y - p1.y) > 0.0;
}

std::string Point::str() const {
    std::ostringstream o;
    o << "(" << x << ", " << y << ", " << z << ")";
    return o.str();
}

std::string CCPoint::str() const { return Point::str() + "(t=" + cctype_name(type) + ")"; }

bool CLPoint::below(const Triangle& t) const { return z < t.bb.maxpt.z; }

std::string CLPoint::str() const { return Point::str() + " cc=" + cc.str(); }

void Bbox::addPoint(const Point& p) {
    if (!initialized) {
        minpt = maxpt = p;
        initialized = true;
        return;
    }
    minpt.x = std::fmin(minpt.x, p.x);
    minpt.y = std::fmin(minpt.y, p.y);
    minpt.z = std::fmin(minpt.z, p.z);
    maxpt.x = std::fmax(maxpt.x, p.x);
    maxpt.y = std::fmax(maxpt.y, p.y);
    maxpt.z = std::fmax(maxpt.z, p.z);
}

bool Bbox::isInside(const Point& p) const {
    if (!initialized) throw Error("Bbox::isInside on an empty box");
    return p.x >= minpt.x && p.x <= maxpt.x && p.y >= minpt.y && p.y <= maxpt.y && p.z >= minpt.z &&
           p.z <= maxpt.z;
}

bool Bbox::overlaps(const Bbox& b) const {
    return !(maxpt.x < b.minpt.x || minpt.x > b.maxpt.x || maxpt.y < b.minpt.y || minpt.y > b.maxpt.y ||
             maxpt.z < b.minpt.z || minpt.z > b.maxpt.z);
}

Triangle::Triangle(const Point& p1, const Point& p2, const Point& p3) {
    p[0] = p1;
    p[1] = p2;
    p[2] = p3;
    calcNormal();
    calcBB();
}

void Triangle::calcNormal() {
    n = (p[0] - p[1]).cross(p[0] - p[2]);
    n.normalize();
}

void Triangle::calcBB() {
    bb.clear();
    for (const Point& q : p) bb.addPoint(q);
}

std::string Triangle::str() const { return "T: " + p[0].str() + " " + p[1].str() + " " + p[2].str() + " n=" + n.str(); }

void STLSurf::addTriangle(const Triangle& t) {
    tris.push_back(t);
    for (const Point& q : t.p) bb.addPoint(q);
}

void STLSurf::rotate(double xr, double yr, double zr) {
    bb.clear();
    for (Triangle& t : tris) {
        for (Point& q : t.p) {
            q.xRotate(xr);
            q.yRotate(yr);
            q.zRotate(zr);
        }
        t.calcNormal();
        t.calcBB();
        for (const Point& q : t.p) bb.addPoint(q);
    }
}

std::vector<double> STLSurf::getBounds() const {
    return {bb.minpt.x, bb.maxpt.x, bb.minpt.y, bb.maxpt.y, bb.minpt.z, bb.maxpt.z};
}

std::vector<double> STLSurf::flat() const {
    std::vector<double> v;
    v.reserve(9 * tris.size());
    for (const Triangle& t : tris)
        for (const Point& q : t.p) {
            v.push_back(q.x);
            v.push_back(q.y);
            v.push_back(q.z);
        }
    return v;
}

std::string STLSurf::str() const { return "STLSurf(N=" + std::to_string(tris.size()) + ")"; }

// ------------------------------------------------------------- STL files

STLReader::STLReader(const std::wstring& path, STLSurf& surface) {
    std::string narrow;
    for (wchar_t c : path) narrow.push_back((char)c);
    read_from_file(narrow, surface);
}

STLReader::STLReader(const std::string& path, STLSurf& surface) { read_from_file(path, surface); }

// The C ABI's reader (ocl_hip_stl_read) fills a flat array; the reference
// returns silently with no triangles when the file cannot be read.
void STLReader::read_from_file(const std::string& path, STLSurf& surface) {
    size_t n = 0;
    if (ocl_hip_stl_read(path.c_str(), nullptr, 0, &n) != OCL_OK || n == 0) return;
    std::vector<double> v(9 * n);
    if (ocl_hip_stl_read(path.c_str(), v.data(), n, &n) != OCL_OK) return;
    for (size_t i = 0; i < n; ++i) {
        const double* q = &v[9 * i];
        surface.addTriangle(Triangle(Point(q[0], q[1], q[2]), Point(q[3], q[4], q[5]), Point(q[6], q[7], q[8])));
    }
}

// ------------------------------------------------------------- paths

Arc::Arc(const Point& a, const Point& b, const Point& centre, bool ccw) : p1(a), p2(b), c(centre), dir(ccw) {
    // arc.cpp:48-56: start and end tangents, radius from the start point
    Point vs = (p1 - c).xyPerp();
    Point ve = (p2 - c).xyPerp();
    radius = vs.xyNorm();
    vs.normalize();
    ve.normalize();
    length = std::fabs(xyIncludedAngle(vs, ve, dir)) * radius;
}

// signed included angle from v1 to v2, turning counter-clockwise if dir
// (arc.cpp:74-95)
double Arc::xyIncludedAngle(const Point& v1, const Point& v2, bool dir) {
    const double kPi = 3.1415926535897932;
    const int sgn = dir ? 1 : -1;
    double a = v1.dot(v2);
    if (a > 1. - 1.0e-10) return 0;
    if (a < -1. + 1.0e-10) {
        a = kPi;
    } else {
        a = std::acos(a);
        const double crs = v1.x * v2.y - v1.y * v2.x;
        if (sgn * crs < 0) a = 2 * kPi - a;
    }
    return sgn * a;
}

Point Arc::getPoint(double t) const {
    if (std::fabs(t) < 1e-14) return p1;
    if (std::fabs(t - 1.0) < 1e-14) return p2;
    double d = t * length;
    if (!dir) d = -d;
    Point v = p1 - c;
    v.xyRotate(d / radius);
    return v + c;
}

Path::Path(const Path& p) { *this = p; }

Path& Path::operator=(const Path& p) {
    if (this == &p) return *this;
    for (Span* s : span_list) delete s;
    span_list.clear();
    for (const Span* s : p.span_list) {
        if (s->type() == LineSpanType)
            span_list.push_back(new LineSpan(static_cast<const LineSpan*>(s)->line));
        else
            span_list.push_back(new ArcSpan(static_cast<const ArcSpan*>(s)->arc));
    }
    return *this;
}

Path::~Path() {
    for (Span* s : span_list) delete s;
}

void Path::append(const Line& l) { span_list.push_back(new LineSpan(l)); }
void Path::append(const Arc& a) { span_list.push_back(new ArcSpan(a)); }

}  // namespace ocl
