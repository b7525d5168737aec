// Geometry value types of the drop-in host API (see ocl/geometry.hpp).
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

#include "ocl/geometry.hpp"

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

static Triangle tri_from(const float v[3][3]) {
    return Triangle(Point(v[0][0], v[0][1], v[0][2]), Point(v[1][0], v[1][1], v[1][2]), Point(v[2][0], v[2][1], v[2][2]));
}

// Text STL: "vertex" lines fill up to three vertices (a fourth overwrites the
// third), "facet" starts a facet, "endfacet" adds it once the third vertex
// slot was reached; keywords are matched on their first five non-blank
// characters (stlreader.cpp:92-146).
static void read_ascii(std::ifstream& in, STLSurf& s) {
    std::string line;
    std::getline(in, line);  // rest of the "solid" line
    float v[3][3] = {};
    int k = 0;
    while (std::getline(in, line)) {
        char key[6] = {0, 0, 0, 0, 0, 0};
        size_t j = 0;
        int i = 0;
        for (; i < 5; ++i, ++j) {
            if (j >= line.size()) break;
            while (j < line.size() && (line[j] == ' ' || line[j] == '\t')) ++j;
            if (j >= line.size()) break;
            key[i] = line[j];
        }
        if (i < 5) continue;
        if (!std::strcmp(key, "verte")) {
            const size_t at = line.find("vertex");
            if (at != std::string::npos)
                std::sscanf(line.c_str() + at + 6, "%f %f %f", &v[k][0], &v[k][1], &v[k][2]);
            k = k + 1 > 2 ? 2 : k + 1;
        } else if (!std::strcmp(key, "facet")) {
            k = 0;
        } else if (!std::strcmp(key, "endfa")) {
            if (k == 2) s.addTriangle(tri_from(v));
        }
    }
}

// Binary STL: 80-byte header, uint32 count, then per facet 12 float32 (normal,
// three vertices) and a 2-byte attribute (stlreader.cpp:70-89).
static void read_binary(std::ifstream& in, STLSurf& s) {
    char header[75];
    in.read(header, 75);
    uint32_t count = 0;
    in.read(reinterpret_cast<char*>(&count), 4);
    for (uint32_t i = 0; i < count && in; ++i) {
        unsigned char rec[50];
        in.read(reinterpret_cast<char*>(rec), 50);
        if (!in) break;
        float v[3][3];
        std::memcpy(v, rec + 12, 36);
        s.addTriangle(tri_from(v));
    }
}

void STLReader::read_from_file(const std::string& path, STLSurf& surface) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return;
    char head[5];
    in.read(head, 5);
    if (in.gcount() < 5) return;
    if (std::memcmp(head, "solid", 5) == 0)
        read_ascii(in, surface);
    else
        read_binary(in, surface);
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
