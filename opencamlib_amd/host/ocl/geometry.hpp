// Geometry value types of the drop-in host API (namespace ocl).
//
// Names, members and meaning follow the reference's value types so that
// caller code compiles unchanged: Point (src/geo/point.hpp), CCType /
// CCPoint (src/geo/ccpoint.hpp:33-68), CLPoint (src/geo/clpoint.hpp), Bbox
// (src/geo/bbox.hpp), Triangle (src/geo/triangle.hpp:35-75), STLSurf
// (src/geo/stlsurf.hpp:38-56), Line / Arc / Path (src/geo/line.hpp,
// arc.hpp, path.hpp) and STLReader (src/geo/stlreader.hpp).
//
// Differences from the reference, all on purpose:
//   - CLPoint holds its CC point by value (the reference keeps an atomic
//     pointer that OpenMP workers swap); getCC() returns it.
//   - Errors throw ocl::Error instead of printing and asserting.
#pragma once

#include <cmath>
#include <list>
#include <stdexcept>
#include <string>
#include <vector>

namespace ocl {

struct Error : std::runtime_error {
    explicit Error(const std::string& m) : std::runtime_error(m) {}
};

struct Point {
    double x = 0, y = 0, z = 0;
    Point() = default;
    Point(double px, double py, double pz) : x(px), y(py), z(pz) {}
    Point(double px, double py) : x(px), y(py), z(0) {}

    double norm() const { return std::sqrt(x * x + y * y + z * z); }
    double xyNorm() const { return std::sqrt(x * x + y * y); }
    // scales to unit length; a zero vector stays zero (point.cpp:75-79)
    void normalize();
    double dot(const Point& p) const { return x * p.x + y * p.y + z * p.z; }
    Point cross(const Point& p) const { return Point(y * p.z - z * p.y, z * p.x - x * p.z, x * p.y - y * p.x); }
    double xyDistance(const Point& p) const { return (*this - p).xyNorm(); }
    Point xyPerp() const { return Point(-y, x, z); }
    void xyRotate(double angle);
    void xRotate(double theta);
    void yRotate(double theta);
    void zRotate(double theta);
    bool isRight(const Point& p1, const Point& p2) const;
    std::string str() const;

    Point operator+(const Point& p) const { return Point(x + p.x, y + p.y, z + p.z); }
    Point operator-(const Point& p) const { return Point(x - p.x, y - p.y, z - p.z); }
    Point operator*(double a) const { return Point(x * a, y * a, z * a); }
    Point& operator+=(const Point& p) { x += p.x; y += p.y; z += p.z; return *this; }
    Point& operator-=(const Point& p) { x -= p.x; y -= p.y; z -= p.z; return *this; }
    Point& operator*=(double a) { x *= a; y *= a; z *= a; return *this; }
    bool operator==(const Point& p) const { return x == p.x && y == p.y && z == p.z; }
    bool operator!=(const Point& p) const { return !(*this == p); }
};
inline Point operator*(double a, const Point& p) { return p * a; }

// CCType numbering of the reference (ccpoint.hpp:33-39).
enum CCType {
    NONE = 0, VERTEX, VERTEX_CYL, EDGE, EDGE_HORIZ, EDGE_SHAFT, EDGE_HORIZ_CYL, EDGE_HORIZ_TOR, EDGE_BALL,
    EDGE_POS, EDGE_NEG, EDGE_CYL, EDGE_CONE, EDGE_CONE_BASE, FACET, FACET_TIP, FACET_CYL, ERROR
};
const char* cctype_name(CCType t);

struct CCPoint : Point {
    CCType type = NONE;
    CCPoint() = default;
    CCPoint(double px, double py, double pz) : Point(px, py, pz) {}
    CCPoint(double px, double py, double pz, CCType t) : Point(px, py, pz), type(t) {}
    CCPoint(const Point& p, CCType t) : Point(p), type(t) {}
    std::string str() const;
};

struct Triangle;

struct CLPoint : Point {
    CCPoint cc;
    CLPoint() = default;
    CLPoint(double px, double py, double pz) : Point(px, py, pz) {}
    CLPoint(double px, double py, double pz, const CCPoint& c) : Point(px, py, pz), cc(c) {}
    CLPoint(const Point& p) : Point(p) {}
    const CCPoint& getCC() const { return cc; }
    // CLPoint::below (clpoint.cpp:63-68): z below the triangle's max z
    bool below(const Triangle& t) const;
    std::string str() const;
};

struct Bbox {
    Point minpt, maxpt;
    bool initialized = false;
    void clear() { initialized = false; }
    void addPoint(const Point& p);
    bool isInside(const Point& p) const;
    bool overlaps(const Bbox& b) const;
};

struct Triangle {
    Point p[3];
    Point n;  // unit normal of (p0-p1) x (p0-p2) (triangle.cpp:65-71)
    Bbox bb;
    Triangle() = default;
    Triangle(const Point& p1, const Point& p2, const Point& p3);
    std::vector<Point> getPoints() const { return {p[0], p[1], p[2]}; }
    void calcNormal();
    void calcBB();
    Point upNormal() const { return n.z < 0 ? n * -1.0 : n; }
    std::string str() const;
};

struct STLSurf {
    std::list<Triangle> tris;
    Bbox bb;
    void addTriangle(const Triangle& t);
    unsigned int size() const { return (unsigned int)tris.size(); }
    std::vector<Triangle> getTriangles() const { return std::vector<Triangle>(tris.begin(), tris.end()); }
    // rotate every vertex about the x, y and z axes in turn (stlsurf.cpp:43-59)
    void rotate(double xr, double yr, double zr);
    // bbox as (minx, maxx, miny, maxy, minz, maxz)
    std::vector<double> getBounds() const;
    // vertices as 9 doubles per triangle, list order: the device upload format
    std::vector<double> flat() const;
    std::string str() const;
};

// Reads binary or ASCII STL into surface (stlreader.cpp:57-150): a file whose
// first five bytes are "solid" is parsed as text, otherwise as binary; the
// coordinates are float32 in both formats.  A missing file adds nothing.
struct STLReader {
    STLReader(const std::wstring& path, STLSurf& surface);
    STLReader(const std::string& path, STLSurf& surface);
    static void read_from_file(const std::string& path, STLSurf& surface);
};

struct Line {
    Point p1, p2;
    Line() = default;
    Line(const Point& a, const Point& b) : p1(a), p2(b) {}
    double length2d() const { return (p2 - p1).xyNorm(); }
    Point getPoint(double t) const { return (p2 - p1) * t + p1; }
};

// Circular arc in XY from p1 to p2 about c, counter-clockwise if dir
// (arc.cpp:30-98).
struct Arc {
    Point p1, p2, c;
    bool dir = true;
    double radius = 0, length = 0;
    Arc() = default;
    Arc(const Point& a, const Point& b, const Point& centre, bool ccw);
    double length2d() const { return length; }
    Point getPoint(double t) const;
    static double xyIncludedAngle(const Point& v1, const Point& v2, bool dir = true);
};

enum SpanType { LineSpanType, ArcSpanType };

struct Span {
    virtual ~Span() = default;
    virtual SpanType type() const = 0;
    virtual double length2d() const = 0;
    virtual Point getPoint(double t) const = 0;
};

struct LineSpan : Span {
    Line line;
    explicit LineSpan(const Line& l) : line(l) {}
    SpanType type() const override { return LineSpanType; }
    double length2d() const override { return line.length2d(); }
    Point getPoint(double t) const override { return line.getPoint(t); }
};

struct ArcSpan : Span {
    Arc arc;
    explicit ArcSpan(const Arc& a) : arc(a) {}
    SpanType type() const override { return ArcSpanType; }
    double length2d() const override { return arc.length2d(); }
    Point getPoint(double t) const override { return arc.getPoint(t); }
};

class Path {
  public:
    Path() = default;
    Path(const Path& p);
    Path& operator=(const Path& p);
    ~Path();
    void append(const Line& l);
    void append(const Arc& a);
    std::list<Span*> span_list;
};

}  // namespace ocl
