// Reference-order FP64 vector helpers shared by the exact drop and push paths.
//
// These restate the arithmetic of ocl::Point (src/geo/point.cpp) operation by
// operation, so that GPU FP64 results match the reference bit for bit where
// only +,-,*,/ and sqrt are involved (all device code is compiled with
// -ffp-contract=off; AMDGPU FP64 division and sqrt are correctly rounded).
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#define OCL_HD __host__ __device__ inline
#else
#define OCL_HD inline
#endif

namespace ocl_amd {

#if !defined(__HIPCC__)
using std::signbit;
using std::sqrt;
using std::fabs;
#endif

struct V3 {
    double x, y, z;
};

OCL_HD V3 v3(double x, double y, double z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
OCL_HD V3 vadd(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
OCL_HD V3 vsub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
// scalar * Point: Point(p) *= s  (point.cpp:270-275, 300-303)
OCL_HD V3 vscale(double s, V3 p) { return v3(p.x * s, p.y * s, p.z * s); }
OCL_HD double sqr(double a) { return a * a; }
OCL_HD double vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
OCL_HD V3 vcross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
OCL_HD double vnorm(V3 p) { return sqrt(sqr(p.x) + sqr(p.y) + sqr(p.z)); }
OCL_HD double vxynorm(V3 p) { return sqrt(sqr(p.x) + sqr(p.y)); }
// Point::normalize multiplies by the reciprocal (point.cpp:75-79)
OCL_HD V3 vunit(V3 p) {
    const double n = vnorm(p);
    return (n != 0.0) ? vscale(1 / n, p) : p;
}
// Point::xyNormalize scales all three components (point.cpp:85-88)
OCL_HD V3 vxyunit(V3 p) {
    if (vxynorm(p) != 0.0) p = vscale(1 / vxynorm(p), p);
    return p;
}
OCL_HD V3 vperp(V3 p) { return v3(-p.y, p.x, p.z); }
OCL_HD bool near0(double a) { return fabs(a) < 1e-7; }  // isZero_tol, numeric.cpp:75-80
OCL_HD double sgn1(double a) { return a < 0.0 ? -1.0 : 1.0; }

// point.cpp:145-163
OCL_HD double xy_line_distance(V3 q, V3 p1, V3 p2) {
    if (p1.x == p2.x && p1.y == p2.y) return -1;
    const V3 n = vunit(v3(p2.y - p1.y, -(p2.x - p1.x), 0));
    return fabs(vdot(n, v3(p1.x - q.x, p1.y - q.y, 0)));
}
// point.cpp:166-172 (3D)
OCL_HD V3 line_closest(V3 q, V3 p1, V3 p2) {
    const V3 v = vsub(p2, p1);
    const double u = vdot(vsub(q, p1), v) / vdot(v, v);
    return vadd(p1, vscale(u, v));
}
// point.cpp:175-201 (XY, z = 0)
OCL_HD V3 xy_line_closest(V3 q, V3 p1, V3 p2) {
    const V3 v = vsub(p2, p1);
    if (near0(vxynorm(v))) return v3(0, 0, 0);
    double u = (q.x - p1.x) * (v.x) + (q.y - p1.y) * (v.y);
    u = u / (v.x * v.x + v.y * v.y);
    return v3(p1.x + u * v.x, p1.y + u * v.y, 0);
}
// point.cpp:337-347
OCL_HD double edge_z_at(V3 q, V3 p1, V3 p2) {
    double t;
    if (fabs(p2.x - p1.x) > fabs(p2.y - p1.y)) t = (q.x - p1.x) / (p2.x - p1.x);
    else t = (q.y - p1.y) / (p2.y - p1.y);
    return p1.z + t * (p2.z - p1.z);
}
// point.cpp:241-253: inclusive 3D segment parameter test (NaN passes)
OCL_HD bool on_segment(V3 q, V3 p1, V3 p2) {
    const V3 d = vsub(p2, p1);
    const double t = vdot(vsub(q, p1), d) / vdot(d, d);
    return !(t > 1.0 || t < 0.0);
}
// point.cpp:224-239: strict XY barycentric test
OCL_HD bool in_triangle_xy(V3 p, V3 a, V3 b, V3 c) {
    const double u = (a.y * c.x - a.x * c.y + (c.y - a.y) * p.x + (a.x - c.x) * p.y) /
                     (a.y * c.x - a.x * c.y + (c.y - a.y) * b.x + (a.x - c.x) * b.y);
    const double v = (a.x * b.y - a.y * b.x + (a.y - b.y) * p.x + (b.x - a.x) * p.y) /
                     (a.x * b.y - a.y * b.x + (a.y - b.y) * c.x + (b.x - a.x) * c.y);
    return u > 0.0 && v > 0.0 && (u + v) < 1.0;
}
// numeric.cpp:118-140
OCL_HD bool solve2x2(double a, double b, double c, double d, double e, double f, double& u, double& v) {
    const double det = a * d - c * b;
    if (near0(det)) return false;
    u = (1.0 / det) * (d * e - b * f);
    v = (1.0 / det) * (-c * e + a * f);
    return true;
}
// numeric.cpp:147-154
OCL_HD bool xy_lines_cross(V3 p1, V3 p2, double& v, V3 p3, V3 p4, double& t) {
    return solve2x2(p2.x - p1.x, -(p4.x - p3.x), p2.y - p1.y, -(p4.y - p3.y), (p3.x - p1.x), p3.y - p1.y, v, t);
}
// numeric.cpp:36-51
OCL_HD double diangle_of(double x, double y) {
    if (y >= 0) return (x >= 0 ? y / (x + y) : 1 - x / (-x + y));
    return (x < 0 ? 2 - y / (-x - y) : 3 + x / (x - y));
}
// ellipseposition.cpp:47-70: (s,t) on the unit circle for a diamond angle
OCL_HD void diangle_st(double dia, double& s, double& t) {
    double d = dia;
    while (d > 4.0) d -= 4.0;
    while (d < 0.0) d += 4.0;
    const V3 p = vunit(v3((d < 2 ? 1 - d : d - 3), (d < 3 ? ((d > 1) ? 2 - d : d) : d - 4), 0.0));
    s = p.x;
    t = p.y;
}

}  // namespace ocl_amd
