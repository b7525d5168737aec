// Backward pass of the drop-cutter: derivatives of the dropped height z at a
// CL point from the active contact the forward pass recorded (triangle,
// feature code, CC point).  There is no counterpart in the reference.
//
// With P(s) the contact on the feature (nothing for a vertex, the segment
// parameter for an edge, the barycentrics for a facet), the drop height is
// z = max_s P(s).z - h(q(s)), q = |P_xy - cl_xy|, h the profile height:
//   - profile smooth at q (ball, bull torus and flat bottom, cone flank, cyl
//     inside its disc): envelope theorem, dz/dx = h'(q) (P.x - x) / q,
//     dz/dP.z = 1, dz/dP.xy = -dz/d(cl.xy), spread over the vertices by the
//     contact's weights, dz/dr = -dh/dr at fixed q;
//   - contact on the rim q = r (cyl edges and facets, cone base): the
//     constraint c = q^2 - r^2 = 0 is active with multiplier
//     lambda = (grad_s P.z . grad_s c) / |grad_s c|^2, so
//     dz/dx = 2 lambda (P.x - x), dz/dP.xy = -2 lambda (P - cl)_xy and
//     dz/dr = -dh(r)/dr + 2 lambda r;
//   - cone tip on a facet (q = 0): z is the plane height under the CL point.
// Composite cutters use the active sub-cutter (a vertex contact: the
// sub-cutter owning its radius, compositecutter.cpp:58-61); their dz/dr is
// NaN.  A cone edge contact in the reference's hyperbola case is
// differentiated as the reference writes it (cone_hyper_edge_grad): its
// formula takes the cutter's length where the cone's contact takes the
// center height, so for a cone with a shank (l > 0) it is not the cone's
// envelope and the envelope theorem does not hold there.
//
// T = float computes facet and vertex contacts in FP32 on coordinates
// relative to the contact point; edge contacts always run in FP64.
#pragma once

#include "cutter.hpp"
#include "drop_exact.hpp"

namespace ocl_amd {

struct GradOut {
    double gx, gy, gr;
    double gv[9];  // dz/d(p0, p1, p2 xyz)
};

template <class T>
OCL_HD void prof_slope(const CutterP& s, T q, T& hp, T& hr) {
    hp = T(0);
    hr = T(0);
    if (s.kind == K_BALL) {
        const T r = (T)s.radius;
        const T w = sqrt(r * r - q * q);
        hp = q / w;
        hr = T(1) - r / w;
    } else if (s.kind == K_BULL) {
        const T r1 = (T)s.r1, r2 = (T)s.r2;
        if (q > r1) {
            const T w = sqrt(r2 * r2 - (q - r1) * (q - r1));
            hp = (q - r1) / w;
            hr = -hp;  // d/d(radius) with the corner radius fixed
        }
    } else if (s.kind == K_CONE) {
        hp = T(1) / (T)s.tan_a;
    }
}

OCL_HD double rim_height_slope(const CutterP& s) {
    return s.kind == K_BALL ? 1.0 : (s.kind == K_CONE ? 1.0 / s.tan_a : 0.0);
}

// The reference's cone edge drop at (x, y) takes its hyperbola case for the
// edge a -> b (conecutter.cpp:115-117).
OCL_HD bool cone_edge_hyper(const CutterP& s, V3 a, V3 b, double x, double y) {
    const double ex = b.x - a.x, ey = b.y - a.y;
    const double lxy = sqrt(ex * ex + ey * ey);
    const double d = fabs((ex * (y - a.y) - ey * (x - a.x)) / lxy);
    const double m = (b.z - a.z) / lxy;
    const double xu = sqrt(s.radius * s.radius - d * d);
    const double mu = (s.ch / s.radius) * xu / sqrt(xu * xu + d * d);
    return fabs(m) <= fabs(mu);
}

// Derivatives of the reference's cone edge height in its hyperbola case
// (conecutter.cpp:109-135), z = a.z + M s + G d: s and d the CL point's
// offset along and distance from the edge's XY line (direction v, left
// normal n, signed offset c = n.(cl - a), d = |c|), M the edge slope,
// G = K M^2 - S / tan, K = r / Q, Q = sqrt(L^2 - r^2 M^2),
// S = sqrt(K^2 M^2 + 1), L = r / tan + l (l, the shank, fixed when r
// varies).  Differentiated by hand: dz/dcl = M v + G sgn(c) n,
// dz/dr = d dG/dr, and for the end points, with P = s + d dG/dM and
// |e| the edge's XY length:
//   dz/db.xy = -P M v / |e| + (M c - G sgn(c) s) n / |e|,  dz/db.z = P / |e|,
//   dz/da.xy = -dz/db.xy - dz/dcl,                        dz/da.z = 1 - P / |e|.
OCL_HD void cone_hyper_edge_grad(const CutterP& cs, V3 a, V3 b, double x, double y, double* g9) {
    const double T = cs.tan_a, l = cs.length - cs.radius / T, R = cs.radius;
    const double ex = b.x - a.x, ey = b.y - a.y;
    const double le = sqrt(ex * ex + ey * ey);
    const double vx = ex / le, vy = ey / le, nx = -vy, ny = vx;
    const double M = (b.z - a.z) / le;
    const double wx = x - a.x, wy = y - a.y;
    const double s = wx * vx + wy * vy, c = wx * nx + wy * ny;
    const double sg = c < 0.0 ? -1.0 : 1.0, d = fabs(c);
    const double L = R / T + l, M2 = M * M;
    const double Q = sqrt(L * L - R * R * M2);
    const double K = R / Q, S = sqrt(K * K * M2 + 1.0);
    const double G = K * M2 - S / T;
    const double KM = R * R * R * M / (Q * Q * Q);               // dK/dM
    const double KR = 1.0 / Q - R * (L / T - R * M2) / (Q * Q * Q);  // dK/dr
    const double GM = KM * M2 + 2.0 * K * M - (K * M2 * KM + K * K * M) / (S * T);
    const double GR = KR * (M2 - K * M2 / (S * T));
    const double P = s + d * GM;
    const double gx = M * vx + G * sg * nx, gy = M * vy + G * sg * ny;
    const double bn = (M * c - G * sg * s) / le;
    const double bx = -P * M * vx / le + bn * nx, by = -P * M * vy / le + bn * ny;
    g9[0] = gx;
    g9[1] = gy;
    g9[2] = d * GR;
    g9[3] = -bx - gx;
    g9[4] = -by - gy;
    g9[5] = 1.0 - P / le;
    g9[6] = bx;
    g9[7] = by;
    g9[8] = P / le;
}

// Sub-cutter of a composite contact: facets and edges carry it in the
// feature code; a vertex contact uses the sub-cutter whose radius window
// holds q (compositecutter.cpp:58-61, 75-83, 111-122).
OCL_HD int grad_sub_index(const CutterDev& cd, int feat, double q) {
    const int k = feat & 15;
    if (!(k >= FT_VTX0 && k < FT_EDGE0)) return (feat >> 4) & 3;
    int n = 0;
#pragma unroll
    for (int j = kMaxSub - 1; j >= 0; --j) {
        const double lo = (j == 0) ? -1E-6 : cd.rlim[j > 0 ? j - 1 : 0] - 1E-6;
        if (j < cd.nsub && lo <= q && q <= cd.rlim[j] + 1e-6) n = j;
    }
    return n;
}

// s: the (sub-)cutter in contact, V: the triangle's vertices, P: the contact
// (FP64, from the forward pass); composite: the caller sets gr to NaN
template <class T>
__host__ __device__ __attribute__((always_inline)) inline GradOut contact_grad(const CutterP& s, const V3* V, int feat,
                                                                              int cctype, double x, double y, V3 P) {
    GradOut o{};
    const int k = feat & 15;
    // local frame: everything relative to the contact point
    const T dx = (T)(P.x - x), dy = (T)(P.y - y);
    const T q = sqrt(dx * dx + dy * dy);
    T w0 = 0, w1 = 0, w2 = 0;
    T fs0 = 0, fs1 = 0, cs0 = 0, cs1 = 0;
    int m = 0;
    if (k >= FT_FACET && k < FT_VTX0) {
        const T e1x = (T)(V[1].x - V[0].x), e1y = (T)(V[1].y - V[0].y), e1z = (T)(V[1].z - V[0].z);
        const T e2x = (T)(V[2].x - V[0].x), e2y = (T)(V[2].y - V[0].y), e2z = (T)(V[2].z - V[0].z);
        const T px = (T)(P.x - V[0].x), py = (T)(P.y - V[0].y);
        const T det = e1x * e2y - e1y * e2x;
        const T u = (px * e2y - py * e2x) / det, v = (e1x * py - e1y * px) / det;
        w0 = T(1) - u - v;
        w1 = u;
        w2 = v;
        m = 2;
        fs0 = e1z;
        fs1 = e2z;
        cs0 = T(2) * (dx * e1x + dy * e1y);
        cs1 = T(2) * (dx * e2x + dy * e2y);
    } else if (k < FT_EDGE0) {
        const int j = k - FT_VTX0;
        w0 = j == 0 ? T(1) : T(0);
        w1 = j == 1 ? T(1) : T(0);
        w2 = j == 2 ? T(1) : T(0);
    }
    T gx, gy, gr, vx, vy;
    if (k >= FT_EDGE0) {
        // edge: FP64
        const int a = k - FT_EDGE0, b = (a + 1) % 3;
        const V3 pa = a == 0 ? V[0] : (a == 1 ? V[1] : V[2]);
        const V3 pb = b == 0 ? V[0] : (b == 1 ? V[1] : V[2]);
        const double ex = pb.x - pa.x, ey = pb.y - pa.y, ez = pb.z - pa.z;
        const double t = ((P.x - pa.x) * ex + (P.y - pa.y) * ey) / (ex * ex + ey * ey);
        const double ddx = P.x - x, ddy = P.y - y;
        const double dq = sqrt(ddx * ddx + ddy * ddy);
        if (s.kind == K_CONE && cone_edge_hyper(s, pa, pb, x, y)) {
            double g9[9];
            cone_hyper_edge_grad(s, pa, pb, x, y, g9);
            o.gx = g9[0];
            o.gy = g9[1];
            o.gr = g9[2];
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int q = 0; q < 3; ++q) o.gv[3 * j + q] = (j == a) ? g9[3 + q] : ((j == b) ? g9[6 + q] : 0.0);
            return o;
        }
        double egx, egy, egr;
        const bool rim = s.kind == K_CYL || (s.kind == K_CONE && dq >= s.radius * (1.0 - 1e-9));
        if (rim) {
            const double c0 = 2.0 * (ddx * ex + ddy * ey);
            const double lam = c0 != 0.0 ? ez / c0 : 0.0;
            egx = 2.0 * lam * ddx;
            egy = 2.0 * lam * ddy;
            egr = -rim_height_slope(s) + 2.0 * lam * s.radius;
        } else {
            double hp, hr;
            prof_slope<double>(s, dq, hp, hr);
            egx = dq > 0.0 ? hp * ddx / dq : 0.0;
            egy = dq > 0.0 ? hp * ddy / dq : 0.0;
            egr = -hr;
        }
        o.gx = egx;
        o.gy = egy;
        o.gr = egr;
        const double wa = 1.0 - t, wb = t;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double wj = (j == a) ? wa : ((j == b) ? wb : 0.0);
            o.gv[3 * j] = -egx * wj;
            o.gv[3 * j + 1] = -egy * wj;
            o.gv[3 * j + 2] = wj;
        }
        return o;
    }
    // conditioning test: near the vertical end of a ball or torus profile
    // (steep facets) sqrt(r^2 - q^2) cancels in FP32; evaluate in FP64
    if (sizeof(T) < sizeof(double) &&
        ((s.kind == K_BALL && q > T(0.9) * (T)s.radius) ||
         (s.kind == K_BULL && q - (T)s.r1 > T(0.9) * (T)s.r2)))
        return contact_grad<double>(s, V, feat, cctype, x, y, P);
    const bool tip = s.kind == K_CONE && m == 2 && cctype == CC_FACET_TIP;
    const bool rim = m > 0 && (s.kind == K_CYL || (s.kind == K_CONE && k == FT_FACET_RIM));
    if (tip) {
        // plane height under the CL point, slope from the facet normal
        const V3 n = vcross(vsub(V[0], V[1]), vsub(V[0], V[2]));
        gx = (T)(-n.x / n.z);
        gy = (T)(-n.y / n.z);
        gr = T(0);
    } else if (rim) {
        const T den = cs0 * cs0 + cs1 * cs1;
        const T lam = den > T(0) ? (fs0 * cs0 + fs1 * cs1) / den : T(0);
        gx = T(2) * lam * dx;
        gy = T(2) * lam * dy;
        gr = (T)(-rim_height_slope(s)) + T(2) * lam * (T)s.radius;
    } else {
        T hp, hr;
        prof_slope<T>(s, q, hp, hr);
        gx = q > T(0) ? hp * dx / q : T(0);
        gy = q > T(0) ? hp * dy / q : T(0);
        gr = -hr;
    }
    vx = -gx;
    vy = -gy;
    o.gx = gx;
    o.gy = gy;
    o.gr = gr;
    const T w[3] = {w0, w1, w2};
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        o.gv[3 * j] = vx * w[j];
        o.gv[3 * j + 1] = vy * w[j];
        o.gv[3 * j + 2] = w[j];
    }
    return o;
}

}  // namespace ocl_amd
