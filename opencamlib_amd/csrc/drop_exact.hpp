// Exact FP64 drop-cutter contact of one CL point against one triangle,
// following the reference operation order (src/cutters/millingcutter.cpp:
// 36-136, 386-409 and the per-cutter overrides).  Used by the GPU for the
// FP64 refine of a selected contact and for escalated points, and by the
// FP64 "exact" precision mode.
#pragma once

#include "cutter.hpp"
#include "geom64.hpp"

namespace ocl_amd {

struct TriX {
    V3 p[3];
    V3 n;      // unit normal as Triangle::calcNormal (triangle.cpp:65-71)
    double maxz;
};

// Running state of one CL point (CLPoint + CCPoint, clpoint.hpp:38-81).
struct DropLane {
    double x, y, z;
    V3 cc;
    int type;
    uint32_t tri;  // rank index of the triangle that made the last lift
    int feat;
    uint32_t cur_tri;
    int cur_feat;
};

OCL_HD bool lane_lift(DropLane& L, double zin, V3 cc, int type) {  // clpoint.cpp:79-99
    if (zin > L.z) {
        L.z = zin;
        L.cc = cc;
        L.type = type;
        L.tri = L.cur_tri;
        L.feat = L.cur_feat;
        return true;
    }
    return false;
}

// vertex k of a triangle without dynamic indexing of a local array (which
// would be placed in scratch memory on the GPU)
OCL_HD double pick3(int k, double a, double b, double c) { return k == 0 ? a : (k == 1 ? b : c); }
OCL_HD V3 tri_vtx(const TriX& t, int k) {
    return v3(pick3(k, t.p[0].x, t.p[1].x, t.p[2].x), pick3(k, t.p[0].y, t.p[1].y, t.p[2].y),
              pick3(k, t.p[0].z, t.p[1].z, t.p[2].z));
}

OCL_HD V3 up_normal(const TriX& t) { return (t.n.z < 0) ? vscale(-1.0, t.n) : t.n; }
OCL_HD bool tri_contains(V3 p, const TriX& t) { return in_triangle_xy(p, t.p[0], t.p[1], t.p[2]); }

// ------------------------------------------------------------- heights
// cylcutter.hpp:52, ballcutter.hpp:51, bullcutter.cpp:54-63, conecutter.cpp:51-54
OCL_HD double prof_height(const CutterP& c, double r) {
    switch (c.kind) {
    case K_CYL: return (r <= c.radius) ? 0.0 : -1.0;
    case K_BALL: return c.radius - sqrt(sqr(c.radius) - sqr(r));
    case K_BULL:
        if (r <= c.r1) return 0.0;
        if (r <= c.radius) return c.r2 - sqrt(sqr(c.r2) - sqr(r - c.r1));
        return -1;
    default: return r / c.tan_a;
    }
}
// compositecutter.cpp:111-122, 75-83, 58-61
OCL_HD int comp_index_of_radius(const CutterDev& cd, double r) {
    for (int n = 0; n < cd.nsub; ++n) {
        const double lo = (n == 0) ? -1E-6 : cd.rlim[n - 1] - 1E-6;
        const double hi = cd.rlim[n] + 1e-6;
        if (lo <= r && r <= hi) return n;
    }
    return 0;
}
OCL_HD double cutter_height(const CutterDev& cd, double r) {
    if (cd.c.kind != K_COMPOSITE) return prof_height(cd.c, r);
    const int i = comp_index_of_radius(cd, r);
    return prof_height(cd.sub[i], r) + cd.zoff[i];
}

// --------------------------------------------------------------- facet
// millingcutter.cpp:51-75
OCL_HD bool facet_generic(const CutterP& c, DropLane& L, const TriX& t) {
    V3 n = up_normal(t);
    if (near0(n.z)) return false;
    if (near0(n.x) && near0(n.y)) {
        const V3 cc = v3(L.x, L.y, t.p[0].z);
        return tri_contains(cc, t) ? lane_lift(L, cc.z, cc, CC_FACET) : false;
    }
    const double d = -vdot(n, t.p[0]);
    n = vunit(n);
    const V3 xyn = vxyunit(v3(n.x, n.y, 0.0));
    const V3 rv = vadd(vscale(c.xyl, xyn), vscale(c.nl, n));
    V3 cc = vsub(v3(L.x, L.y, L.z), rv);
    cc.z = (1.0 / n.z) * (-d - n.x * cc.x - n.y * cc.y);
    const double tip = cc.z + rv.z - c.ch;
    return tri_contains(cc, t) ? lane_lift(L, tip, cc, CC_FACET) : false;
}

// conecutter.cpp:70-105: tip contact first, rim only if the tip did not lift
OCL_HD bool facet_cone(const CutterP& c, DropLane& L, const TriX& t) {
    V3 n = up_normal(t);
    if (near0(n.z)) return false;
    if (near0(n.x) && near0(n.y)) {
        const V3 cc = v3(L.x, L.y, t.p[0].z);
        return tri_contains(cc, t) ? lane_lift(L, cc.z, cc, CC_FACET_TIP) : false;
    }
    const double a = n.x, b = n.y, cz = n.z;
    const double d = -vdot(n, t.p[0]);
    n = vxyunit(n);
    V3 rim = vsub(v3(L.x, L.y, L.z), vscale(c.radius, n));
    rim.z = (1.0 / cz) * (-d - a * rim.x - b * rim.y);
    const double rim_z = rim.z - c.length;
    V3 tip = v3(L.x, L.y, 0.0);
    tip.z = (1.0 / cz) * (-d - a * tip.x - b * tip.y);
    if (tri_contains(tip, t) && lane_lift(L, tip.z, tip, CC_FACET_TIP)) return true;
    const int f0 = L.cur_feat;
    L.cur_feat = f0 + FT_FACET_RIM;
    const bool r = tri_contains(rim, t) ? lane_lift(L, rim_z, rim, CC_FACET_CYL) : false;
    L.cur_feat = f0;
    return r;
}

OCL_HD bool facet_simple(const CutterP& c, DropLane& L, const TriX& t) {
    return (c.kind == K_CONE) ? facet_cone(c, L, t) : facet_generic(c, L, t);
}

// compositecutter.cpp:124-147: xy distance of the sub-cutter cc from the CL
OCL_HD bool comp_cc_in_window(const CutterDev& cd, int n, const DropLane& s) {
    if (s.type == CC_NONE) return false;
    const double dist = vxynorm(vsub(v3(s.x, s.y, s.z), s.cc));
    const double lo = (n == 0) ? -1E-6 : cd.rlim[n - 1] - 1E-6;
    const double hi = cd.rlim[n] + 1e-6;
    return !(dist < lo) && !(dist > hi);
}

OCL_HD DropLane sub_lane(const DropLane& L, double zoff, int feat_base) {
    DropLane s;
    s.x = L.x + 0;
    s.y = L.y + 0;
    s.z = L.z + zoff;
    s.cc = v3(0, 0, 0);
    s.type = CC_NONE;
    s.tri = 0xffffffffu;
    s.feat = FT_NONE;
    s.cur_tri = L.cur_tri;
    s.cur_feat = feat_base;
    return s;
}

// compositecutter.cpp:152-172
OCL_HD bool facet_composite(const CutterDev& cd, DropLane& L, const TriX& t) {
    bool r = false;
    for (int n = 0; n < cd.nsub; ++n) {
        DropLane s = sub_lane(L, cd.zoff[n], (n << 4) | FT_FACET);
        if (facet_simple(cd.sub[n], s, t) && comp_cc_in_window(cd, n, s)) {
            const int f0 = L.cur_feat;
            L.cur_feat = s.feat;
            if (lane_lift(L, s.z - cd.zoff[n], s.cc, CC_FACET)) r = true;
            L.cur_feat = f0;
        }
    }
    return r;
}

// -------------------------------------------------------------- vertex
// millingcutter.cpp:36-47
OCL_HD bool vertex_drop(const CutterDev& cd, DropLane& L, const TriX& t) {
    bool r = false;
    const int f0 = L.cur_feat;
    for (int k = 0; k < 3; ++k) {
        const V3 p = tri_vtx(t, k);
        const double q = vxynorm(vsub(v3(L.x, L.y, L.z), p));
        if (q <= cd.c.radius) {
            L.cur_feat = f0 + FT_VTX0 + k;
            if (lane_lift(L, p.z - cutter_height(cd, q), p, CC_VERTEX)) r = true;
        }
    }
    L.cur_feat = f0;
    return r;
}

// ---------------------------------------------------------------- edge
struct CanonPair { double cc_u, cl_z; };

// cylcutter.cpp:52-70
OCL_HD CanonPair canon_cyl(const CutterP& c, V3 u1, V3 u2) {
    const double s = sqrt(sqr(c.radius) - sqr(u1.y));
    const double z1 = edge_z_at(v3(s, u1.y, 0), u1, u2);
    const double z2 = edge_z_at(v3(-s, u1.y, 0), u1, u2);
    CanonPair r;
    if (z1 > z2) { r.cc_u = s; r.cl_z = z1; }
    else { r.cc_u = -s; r.cl_z = z2; }
    return r;
}
// ballcutter.cpp:50-63
OCL_HD CanonPair canon_ball(const CutterP& c, V3 u1, V3 u2) {
    const double s = sqrt(sqr(c.radius) - sqr(u1.y));
    V3 nrm = vxyunit(v3(u2.z - u1.z, -(u2.x - u1.x), 0));
    if (nrm.y < 0) nrm = vscale(-1, nrm);
    V3 cc = v3(-s * nrm.x, u1.y, 0);
    cc.z = edge_z_at(cc, u1, u2);
    CanonPair r;
    r.cc_u = cc.x;
    r.cl_z = cc.z + s * nrm.y - c.radius;
    return r;
}
// conecutter.cpp:109-135
OCL_HD CanonPair canon_cone(const CutterP& c, V3 u1, V3 u2) {
    const double d = u1.y;
    const double m = (u2.z - u1.z) / (u2.x - u1.x);
    const double xu = sqrt(sqr(c.radius) - sqr(u1.y));
    const double mu = (c.ch / c.radius) * xu / sqrt(sqr(xu) + sqr(d));
    const bool hyper = (fabs(m) <= fabs(mu));
    double ccu;
    if (hyper) ccu = sgn1(m) * sqrt(sqr(c.radius) * sqr(m) * sqr(d) / (sqr(c.length) - sqr(c.radius) * sqr(m)));
    else ccu = sgn1(m) * xu;
    const double ccz = edge_z_at(v3(ccu, d, 0.0), u1, u2);
    CanonPair r;
    r.cc_u = ccu;
    if (hyper) r.cl_z = ccz - c.ch + (c.radius - sqrt(sqr(ccu) + sqr(d))) / c.tan_a;
    else r.cl_z = ccz - c.ch;
    return r;
}

// ---- BullCutter offset-ellipse solve (bullcutter.cpp:82-105, ellipse.cpp)
// Ellipse centred at (0, d) with axes a (along x) and b, offset r1.  The
// error is the y coordinate of the offset-ellipse point (ellipse.cpp:321-324).
struct OffEllipse {
    V3 center;
    double a, b, off;
};
OCL_HD V3 oe_point(const OffEllipse& e, double s, double t) {  // ellipse.cpp:60-107
    V3 p = e.center;
    p.x += e.a * s;
    p.y += e.b * t;
    const V3 n = vunit(v3(e.b * s, e.a * t, 0));
    return vadd(p, vscale(e.off, n));
}
OCL_HD double oe_error_dia(const OffEllipse& e, double dia) {
    double s, t;
    diangle_st(dia, s, t);
    return oe_point(e, s, t).y;
}

// brent_zero.hpp:36-123 with an iteration cap (the reference loop is
// unbounded; converged solves here take < 60 iterations).  `iters` returns
// the count.
OCL_HD double brent_root(const OffEllipse& e, double a, double b, double eps, double tolv, int& iters) {
    double c, d, ee, fa, fb, fc, m, p, q, r, s, tol;
    fa = oe_error_dia(e, a);
    fb = oe_error_dia(e, b);
    c = a;
    fc = fa;
    ee = b - a;
    d = ee;
    iters = 0;
    while (iters < 200) {
        ++iters;
        if (fabs(fc) < fabs(fb)) {
            a = b; b = c; c = a;
            fa = fb; fb = fc; fc = fa;
        }
        tol = 2.0 * eps * fabs(b) + tolv;
        m = 0.5 * (c - b);
        if ((fabs(m) <= tol) || (fb == 0.0)) break;
        if ((fabs(ee) < tol) || (fabs(fa) <= fabs(fb))) {
            ee = m;
            d = ee;
        } else {
            s = fb / fa;
            if (a == c) {
                p = 2.0 * m * s;
                q = 1.0 - s;
            } else {
                q = fa / fc;
                r = fb / fc;
                p = s * (2.0 * m * a * (q - r) - (b - a) * (r - 1.0));
                q = (q - 1.0) * (r - 1.0) * (s - 1.0);
            }
            if (p > 0.0) q = -q;
            else p = -p;
            s = ee;
            ee = d;
            if ((2.0 * p < (3.0 * m * q - fabs(tol * q))) && (p < fabs(0.5 * s * q))) d = p / q;
            else { ee = m; d = ee; }
        }
        a = b;
        fa = fb;
        if (fabs(d) > tol) b = b + d;
        else if (m > 0.0) b = b + tol;
        else b = b - tol;
        fb = oe_error_dia(e, b);
        if (signbit(fb) == signbit(fc)) {
            c = a;
            fc = fa;
            ee = b - a;
            d = ee;
        }
    }
    return b;
}

// ellipse.cpp:120-181: the second (symmetric) solution
OCL_HD void oe_second(const OffEllipse& e, double s1, double t1, double& s2, double& t2) {
    const double err1 = fabs(oe_point(e, s1, t1).y);
    const double sgs[4] = {1, -1, -1, 1}, sgt[4] = {-1, 1, -1, 1};
    for (int k = 0; k < 4; ++k) {
        s2 = sgs[k] * s1;
        t2 = sgt[k] * t1;
        if (fabs(oe_point(e, s2, t2).y) < err1 + 1e-10) {
            if (k == 3) return;
            if ((fabs(s2 - s1) > 1E-8) || (fabs(t2 - t1) > 1E-8)) return;
        }
    }
}

OCL_HD CanonPair canon_bull(const CutterP& c, V3 u1, V3 u2, int& iters) {
    CanonPair r;
    iters = 0;
    if (near0(u1.z - u2.z)) {
        r.cc_u = 0;
        r.cl_z = u1.z - prof_height(c, u1.y);
        return r;
    }
    const double theta = atan((u2.z - u1.z) / (u2.x - u1.x));
    OffEllipse e;
    e.center = v3(0, u1.y, 0);
    e.a = fabs(c.r2 / sin(theta));
    e.b = c.r2;
    e.off = c.r1;
    double s1, t1, s2, t2;
    diangle_st(0.0, s1, t1);
    double sb, tb;
    diangle_st(3.0, sb, tb);
    if (fabs(oe_point(e, s1, t1).y) < 1e-10) {
        iters = 1;
    } else if (fabs(oe_point(e, sb, tb).y) < 1e-10) {
        s1 = sb; t1 = tb;
        iters = 1;
    } else {
        const double dia = brent_root(e, 0.0, 3.0, 3E-16, 1e-10, iters);
        diangle_st(dia, s1, t1);
    }
    oe_second(e, s1, t1, s2, t2);
    // ellipse.cpp:335-356: keep the solution whose ellipse centre is higher
    const V3 cle1 = oe_point(e, s1, t1), cle2 = oe_point(e, s2, t2);
    const double ta = (e.center.x + (-cle1.x) - u1.x) / (u2.x - u1.x);
    const double tb2 = (e.center.x + (-cle2.x) - u1.x) / (u2.x - u1.x);
    const V3 ec1 = vadd(u1, vscale(ta, vsub(u2, u1)));
    const V3 ec2 = vadd(u1, vscale(tb2, vsub(u2, u1)));
    double sh, th;
    V3 cen;
    if (ec1.z >= ec2.z) { sh = s1; th = t1; cen = ec1; }
    else { sh = s2; th = t2; cen = ec2; }
    V3 ep = cen;
    ep.x += e.a * sh;
    ep.y += e.b * th;
    r.cc_u = line_closest(ep, u1, u2).x;
    r.cl_z = cen.z - c.r2;
    return r;
}

// millingcutter.cpp:121-136
OCL_HD bool single_edge_drop(const CutterP& c, DropLane& L, V3 p1, V3 p2, double d, int& iters) {
    const V3 v = vsub(p2, p1);
    const V3 vxy = vxyunit(v3(v.x, v.y, 0.0));
    const V3 sc = xy_line_closest(v3(L.x, L.y, L.z), p1, p2);
    const V3 u1 = v3(vdot(vsub(p1, sc), vxy), d, p1.z);
    const V3 u2 = v3(vdot(vsub(p2, sc), vxy), d, p2.z);
    CanonPair k;
    iters = 0;
    switch (c.kind) {
    case K_CYL: k = canon_cyl(c, u1, u2); break;
    case K_BALL: k = canon_ball(c, u1, u2); break;
    case K_BULL: k = canon_bull(c, u1, u2, iters); break;
    default: k = canon_cone(c, u1, u2); break;
    }
    V3 cc = vadd(sc, vscale(k.cc_u, vxy));
    cc.z = edge_z_at(cc, p1, p2);
    return on_segment(cc, p1, p2) ? lane_lift(L, k.cl_z, cc, CC_EDGE) : false;
}

// millingcutter.cpp:79-94
OCL_HD bool edge_simple(const CutterP& c, DropLane& L, const TriX& t, int& iters) {
    bool r = false;
    const int f0 = L.cur_feat;
    for (int n = 0; n < 3; n++) {
        const V3 p1 = tri_vtx(t, n), p2 = tri_vtx(t, n == 2 ? 0 : n + 1);
        if (!near0(p1.x - p2.x) || !near0(p1.y - p2.y)) {
            const double d = xy_line_distance(v3(L.x, L.y, L.z), p1, p2);
            if (d <= c.radius) {
                L.cur_feat = f0 + FT_EDGE0 + n;
                int it = 0;
                if (single_edge_drop(c, L, p1, p2, d, it)) r = true;
                iters += it;
            }
        }
    }
    L.cur_feat = f0;
    return r;
}

// compositecutter.cpp:175-194
OCL_HD bool edge_composite(const CutterDev& cd, DropLane& L, const TriX& t, int& iters) {
    bool r = false;
    for (int n = 0; n < cd.nsub; ++n) {
        DropLane s = sub_lane(L, cd.zoff[n], n << 4);
        if (edge_simple(cd.sub[n], s, t, iters) && comp_cc_in_window(cd, n, s)) {
            const int f0 = L.cur_feat;
            L.cur_feat = s.feat;
            if (lane_lift(L, s.z - cd.zoff[n], s.cc, CC_EDGE)) r = true;
            L.cur_feat = f0;
        }
    }
    return r;
}

// Exact FP64 contact of ONE feature of a triangle (a selection-pass winner),
// computed as the reference computes that feature inside dropCutter.  The
// lane must start below every candidate (z = -inf).  Not for composites.
OCL_HD void exact_feature(const CutterP& c, DropLane& L, const TriX& t, int feat, int& iters) {
    const int kf = feat & 15;
    L.cur_feat = feat;
    iters = 0;
    if (kf == FT_FACET || kf == FT_FACET_RIM) {
        if (c.kind != K_CONE) {
            facet_generic(c, L, t);
            return;
        }
        // conecutter.cpp:70-105, one of the two contacts
        V3 n = up_normal(t);
        if (near0(n.z)) return;
        if (near0(n.x) && near0(n.y)) {
            if (kf != FT_FACET) return;
            const V3 cc = v3(L.x, L.y, t.p[0].z);
            if (tri_contains(cc, t)) lane_lift(L, cc.z, cc, CC_FACET_TIP);
            return;
        }
        const double a = n.x, b = n.y, cz = n.z;
        const double d = -vdot(n, t.p[0]);
        n = vxyunit(n);
        if (kf == FT_FACET) {
            V3 tip = v3(L.x, L.y, 0.0);
            tip.z = (1.0 / cz) * (-d - a * tip.x - b * tip.y);
            if (tri_contains(tip, t)) lane_lift(L, tip.z, tip, CC_FACET_TIP);
        } else {
            V3 rim = vsub(v3(L.x, L.y, L.z), vscale(c.radius, n));
            rim.z = (1.0 / cz) * (-d - a * rim.x - b * rim.y);
            if (tri_contains(rim, t)) lane_lift(L, rim.z - c.length, rim, CC_FACET_CYL);
        }
        return;
    }
    if (kf < FT_EDGE0) {
        const V3 p = tri_vtx(t, kf - FT_VTX0);
        const double q = vxynorm(vsub(v3(L.x, L.y, L.z), p));
        if (q <= c.radius) lane_lift(L, p.z - prof_height(c, q), p, CC_VERTEX);
        return;
    }
    const int n = kf - FT_EDGE0;
    const V3 p1 = tri_vtx(t, n), p2 = tri_vtx(t, n == 2 ? 0 : n + 1);
    if (!near0(p1.x - p2.x) || !near0(p1.y - p2.y)) {
        const double d = xy_line_distance(v3(L.x, L.y, L.z), p1, p2);
        if (d <= c.radius) single_edge_drop(c, L, p1, p2, d, iters);
    }
}

// Exact FP64 contact of ONE feature of a composite cutter's triangle (a
// selection-pass winner), as the reference computes it inside its composite
// drop (compositecutter.cpp:152-194, 58-61): the facet of sub-cutter n
// (feature bits 4-5) with its window test, a vertex with the composite
// profile, or edge j of sub-cutter n alone with its window test (the
// selection found it to be that sub-cutter's highest edge contact).  The
// lane must start below every candidate (z = -inf).
OCL_HD void exact_feature_comp(const CutterDev& cd, DropLane& L, const TriX& t, int feat, int& iters) {
    const int kf = feat & 15, n = (feat >> 4) & 3;
    iters = 0;
    L.cur_feat = feat;
    if (kf == FT_FACET) {
        DropLane s = sub_lane(L, cd.zoff[n], feat);
        if (facet_simple(cd.sub[n], s, t) && comp_cc_in_window(cd, n, s)) lane_lift(L, s.z - cd.zoff[n], s.cc, CC_FACET);
        return;
    }
    if (kf < FT_EDGE0) {
        const V3 p = tri_vtx(t, kf - FT_VTX0);
        const double q = vxynorm(vsub(v3(L.x, L.y, L.z), p));
        if (q <= cd.c.radius) lane_lift(L, p.z - cutter_height(cd, q), p, CC_VERTEX);
        return;
    }
    const int j = kf - FT_EDGE0;
    const V3 p1 = tri_vtx(t, j), p2 = tri_vtx(t, j == 2 ? 0 : j + 1);
    if (!near0(p1.x - p2.x) || !near0(p1.y - p2.y)) {
        const double d = xy_line_distance(v3(L.x, L.y, L.z), p1, p2);
        if (d <= cd.sub[n].radius) {
            DropLane s = sub_lane(L, cd.zoff[n], feat);
            if (single_edge_drop(cd.sub[n], s, p1, p2, d, iters) && comp_cc_in_window(cd, n, s))
                lane_lift(L, s.z - cd.zoff[n], s.cc, CC_EDGE);
        }
    }
}

// millingcutter.cpp:422-433 (XY box of the cutter vs triangle bbox)
OCL_HD bool cutter_overlaps(double radius, double x, double y, const TriX& t) {
    double mnx = t.p[0].x, mxx = t.p[0].x, mny = t.p[0].y, mxy = t.p[0].y;
    if (t.p[1].x > mxx) mxx = t.p[1].x;
    if (t.p[1].x < mnx) mnx = t.p[1].x;
    if (t.p[1].y > mxy) mxy = t.p[1].y;
    if (t.p[1].y < mny) mny = t.p[1].y;
    if (t.p[2].x > mxx) mxx = t.p[2].x;
    if (t.p[2].x < mnx) mnx = t.p[2].x;
    if (t.p[2].y > mxy) mxy = t.p[2].y;
    if (t.p[2].y < mny) mny = t.p[2].y;
    if (mxx < x - radius) return false;
    if (mnx > x + radius) return false;
    if (mxy < y - radius) return false;
    if (mny > y + radius) return false;
    return true;
}

// batchdropcutter.cpp:214-220 + millingcutter.cpp:386-409: one candidate
// triangle of the KD search applied to the running CL state.  Returns the
// Brent iteration count spent (Bull edges).
OCL_HD int drop_triangle_exact(const CutterDev& cd, DropLane& L, const TriX& t) {
    int iters = 0;
    if (!cutter_overlaps(cd.c.radius, L.x, L.y, t)) return 0;
    if (!(L.z < t.maxz)) return 0;
    L.cur_feat = 0;
    const bool facet = (cd.c.kind == K_COMPOSITE) ? facet_composite(cd, L, t) : facet_simple(cd.c, L, t);
    if (!facet) {
        vertex_drop(cd, L, t);
        if (L.z < t.maxz) {
            if (cd.c.kind == K_COMPOSITE) edge_composite(cd, L, t, iters);
            else edge_simple(cd.c, L, t, iters);
        }
    }
    return iters;
}

}  // namespace ocl_amd
