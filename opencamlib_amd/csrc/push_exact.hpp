// Reference-order FP64 push-cutter for one (fiber, triangle) pair.
//
// Restates MillingCutter::pushCutter (src/cutters/millingcutter.cpp:139-382)
// and the cutter overrides it dispatches to (cylcutter.cpp:73-91,
// ballcutter.cpp:66-134, bullcutter.cpp:109-155, conecutter.cpp:137-320,
// compositecutter.cpp:196-270) with the Interval update rules of
// src/algo/interval.cpp:46-118.  Operation order follows the reference so
// that the device results match its x86-64 FP64 arithmetic (all code is
// compiled with -ffp-contract=off); only the Bull edge solve differs, by the
// device atan/sin and the Brent iteration cap.
#pragma once

#include "cutter.hpp"
#include "drop_exact.hpp"
#include "geom64.hpp"

namespace ocl_amd {

// Interval of one pair (interval.cpp:31-118): lower/upper fiber parameter,
// the CC point and CC type at each end.  "Empty" = never updated, or both
// ends exactly 0 (Interval::empty).
struct PIval {
    double lo, hi;
    V3 lcc, ucc;
    int lty, uty;
};

OCL_HD void piv_init(PIval& i) {
    i.lo = 0.0;
    i.hi = 0.0;
    i.lcc = v3(0, 0, 0);
    i.ucc = v3(0, 0, 0);
    i.lty = CC_NONE;
    i.uty = CC_NONE;
}
OCL_HD bool piv_empty(const PIval& i) { return i.lo == 0.0 && i.hi == 0.0; }

// interval.cpp:61-76
OCL_HD void piv_upper(PIval& i, double t, V3 p, int type) {
    if (i.uty == CC_NONE) {
        i.hi = t;
        i.lo = t;
        i.ucc = p;
        i.uty = type;
        i.lcc = p;
        i.lty = type;
    }
    if (t > i.hi) {
        i.hi = t;
        i.ucc = p;
        i.uty = type;
    }
}
// interval.cpp:78-93
OCL_HD void piv_lower(PIval& i, double t, V3 p, int type) {
    if (i.lty == CC_NONE) {
        i.lo = t;
        i.hi = t;
        i.lcc = p;
        i.lty = type;
        i.ucc = p;
        i.uty = type;
    }
    if (t < i.lo) {
        i.lo = t;
        i.lcc = p;
        i.lty = type;
    }
}
// interval.cpp:46-49
OCL_HD void piv_update(PIval& i, double t, V3 p, int type) {
    piv_upper(i, t, p, type);
    piv_lower(i, t, p, type);
}
// interval.cpp:51-59 (CC inside the edge segment, 3-D parameter test)
OCL_HD bool piv_update_if(PIval& i, double t, V3 cc, int type, V3 p1, V3 p2, bool cond) {
    if (on_segment(cc, p1, p2) && cond) {
        piv_update(i, t, cc, type);
        return true;
    }
    return false;
}

// Fiber (fiber.cpp:29-106): p1, p2 at one z, dir = normalized p2 - p1.
struct FibX {
    V3 p1, p2, dir;
};
OCL_HD FibX make_fiber(const double* q) {
    FibX f;
    f.p1 = v3(q[0], q[1], q[2]);
    f.p2 = v3(q[3], q[4], q[5]);
    f.dir = vunit(vsub(f.p2, f.p1));
    return f;
}
OCL_HD double fib_tval(const FibX& f, V3 p) {
    const V3 d = vsub(f.p2, f.p1);
    return vdot(vsub(p, f.p1), d) / vdot(d, d);
}
OCL_HD V3 fib_point(const FibX& f, double t) { return vadd(f.p1, vscale(t, vsub(f.p2, f.p1))); }

// width(h): cylcutter.hpp:53, ballcutter.hpp:52, bullcutter.cpp:66-68, conecutter.cpp:56-60
OCL_HD double prof_width(const CutterP& c, double h) {
    switch (c.kind) {
    case K_CYL: return c.radius;
    case K_BALL: return (h >= c.radius) ? c.radius : sqrt(sqr(c.radius) - sqr(c.radius - h));
    case K_BULL: return (h >= c.r2) ? c.radius : c.r1 + sqrt(sqr(c.r2) - sqr(c.r2 - h));
    default: return (h < c.ch) ? h * c.tan_a : c.radius;
    }
}
// compositecutter.cpp:97-108, 64-69
OCL_HD int comp_index_of_height(const CutterDev& cd, double h) {
    for (int n = 0; n < cd.nsub; ++n) {
        const double lo = (n == 0) ? -1E-6 : cd.hlim[n - 1] - 1E-6;
        const double hi = cd.hlim[n] + 1e-6;
        if (lo <= h && h <= hi) return n;
    }
    return cd.nsub - 1;
}

// millingcutter.cpp:148-167
OCL_HD bool single_vertex_push(const CutterP& c, const FibX& f, PIval& i, V3 p, int type) {
    if ((p.z >= f.p1.z) && (p.z <= (f.p1.z + c.length))) {
        const V3 pq = xy_line_closest(p, f.p1, f.p2);
        const double q = vxynorm(vsub(p, pq));
        const double h = p.z - f.p1.z;
        const double cw = prof_width(c, h);
        if (q <= cw) {
            const double ofs = sqrt(sqr(cw) - sqr(q));
            const V3 start = vsub(pq, vscale(ofs, f.dir));
            const V3 stop = vadd(pq, vscale(ofs, f.dir));
            piv_upper(i, fib_tval(f, stop), p, type);
            piv_lower(i, fib_tval(f, start), p, type);
            return true;
        }
    }
    return false;
}

// Triangle::zslice_verts (triangle.cpp:77-128)
OCL_HD bool zslice_verts(const TriX& t, double zcut, V3& q1, V3& q2) {
    const double zmin = fmin(fmin(t.p[0].z, t.p[1].z), t.p[2].z);
    if ((zcut <= zmin) || (zcut >= t.maxz)) return false;
    // vertices below/above the cut, in vertex order (named registers, no
    // dynamically indexed arrays)
    V3 b0 = t.p[0], b1 = t.p[0], a0 = t.p[0], a1 = t.p[0];
    int nb = 0, na = 0;
    for (int m = 0; m < 3; ++m) {
        const V3 p = tri_vtx(t, m);
        if (p.z <= zcut) {
            if (nb == 0) b0 = p;
            else b1 = p;
            ++nb;
        } else {
            if (na == 0) a0 = p;
            else a1 = p;
            ++na;
        }
    }
    if (nb == 2) {
        const double t1 = (zcut - a0.z) / (b0.z - a0.z);
        const double t2 = (zcut - a0.z) / (b1.z - a0.z);
        q1 = vadd(a0, vscale(t1, vsub(b0, a0)));
        q2 = vadd(a0, vscale(t2, vsub(b1, a0)));
        return true;
    } else if (nb == 1) {
        const double t1 = (zcut - a0.z) / (b0.z - a0.z);
        const double t2 = (zcut - a1.z) / (b0.z - a1.z);
        q1 = vadd(a0, vscale(t1, vsub(b0, a0)));
        q2 = vadd(a1, vscale(t2, vsub(b0, a1)));
        return true;
    }
    return false;
}

// millingcutter.cpp:139-146 and CylCutter::vertexPush (cylcutter.cpp:73-91)
OCL_HD bool base_vertex_push(const CutterP& c, const FibX& f, PIval& i, const TriX& t) {
    bool r = false;
    for (int k = 0; k < 3; ++k)
        if (single_vertex_push(c, f, i, tri_vtx(t, k), CC_VERTEX)) r = true;
    if (c.kind == K_CYL) {
        V3 a, b;
        if (zslice_verts(t, f.p1.z, a, b)) {
            a.z = f.p1.z;
            b.z = f.p1.z;
            if (single_vertex_push(c, f, i, a, CC_VERTEX_CYL)) r = true;
            if (single_vertex_push(c, f, i, b, CC_VERTEX_CYL)) r = true;
        }
    }
    return r;
}

// millingcutter.cpp:177-275 (exact zero test of the normal: point.cpp:316-323)
OCL_HD bool general_facet_push(double nl, double ch, double xyl, const FibX& fib, PIval& i, const TriX& t) {
    V3 normal = up_normal(t);
    if (normal.x == 0.0 && normal.y == 0.0) return false;
    normal = vunit(normal);
    V3 xyn = normal;
    xyn.z = 0;
    xyn = vxyunit(xyn);
    const V3 p0 = t.p[0], p1 = t.p[1], p2 = t.p[2];
    double a, b, e, u = 0, v = 0;
    const double c = p1.z - p0.z;
    const double d = p2.z - p0.z;
    const double f = -p0.z - nl * normal.z + fib.p1.z + ch;
    if (fib.p1.y == fib.p2.y) {
        a = p1.y - p0.y;
        b = p2.y - p0.y;
        e = -p0.y - nl * normal.y - xyl * xyn.y + fib.p1.y;
        if (!solve2x2(a, b, c, d, e, f, u, v)) return false;
        const V3 cc = vadd(vadd(p0, vscale(u, vsub(p1, p0))), vscale(v, vsub(p2, p0)));
        if (!tri_contains(cc, t)) return false;
        const double tval = (1.0 / (fib.p2.x - fib.p1.x)) *
                            (p0.x + nl * normal.x + xyl * xyn.x - fib.p1.x + u * (p1.x - p0.x) + v * (p2.x - p0.x));
        piv_update(i, tval, cc, CC_FACET);
        return true;
    } else if (fib.p1.x == fib.p2.x) {
        a = p1.x - p0.x;
        b = p2.x - p0.x;
        e = -p0.x - nl * normal.x - xyl * xyn.x + fib.p1.x;
        if (!solve2x2(a, b, c, d, e, f, u, v)) return false;
        const V3 cc = vadd(vadd(p0, vscale(u, vsub(p1, p0))), vscale(v, vsub(p2, p0)));
        if (!tri_contains(cc, t)) return false;
        const double tval = (1.0 / (fib.p2.y - fib.p1.y)) *
                            (p0.y + nl * normal.y + xyl * xyn.y - fib.p1.y + u * (p1.y - p0.y) + v * (p2.y - p0.y));
        piv_update(i, tval, cc, CC_FACET);
        return true;
    }
    return false;
}

// millingcutter.cpp:370-375
OCL_HD bool calc_cc_update(double t, double u, V3 q, V3 p1, V3 p2, PIval& i, double height, int type) {
    const V3 cc = vadd(q, vscale(u, vsub(p2, p1)));
    return piv_update_if(i, t, cc, type, p1, p2, cc.z >= height);
}

// millingcutter.cpp:335-368
OCL_HD bool horiz_edge_push(const CutterP& c, const FibX& f, PIval& i, V3 p1, V3 p2) {
    bool r = false;
    const double h = p1.z - f.p1.z;
    if (h > 0.0) {
        if (near0(p2.z - p1.z)) {
            const double eff = prof_width(c, h);
            double qt = 0, qv = 0;
            if (xy_lines_cross(p1, p2, qv, f.p1, f.p2, qt)) {
                const V3 q = vadd(p1, vscale(qv, vsub(p2, p1)));
                double ccv = 0, clt = 0;
                V3 tang = vsub(p2, p1);
                tang.z = 0;
                tang = vxyunit(tang);
                const V3 nrm = vperp(tang);
                const V3 q1 = vadd(q, vscale(eff, nrm));
                const V3 q2 = vadd(q1, vsub(p2, p1));
                if (xy_lines_cross(q1, q2, ccv, f.p1, f.p2, clt)) {
                    const double t1 = clt;
                    const double t2 = qt + (qt - clt);
                    if (calc_cc_update(t1, ccv, q, p1, p2, i, f.p1.z, CC_EDGE_HORIZ)) r = true;
                    if (calc_cc_update(t2, -ccv, q, p1, p2, i, f.p1.z, CC_EDGE_HORIZ)) r = true;
                }
            }
        }
    }
    return r;
}

// millingcutter.cpp:304-332
OCL_HD bool shaft_edge_push(const CutterP& c, const FibX& f, PIval& i, V3 p1, V3 p2) {
    bool r = false;
    double u = 0, v = 0;
    if (xy_lines_cross(p1, p2, u, f.p1, f.p2, v)) {
        const V3 q = vadd(p1, vscale(u, vsub(p2, p1)));
        V3 tang = vsub(p2, p1);
        tang.z = 0;
        tang = vxyunit(tang);
        const V3 nrm = vperp(tang);
        const V3 q1 = vadd(q, vscale(c.radius, nrm));
        const V3 q2 = vadd(q1, vsub(p2, p1));
        double u_cc = 0, t_cl = 0;
        if (xy_lines_cross(q1, q2, u_cc, f.p1, f.p2, t_cl)) {
            const double t1 = t_cl;
            const double t2 = v + (v - t_cl);
            if (calc_cc_update(t1, u_cc, q, p1, p2, i, f.p1.z + c.ch, CC_EDGE_SHAFT)) r = true;
            if (calc_cc_update(t2, -u_cc, q, p1, p2, i, f.p1.z + c.ch, CC_EDGE_SHAFT)) r = true;
        }
    }
    return r;
}

// ballcutter.cpp:127-134
OCL_HD bool ball_cc_update(const CutterP& c, double t, V3 p1, V3 p2, const FibX& f, PIval& i) {
    const V3 center = vadd(fib_point(f, t), v3(0, 0, c.radius));
    const V3 cc = line_closest(center, p1, p2);
    return piv_update_if(i, t, cc, CC_EDGE_BALL, p1, p2, (center.z - cc.z >= 0));
}

// ballcutter.cpp:66-124: the fiber-line vs cylinder-around-the-edge quadratic
OCL_HD bool ball_general_edge_push(const CutterP& c, const FibX& f, PIval& i, V3 p1, V3 p2) {
    bool r = false;
    if (near0(vxynorm(vsub(p2, p1)))) return false;
    const V3 ufp1 = vadd(f.p1, v3(0, 0, c.radius));
    const V3 ufp2 = vadd(f.p2, v3(0, 0, c.radius));
    const V3 ab = vsub(p2, p1);
    const V3 ao = vsub(ufp1, p1);
    const V3 ao_x_ab = vcross(ao, ab);
    const V3 v_x_ab = vcross(vsub(ufp2, ufp1), ab);
    const double ab2 = vdot(ab, ab);
    const double a = vdot(v_x_ab, v_x_ab);
    const double b = 2 * (vdot(v_x_ab, ao_x_ab));
    const double cq = vdot(ao_x_ab, ao_x_ab) - (c.radius * c.radius * ab2);
    const double discr = b * b - 4 * a * cq;
    if (near0(discr)) {
        const double t1 = -b / (2 * a);
        if (ball_cc_update(c, t1, p1, p2, f, i)) r = true;
    } else if (discr > 0.0) {
        const double t1 = (-b + sqrt(discr)) / (2 * a);
        const double t2 = (-b - sqrt(discr)) / (2 * a);
        if (ball_cc_update(c, t1, p1, p2, f, i)) r = true;
        if (ball_cc_update(c, t2, p1, p2, f, i)) r = true;
    }
    return r;
}

// AlignedEllipse (ellipse.cpp:48-57, 88-114, 212-319): ellipse in the plane
// of the fiber, axes along the edge's XY direction and its perpendicular.
struct AEll {
    V3 center, major, minor, target, edir;
    double a, b, off;
};
OCL_HD V3 aell_epoint(const AEll& e, double s, double t) {
    return vadd(vadd(e.center, vscale(e.a * s, e.major)), vscale(e.b * t, e.minor));
}
OCL_HD V3 aell_normal(const AEll& e, double s, double t) {
    return vunit(vadd(vscale(s * e.b, e.major), vscale(t * e.a, e.minor)));
}
OCL_HD V3 aell_oepoint(const AEll& e, double s, double t) {
    return vadd(aell_epoint(e, s, t), vscale(e.off, aell_normal(e, s, t)));
}
OCL_HD double aell_error(const AEll& e, double dia) {
    double s, t;
    diangle_st(dia, s, t);
    return vdot(vsub(e.target, aell_oepoint(e, s, t)), e.edir);
}

// brent_zero.hpp:36-123 on the aligned-ellipse error, capped at 200 steps
OCL_HD double aell_brent(const AEll& el, double a, double b, double eps, double tolv, int& iters) {
    double c, d, e, fa, fb, fc, m, p, q, r, s, tol;
    fa = aell_error(el, a);
    fb = aell_error(el, b);
    c = a;
    fc = fa;
    e = b - a;
    d = e;
    int it = 0;
    while (it < 200) {
        ++it;
        if (fabs(fc) < fabs(fb)) {
            a = b;
            b = c;
            c = a;
            fa = fb;
            fb = fc;
            fc = fa;
        }
        tol = 2.0 * eps * fabs(b) + tolv;
        m = 0.5 * (c - b);
        if ((fabs(m) <= tol) || (fb == 0.0)) break;
        if ((fabs(e) < tol) || (fabs(fa) <= fabs(fb))) {
            e = m;
            d = e;
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
            s = e;
            e = d;
            if ((2.0 * p < (3.0 * m * q - fabs(tol * q))) && (p < fabs(0.5 * s * q))) {
                d = p / q;
            } else {
                e = m;
                d = e;
            }
        }
        a = b;
        fa = fb;
        if (fabs(d) > tol) b = b + d;
        else if (m > 0.0) b = b + tol;
        else b = b - tol;
        fb = aell_error(el, b);
        if (signbit(fb) == signbit(fc)) {
            c = a;
            fc = fa;
            e = b - a;
            d = e;
        }
    }
    iters += it;
    return b;
}

// aell_brent as a state machine that asks for one error evaluation per
// step, so that a wavefront can keep every lane on some solve: a lane whose
// solve ends takes the next one while the others iterate.  The arithmetic
// is aell_brent's, in the same order, so the roots are the same.
struct BrentState {
    double a, b, c, d, e, fa, fb, fc;
    int it, phase;  // phase 0: f(a) pending, 1: f(b) pending, 2: f(b) of a step pending
};
OCL_HD void brent_start(BrentState& B, double a, double b) {
    B.a = a;
    B.b = b;
    B.it = 0;
    B.phase = 0;
}
// the point whose error the machine needs next
OCL_HD double brent_x(const BrentState& B) { return B.phase == 0 ? B.a : B.b; }
// consumes f(brent_x); true when the solve has ended (root B.b, B.it steps)
OCL_HD bool brent_step(BrentState& B, double fx, double eps, double tolv) {
    if (B.phase == 0) {
        B.fa = fx;
        B.phase = 1;
        return false;
    }
    B.fb = fx;
    if (B.phase == 1) {
        B.c = B.a;
        B.fc = B.fa;
        B.e = B.b - B.a;
        B.d = B.e;
    } else if (signbit(B.fb) == signbit(B.fc)) {
        B.c = B.a;
        B.fc = B.fa;
        B.e = B.b - B.a;
        B.d = B.e;
    }
    B.phase = 2;
    if (B.it >= 200) return true;
    ++B.it;
    if (fabs(B.fc) < fabs(B.fb)) {
        B.a = B.b;
        B.b = B.c;
        B.c = B.a;
        B.fa = B.fb;
        B.fb = B.fc;
        B.fc = B.fa;
    }
    const double tol = 2.0 * eps * fabs(B.b) + tolv;
    const double m = 0.5 * (B.c - B.b);
    if ((fabs(m) <= tol) || (B.fb == 0.0)) return true;
    if ((fabs(B.e) < tol) || (fabs(B.fa) <= fabs(B.fb))) {
        B.e = m;
        B.d = B.e;
    } else {
        double p, q, r;
        double s = B.fb / B.fa;
        if (B.a == B.c) {
            p = 2.0 * m * s;
            q = 1.0 - s;
        } else {
            q = B.fa / B.fc;
            r = B.fb / B.fc;
            p = s * (2.0 * m * B.a * (q - r) - (B.b - B.a) * (r - 1.0));
            q = (q - 1.0) * (r - 1.0) * (s - 1.0);
        }
        if (p > 0.0) q = -q;
        else p = -p;
        s = B.e;
        B.e = B.d;
        if ((2.0 * p < (3.0 * m * q - fabs(tol * q))) && (p < fabs(0.5 * s * q))) {
            B.d = p / q;
        } else {
            B.e = m;
            B.d = B.e;
        }
    }
    B.a = B.b;
    B.fa = B.fb;
    if (fabs(B.d) > tol) B.b = B.b + B.d;
    else if (m > 0.0) B.b = B.b + tol;
    else B.b = B.b - tol;
    return false;
}

// bullcutter.cpp:109-155 with AlignedEllipse::aligned_solver (ellipse.cpp:212-311),
// in three steps so that a wavefront can run the Brent solves of many
// (fiber, triangle, edge) triples side by side:
//   bull_edge_prep:    the offset ellipse of the edge in the plane
//                      z = fiber z + r2 (false: no contact can reach the
//                      edge segment, or a degenerate edge);
//   bull_edge_bracket: the diamond-angle bracket of the two roots (false:
//                      the error has no sign change, no contact);
//   aell_brent twice:  d1 on (lolim, hilim), d2 on (hilim - 4, lolim);
//   bull_edge_finish:  the two contacts and the interval updates.
OCL_HD bool bull_edge_prep(const CutterP& c, const FibX& f, V3 p1, V3 p2, AEll& e) {
    if (near0(vxynorm(vsub(p2, p1)))) return false;
    if (near0(p2.z - p1.z)) return false;
    const double tplane = (f.p1.z + c.r2 - p1.z) / (p2.z - p1.z);
    {
        // A contact is kept only if its CC, the projection of the ellipse
        // point onto the edge line, lies on the segment (t in [0, 1],
        // update_ifCCinEdgeAndTrue).  The ellipse lies in the plane
        // z = f.z + r2, centred at parameter tplane, with semi-axis
        // r2 / |sin theta| along the edge's XY direction and r2 across it,
        // so the projection stays within r2 |e_xy| / (|e_z| |e|) of tplane.
        // Beyond that (plus a rounding margin) no root can update the
        // interval, and the two Brent solves are skipped.
        const V3 ed = vsub(p2, p1);
        const double reach = c.r2 * vxynorm(ed) / (fabs(ed.z) * vnorm(ed));
        const double margin = 1e-9 * (1.0 + fabs(tplane) + reach);
        if (tplane + reach < -margin || tplane - reach > 1.0 + margin) return false;
    }
    e.center = vadd(p1, vscale(tplane, vsub(p2, p1)));
    V3 major = vsub(p2, p1);
    major.z = 0;
    e.major = vxyunit(major);
    e.minor = vperp(e.major);
    const double theta = atan((p2.z - p1.z) / vxynorm(vsub(p2, p1)));
    e.a = fabs(c.r2 / sin(theta));
    e.b = c.r2;
    e.off = c.r1;
    // aligned_solver
    e.edir = vperp(f.dir);
    e.target = f.p1;
    return true;
}

OCL_HD bool bull_edge_bracket(const AEll& e, const FibX& f, double& lolim, double& hilim) {
    double t1 = 0.0;
    if (f.p1.y == f.p2.y)
        t1 = sqrt(sqr(e.b * e.minor.y) / (sqr(e.a * e.major.y) + sqr(e.b * e.minor.y)));
    else if (f.p1.x == f.p2.x)
        t1 = sqrt(sqr(e.b * e.minor.x) / (sqr(e.a * e.major.x) + sqr(e.b * e.minor.x)));
    const double s1 = sqrt(1.0 - sqr(t1));
    bool fpos = false, fneg = false;
    double adia = 0.0, bdia = 0.0;
    for (int k = 0; k < 4; ++k) {
        const double sx = (k < 2) ? 1.0 : -1.0, ty = (k % 2 == 0) ? 1.0 : -1.0;
        const double dia = diangle_of(sx * s1, ty * t1);
        const double err = aell_error(e, dia);
        if (err > 0) {
            fpos = true;
            adia = dia;
        } else if (err < 0) {
            fneg = true;
            bdia = dia;
        }
    }
    if (!(fpos && fneg)) return false;
    lolim = 0.0;
    hilim = 0.0;
    if (adia > bdia) {
        lolim = bdia;
        hilim = adia;
    } else if (bdia > adia) {
        hilim = bdia;
        lolim = adia;
    }
    return true;
}

// root `which` (0: d1, 1: d2) of the bracket
OCL_HD double bull_edge_solve(const AEll& e, double lolim, double hilim, int which, int& iters) {
    return which == 0 ? aell_brent(e, lolim, hilim, 3E-16, 1e-10, iters)
                      : aell_brent(e, hilim - 4.0, lolim, 3E-16, 1e-10, iters);
}

OCL_HD bool bull_edge_finish(const CutterP& c, const FibX& f, PIval& i, V3 p1, V3 p2, const AEll& e, double d1,
                             double d2) {
    bool r = false;
    double s_1, t_1, s_2, t_2;
    diangle_st(d1, s_1, t_1);
    diangle_st(d2, s_2, t_2);
    const V3 cc = line_closest(aell_epoint(e, s_1, t_1), p1, p2);
    const V3 cc2 = line_closest(aell_epoint(e, s_2, t_2), p1, p2);
    const V3 cl = vsub(aell_oepoint(e, s_1, t_1), v3(0, 0, c.ch));
    const V3 cl2 = vsub(aell_oepoint(e, s_2, t_2), v3(0, 0, c.ch));
    if (piv_update_if(i, fib_tval(f, cl), cc, CC_EDGE_POS, p1, p2, true)) r = true;
    if (piv_update_if(i, fib_tval(f, cl2), cc2, CC_EDGE_POS, p1, p2, true)) r = true;
    return r;
}

OCL_HD bool bull_general_edge_push(const CutterP& c, const FibX& f, PIval& i, V3 p1, V3 p2, int& iters) {
    AEll e;
    double lolim, hilim;
    if (!bull_edge_prep(c, f, p1, p2, e) || !bull_edge_bracket(e, f, lolim, hilim)) return false;
    const double d1 = bull_edge_solve(e, lolim, hilim, 0, iters);
    const double d2 = bull_edge_solve(e, lolim, hilim, 1, iters);
    return bull_edge_finish(c, f, i, p1, p2, e, d1, d2);
}

// conecutter.cpp:300-308
OCL_HD bool cone_circle_cc(const CutterP& c, double t, V3 p1, V3 p2, const FibX& f, PIval& i) {
    const double t_cc = (f.p1.z + c.ch - p1.z) / (p2.z - p1.z);
    const V3 cc = vadd(p1, vscale(t_cc, vsub(p2, p1)));
    return piv_update_if(i, t, cc, CC_EDGE_CONE_BASE, p1, p2, true);
}
// conecutter.cpp:279-295
OCL_HD bool cone_cone_cc(V3 tang, V3 tip, V3 base, V3 p1, V3 p2, const FibX& f, PIval& i) {
    double u = 0, t = 0;
    if (xy_lines_cross(f.p1, f.p2, u, tang, tip, t)) {
        if ((t >= 0.0) && (t <= 1.0)) {
            V3 cc = vadd(base, vscale(t, vsub(tip, base)));
            cc.z = edge_z_at(cc, p1, p2);
            return piv_update_if(i, u, cc, CC_EDGE_CONE, p1, p2, true);
        }
    }
    return false;
}
// conecutter.cpp:150-275 (exact discr == 0.0 test as in the reference)
OCL_HD bool cone_general_edge_push(const CutterP& c, const FibX& f, PIval& i, V3 p1, V3 p2) {
    bool r = false;
    if (near0(p2.z - p1.z)) return false;
    const double t_tip = (f.p1.z - p1.z) / (p2.z - p1.z);
    const V3 p_tip = vadd(p1, vscale(t_tip, vsub(p2, p1)));
    const double t_base = (f.p1.z + c.ch - p1.z) / (p2.z - p1.z);
    V3 p_base = vadd(p1, vscale(t_base, vsub(p2, p1)));
    p_base.z = f.p1.z;
    const double L = vxynorm(vsub(p_base, p_tip));
    const double d = xy_line_distance(p_base, f.p1, f.p2);
    if (d <= c.radius) {
        const double dx = f.p2.x - f.p1.x;
        const double dy = f.p2.y - f.p1.y;
        const double dr = sqrt(sqr(dx) + sqr(dy));
        const double det = (f.p1.x - p_base.x) * (f.p2.y - p_base.y) - (f.p2.x - p_base.x) * (f.p1.y - p_base.y);
        const double discr = sqr(c.radius) * sqr(dr) - sqr(det);
        if (discr >= 0.0) {
            if (discr == 0.0) {
                const double xt = (det * dy) / sqr(dr);
                const double yt = -(det * dx) / sqr(dr);
                const V3 pt = v3(xt + p_base.x, yt + p_base.y, 0);
                if (cone_circle_cc(c, fib_tval(f, pt), p1, p2, f, i)) r = true;
            } else {
                const double xp = (det * dy + sgn1(dy) * dx * sqrt(discr)) / sqr(dr);
                const double yp = (-det * dx + fabs(dy) * sqrt(discr)) / sqr(dr);
                const V3 clp = v3(xp + p_base.x, yp + p_base.y, 0);
                const double tp = fib_tval(f, clp);
                const double xn = (det * dy - sgn1(dy) * dx * sqrt(discr)) / sqr(dr);
                const double yn = (-det * dx - fabs(dy) * sqrt(discr)) / sqr(dr);
                const V3 cln = v3(xn + p_base.x, yn + p_base.y, 0);
                const double tn = fib_tval(f, cln);
                if (cone_circle_cc(c, tp, p1, p2, f, i)) r = true;
                if (cone_circle_cc(c, tn, p1, p2, f, i)) r = true;
            }
        }
    }
    if (L > c.radius) {
        V3 p_mid = vscale(0.5, vadd(p_base, p_tip));
        p_mid.z = f.p1.z;
        const double r_tang = L / 2;
        V3 pd = vsub(p_mid, p_base);
        pd.z = 0;
        const double dist = vxynorm(pd);
        const bool case1 = (near0(dist) && near0(fabs(c.radius - r_tang)));
        const bool case2 = (dist > (c.radius + r_tang));
        const bool case3 = (dist < fabs(c.radius - r_tang));
        const bool case4 = (near0(dist - (c.radius + r_tang)));
        if (!(case1 || case2 || case3 || case4)) {
            const double a = (-sqr(r_tang) + sqr(c.radius) + sqr(dist)) / (2.0 * dist);
            const V3 v2 = vadd(p_base, vscale(a / dist, pd));
            const double h = sqrt(sqr(c.radius) - sqr(a));
            const V3 ofs = v3(-pd.y * (h / dist), pd.x * (h / dist), 0);
            const V3 tang1 = vadd(v2, ofs);
            const V3 tang2 = vsub(v2, ofs);
            if (cone_cone_cc(tang1, p_tip, p_base, p1, p2, f, i)) r = true;
            if (cone_cone_cc(tang2, p_tip, p_base, p1, p2, f, i)) r = true;
        }
    }
    return r;
}

OCL_HD bool general_edge_push(const CutterP& c, const FibX& f, PIval& i, V3 p1, V3 p2, int& iters) {
    switch (c.kind) {
    case K_BALL: return ball_general_edge_push(c, f, i, p1, p2);
    case K_BULL: return bull_general_edge_push(c, f, i, p1, p2, iters);
    case K_CONE: return cone_general_edge_push(c, f, i, p1, p2);
    default: return false;
    }
}

// millingcutter.cpp:277-301: a horizontal edge pre-empts the shaft and
// general contacts of that edge
OCL_HD bool base_edge_push(const CutterP& c, const FibX& f, PIval& i, const TriX& t, int& iters) {
    bool r = false;
    for (int n = 0; n < 3; n++) {
        const V3 p1 = tri_vtx(t, n), p2 = tri_vtx(t, (n + 1) % 3);
        if (horiz_edge_push(c, f, i, p1, p2)) {
            r = true;
        } else {
            if (shaft_edge_push(c, f, i, p1, p2)) r = true;
            if (general_edge_push(c, f, i, p1, p2, iters)) r = true;
        }
    }
    return r;
}

// millingcutter.cpp:169-175 and ConeCutter::facetPush (conecutter.cpp:137-147)
OCL_HD bool simple_facet_push(const CutterP& c, const FibX& f, PIval& i, const TriX& t) {
    if (c.kind == K_CONE) {
        bool r = false;
        if (general_facet_push(0, 0, 0, f, i, t)) r = true;
        if (general_facet_push(0, c.ch, c.xyl, f, i, t)) r = true;
        return r;
    }
    return general_facet_push(c.nl, c.ch, c.xyl, f, i, t);
}

// one part (0 vertex, 1 facet, 2 edge) of a simple cutter
OCL_HD bool simple_push_part(const CutterP& c, const FibX& f, PIval& i, const TriX& t, int part, int& iters) {
    if (part == 0) return base_vertex_push(c, f, i, t);
    if (part == 1) return simple_facet_push(c, f, i, t);
    return base_edge_push(c, f, i, t, iters);
}

// CompositeCutter::vertexPush/facetPush/edgePush (compositecutter.cpp:196-270):
// each sub-cutter runs on a fiber raised by its z offset; a contact end is
// kept when its CC height maps back to that sub-cutter.
OCL_HD bool comp_push_part(const CutterDev& cd, const FibX& f, PIval& i, const TriX& t, int part, int& iters) {
    double ct[2 * kMaxSub];
    V3 cp[2 * kMaxSub];
    int cty[2 * kMaxSub];
    int nc = 0;
    for (int n = 0; n < cd.nsub; ++n) {
        PIval ci;
        piv_init(ci);
        FibX cf = f;
        cf.p1.z = f.p1.z + cd.zoff[n];
        cf.p2.z = f.p2.z + cd.zoff[n];
        if (simple_push_part(cd.sub[n], cf, ci, t, part, iters)) {
            if (n == comp_index_of_height(cd, ci.ucc.z - f.p1.z)) {
                ct[nc] = ci.hi;
                cp[nc] = ci.ucc;
                cty[nc] = ci.uty;
                ++nc;
            }
            if (n == comp_index_of_height(cd, ci.lcc.z - f.p1.z)) {
                ct[nc] = ci.lo;
                cp[nc] = ci.lcc;
                cty[nc] = ci.lty;
                ++nc;
            }
        }
    }
    for (int n = 0; n < nc; ++n) piv_update(i, ct[n], cp[n], cty[n]);
    return nc > 0;
}

// ------------------------------------------------------------ pair bounds
// A bound on the interval one (fiber, triangle) pair can add, used to leave
// out pairs that cannot change the fiber (no counterpart in the reference,
// whose pushCutter3 notes the idea as a to-do, batchpushcutter.cpp:177).
// Fiber::addInterval ignores an interval strictly inside an interval the
// fiber already holds (Fiber::contains with the strict Interval::inside,
// fiber.cpp:58-66, interval.cpp:106-111), and the fiber's intervals only
// grow and merge, so a pair whose bound lies strictly inside one of them
// when it is queued changes nothing whenever it would have been pushed.
//
// Every contact of the push code has its CC point on the triangle and its
// CL point within the cutter's profile width at the CC height (at most the
// radius) of the CC point in XY: vertices and the CylCutter's z-slice
// points are on the triangle; edge contacts pass isInside(p1, p2) on the
// edge's line; BullCutter edge contacts are offset-ellipse points, r1 from
// a point r2 from the edge line, for any root the solve returns; facet
// contacts are tested with the XY barycentric test (tri_contains).  That
// test is only trustworthy for a triangle whose XY projection is well
// conditioned: on a vertical or sliver triangle the reference accepts CC
// points far outside it, so such triangles get no bound.  Every contact
// has its CC point at or above the fiber (vertex and edge height tests,
// facet heights fz + ch - nl nz), so a well-conditioned triangle entirely
// below the fiber adds nothing.
//
// Record (three values per triangle, in the push tree's record padding):
// conditioned flag, lowest and highest vertex coordinate along the fiber
// axis (0 = x for X fibers, 1 = y for Y fibers).
OCL_HD void push_bound_rec(const double* p9, int axis, double out[3]) {
    const double ax = p9[0], ay = p9[1], bx = p9[3], by = p9[4], cx = p9[6], cy = p9[7];
    // the denominators of the barycentric test, as tri_contains computes them
    const double du = ay * cx - ax * cy + (cy - ay) * bx + (ax - cx) * by;
    const double dv = ax * by - ay * bx + (ay - by) * cx + (bx - ax) * cy;
    double m = 1.0;
    for (int k = 0; k < 3; ++k) m = fmax(m, fmax(fabs(p9[3 * k]), fabs(p9[3 * k + 1])));
    const double m2 = m * m;
    const double e0 = sqrt(sqr(bx - ax) + sqr(by - ay)), e1 = sqrt(sqr(cx - bx) + sqr(cy - by));
    const double e2 = sqrt(sqr(ax - cx) + sqr(ay - cy));
    // Rounding moves the test's decisions by about 50 eps m^2 / (shortest
    // edge) in XY and its denominators by 8 eps m^2: with these floors an
    // accepted CC point lies within 6e-5 mm of the triangle, well inside the
    // bound's 1e-3 mm margin.
    const bool ok = fabs(du) >= 1e-9 * m2 && fabs(dv) >= 1e-9 * m2 && fmin(e0, fmin(e1, e2)) >= 1e-10 * m2;
    out[0] = ok ? 1.0 : 0.0;
    out[1] = fmin(fmin(p9[axis], p9[3 + axis]), p9[6 + axis]);
    out[2] = fmax(fmax(p9[axis], p9[3 + axis]), p9[6 + axis]);
}

// Bound of one pair from the triangle's record (rec3 = push_bound_rec, zmax =
// highest vertex, c0/c1 = lowest/highest vertex coordinate across the
// fiber) for the fiber from a to b along its axis, at height fz and
// coordinate ac across it.  Returns 0: no bound (composite cutters,
// ill-conditioned triangles); 1: the pair adds no interval (the triangle
// lies below the fiber, or farther than the radius from it across: a CL
// point on the fiber is within the radius of its CC point in XY, so its CC
// point is within the radius of the fiber across it; the reference's KD
// search returns such triangles, kdtree.hpp:273-304); 2: any interval the
// pair adds lies in [tl, tu] (fiber parameter).  Ball and cone bounds use
// the profile width at the highest reachable CC height (their contacts are
// exact geometric tangencies); the other cutters use the radius.
OCL_HD int push_bound(const CutterP& c, const double rec3[3], double zmax, double c0, double c1, double fz, double ac,
                      double a, double b, double& tl, double& tu) {
    if (c.kind == K_COMPOSITE || rec3[0] == 0.0 || !(b != a)) return 0;
    const double hb = zmax - fz + 1e-4 + 1e-9 * (fabs(zmax) + fabs(fz));
    if (hb < 0.0) return 1;
    const double ma = c.radius + 1e-3 + 1e-9 * (fabs(c0) + fabs(c1) + fabs(ac) + c.radius);
    if (c1 < ac - ma || c0 > ac + ma) return 1;
    double w = c.radius;
    if (c.kind == K_BALL || c.kind == K_CONE) w = fmin(prof_width(c, hb), c.radius);
    w += 1e-3 + 1e-9 * (fabs(rec3[1]) + fabs(rec3[2]) + c.radius);
    const double inv = 1.0 / (b - a);
    double t0 = (rec3[1] - w - a) * inv, t1 = (rec3[2] + w - a) * inv;
    if (t0 > t1) {
        const double q = t0;
        t0 = t1;
        t1 = q;
    }
    tl = t0 - 1e-12 * (1.0 + fabs(t0));
    tu = t1 + 1e-12 * (1.0 + fabs(t1));
    return 2;
}

// MillingCutter::pushCutter (millingcutter.cpp:377-382): vertices, facet,
// edges, all applied to the same interval.  K = the cutter kind, fixed at
// compile time so each kernel carries only its cutter's code.
template <int K>
OCL_HD void push_pair(const CutterDev& cd, const FibX& f, const TriX& t, PIval& i, int& iters) {
    if (K == K_COMPOSITE) {
        for (int part = 0; part < 3; ++part) comp_push_part(cd, f, i, t, part, iters);
    } else {
        CutterP c = cd.c;
        c.kind = K;
        for (int part = 0; part < 3; ++part) simple_push_part(c, f, i, t, part, iters);
    }
}

}  // namespace ocl_amd
