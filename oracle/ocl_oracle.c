/*
 * ocl_oracle.c -- FP64 CPU restatement of OpenCAMLib's batch drop-cutter and
 * batch push-cutter.  TEST INFRASTRUCTURE ONLY: the parity checker of the HIP
 * path and the timed CPU baseline ("port").  The product library never links
 * or calls it.  See ocl_oracle.h for what pins each part.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).  Operation
 * order follows the reference so that FP64 results match its x86-64 build.
 */
#include "ocl_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TOL 0.0000001 /* numeric.cpp:34 */

/* ------------------------------------------------------------------ point */
typedef struct { double x, y, z; } P3;

static inline P3 p3(double x, double y, double z) { P3 r; r.x = x; r.y = y; r.z = z; return r; }
static inline P3 padd(P3 a, P3 b) { return p3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline P3 psub(P3 a, P3 b) { return p3(a.x - b.x, a.y - b.y, a.z - b.z); }
/* Point *= a  (point.cpp:270-275) */
static inline P3 pmul(double a, P3 p) { return p3(p.x * a, p.y * a, p.z * a); }
static inline double sq(double x) { return x * x; }
/* point.cpp:60-62 */
static inline double pnorm(P3 p) { return sqrt(sq(p.x) + sq(p.y) + sq(p.z)); }
/* point.cpp:71-73 */
static inline double pdot(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* point.cpp:64-69 */
static inline P3 pcross(P3 a, P3 b) {
    return p3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
/* point.cpp:75-79 */
static inline P3 pnormalize(P3 p) {
    double n = pnorm(p);
    if (n != 0.0) p = pmul(1 / n, p);
    return p;
}
/* point.cpp:81-83 */
static inline double pxynorm(P3 p) { return sqrt(sq(p.x) + sq(p.y)); }
/* point.cpp:85-88 */
static inline P3 pxynormalize(P3 p) {
    if (pxynorm(p) != 0.0) p = pmul(1 / pxynorm(p), p);
    return p;
}
/* point.cpp:90-93 */
static inline P3 pxyperp(P3 p) { return p3(-p.y, p.x, p.z); }
static inline int iszero_tol(double x) { return fabs(x) < TOL; } /* numeric.cpp:75-80 */
static inline double sgn(double x) { return x < 0.0 ? -1 : 1; }  /* numeric.cpp:54-59 */

/* point.cpp:131-135 */
static inline double pxydist(P3 a, P3 b) { return pxynorm(psub(a, b)); }

/* point.cpp:140-163 */
static double xy_distance_to_line(P3 self, P3 p1, P3 p2) {
    if ((p1.x == p2.x) && (p1.y == p2.y)) return -1;
    P3 v = pnormalize(p3(p2.y - p1.y, -(p2.x - p1.x), 0));
    P3 r = p3(p1.x - self.x, p1.y - self.y, 0);
    return fabs(pdot(v, r));
}

/* point.cpp:166-172 */
static P3 closest_point(P3 self, P3 p1, P3 p2) {
    P3 v = psub(p2, p1);
    double u = pdot(psub(self, p1), v) / pdot(v, v);
    return padd(p1, pmul(u, v));
}

/* point.cpp:175-201 */
static P3 xy_closest_point(P3 self, P3 p1, P3 p2) {
    P3 v = psub(p2, p1);
    if (iszero_tol(pxynorm(v))) return p3(0, 0, 0);
    double u = (self.x - p1.x) * (v.x) + (self.y - p1.y) * (v.y);
    u = u / (v.x * v.x + v.y * v.y);
    return p3(p1.x + u * v.x, p1.y + u * v.y, 0);
}

/* point.cpp:337-347 */
static inline double z_project_onto_edge(P3 self, P3 p1, P3 p2) {
    double t;
    if (fabs(p2.x - p1.x) > fabs(p2.y - p1.y))
        t = (self.x - p1.x) / (p2.x - p1.x);
    else
        t = (self.y - p1.y) / (p2.y - p1.y);
    return p1.z + t * (p2.z - p1.z);
}

/* point.cpp:241-253, segment test with inclusive t in [0,1]; NaN passes */
static inline int inside_points(P3 self, P3 p1, P3 p2) {
    P3 d = psub(p2, p1);
    P3 s = psub(self, p1);
    double t = pdot(s, d) / pdot(d, d);
    if (t > 1.0 || t < 0.0) return 0;
    return 1;
}

/* --------------------------------------------------------------- triangle */
typedef struct {
    P3 p[3];
    P3 n;
    double bb[6]; /* minx maxx miny maxy minz maxz (bbox.cpp:92-118) */
} Tri;

/* triangle.cpp:39-75, bbox.cpp:53-84 */
static void tri_init(Tri* t, P3 a, P3 b, P3 c) {
    t->p[0] = a; t->p[1] = b; t->p[2] = c;
    P3 v1 = psub(a, b), v2 = psub(a, c);
    t->n = pnormalize(pcross(v1, v2));
    double mn[3] = {a.x, a.y, a.z}, mx[3] = {a.x, a.y, a.z};
    for (int k = 1; k < 3; ++k) {
        P3 p = t->p[k];
        double c3[3] = {p.x, p.y, p.z};
        for (int d = 0; d < 3; ++d) {
            if (c3[d] > mx[d]) mx[d] = c3[d];
            if (c3[d] < mn[d]) mn[d] = c3[d];
        }
    }
    t->bb[0] = mn[0]; t->bb[1] = mx[0]; t->bb[2] = mn[1];
    t->bb[3] = mx[1]; t->bb[4] = mn[2]; t->bb[5] = mx[2];
}

/* triangle.cpp:73-75 */
static inline P3 up_normal(const Tri* t) { return (t->n.z < 0) ? pmul(-1.0, t->n) : t->n; }

/* point.cpp:224-239, strict barycentric test in XY */
static int inside_tri(P3 p, const Tri* t) {
    P3 a = t->p[0], b = t->p[1], c = t->p[2];
    double u = (a.y * c.x - a.x * c.y + (c.y - a.y) * p.x + (a.x - c.x) * p.y) /
               (a.y * c.x - a.x * c.y + (c.y - a.y) * b.x + (a.x - c.x) * b.y);
    double v = (a.x * b.y - a.y * b.x + (a.y - b.y) * p.x + (b.x - a.x) * p.y) /
               (a.x * b.y - a.y * b.x + (a.y - b.y) * c.x + (b.x - a.x) * c.y);
    return u > 0.0 && v > 0.0 && (u + v) < 1.0;
}

struct oc_mesh {
    long n;
    Tri* t;
    /* lazily built KD trees, one per dimension set */
    struct kdtree* kd[3];
    unsigned kd_bucket[3];
};

/* ------------------------------------------------------------- CL record */
typedef struct {
    double x, y, z;
    P3 cc;
    int type;
    int tri, feat;   /* last lift, for parity diagnostics */
    int cur_tri, cur_feat;
    int diag;        /* record quirk bits (slower; off in timed runs) */
    int quirk;       /* OC_Q_* bits: where a reference rule discarded a contact */
} CL;

/* clpoint.cpp:63-68 */
static inline int cl_below(const CL* cl, const Tri* t) { return cl->z < t->bb[5]; }

/* clpoint.cpp:79-99 */
static inline int cl_lift(CL* cl, double zin, P3 cc, int type) {
    if (zin > cl->z) {
        cl->z = zin;
        cl->cc = cc;
        cl->type = type;
        cl->tri = cl->cur_tri;
        cl->feat = cl->cur_feat;
        return 1;
    }
    return 0;
}

/* ------------------------------------------------------- cutter profiles */
void oc_cutter_cyl(oc_cutter* c, double d, double l) {
    memset(c, 0, sizeof(*c));
    c->kind = OC_CYL;
    c->diameter = d; c->radius = d / 2.0; c->length = l;
    c->xy_normal_length = c->radius; c->normal_length = 0.0; c->center_height = 0.0;
}
void oc_cutter_ball(oc_cutter* c, double d, double l) {
    memset(c, 0, sizeof(*c));
    c->kind = OC_BALL;
    c->diameter = d; c->radius = d / 2.0; c->length = l;
    c->normal_length = c->radius; c->xy_normal_length = 0.0; c->center_height = c->radius;
}
void oc_cutter_bull(oc_cutter* c, double d, double r, double l) {
    memset(c, 0, sizeof(*c));
    c->kind = OC_BULL;
    c->diameter = d; c->radius = d / 2.0;
    c->radius1 = d / 2.0 - r; c->radius2 = r; c->length = l;
    c->xy_normal_length = c->radius1; c->normal_length = c->radius2; c->center_height = c->radius2;
}
void oc_cutter_cone(oc_cutter* c, double d, double a, double l) {
    memset(c, 0, sizeof(*c));
    c->kind = OC_CONE;
    c->diameter = d; c->radius = d / 2.0; c->angle = a;
    c->length = c->radius / tan(a) + l;
    c->center_height = c->radius / tan(a);
    c->xy_normal_length = c->radius; c->normal_length = 0.0;
}
void oc_cutter_composite(oc_cutter* c) {
    memset(c, 0, sizeof(*c));
    c->kind = OC_COMPOSITE;
}
void oc_composite_add(oc_cutter* c, oc_cutter* sub, double r, double h, double zoff) {
    int n = c->nsub++;
    c->sub[n] = sub; c->radiusvec[n] = r; c->heightvec[n] = h; c->zoffset[n] = zoff;
    if (r > c->radius) { c->radius = r; c->diameter = 2 * r; }
}

/* compositecutter.cpp:111-122 */
static int comp_valid_radius(const oc_cutter* c, int n, double r) {
    double lo = (n == 0) ? -1E-6 : c->radiusvec[n - 1] - 1E-6;
    double hi = c->radiusvec[n] + 1e-6;
    return (lo <= r) && (r <= hi);
}
/* compositecutter.cpp:97-108 */
static int comp_valid_height(const oc_cutter* c, int n, double h) {
    double lo = (n == 0) ? -1E-6 : c->heightvec[n - 1] - 1E-6;
    double hi = c->heightvec[n] + 1e-6;
    return (lo <= h) && (h <= hi);
}
/* compositecutter.cpp:64-83 */
static int comp_radius_to_index(const oc_cutter* c, double r) {
    for (int n = 0; n < c->nsub; ++n)
        if (comp_valid_radius(c, n, r)) return n;
    return 0;
}
static int comp_height_to_index(const oc_cutter* c, double h) {
    for (int n = 0; n < c->nsub; ++n)
        if (comp_valid_height(c, n, h)) return n;
    return c->nsub - 1;
}

/* height(r): cylcutter.hpp:52, ballcutter.hpp:51, bullcutter.cpp:54-63,
   conecutter.cpp:51-54, compositecutter.cpp:58-61 */
static double cutter_height(const oc_cutter* c, double r) {
    switch (c->kind) {
    case OC_CYL: return (r <= c->radius) ? 0.0 : -1.0;
    case OC_BALL: return c->radius - sqrt(sq(c->radius) - sq(r));
    case OC_BULL:
        if (r <= c->radius1) return 0.0;
        else if (r <= c->radius) return c->radius2 - sqrt(sq(c->radius2) - sq(r - c->radius1));
        return -1;
    case OC_CONE: return r / tan(c->angle);
    default: {
        int idx = comp_radius_to_index(c, r);
        return cutter_height(c->sub[idx], r) + c->zoffset[idx];
    }
    }
}

/* width(h): cylcutter.hpp:53, ballcutter.hpp:52, bullcutter.cpp:66-68,
   conecutter.cpp:56-60, compositecutter.cpp:64-69 */
static double cutter_width(const oc_cutter* c, double h) {
    switch (c->kind) {
    case OC_CYL: return c->radius;
    case OC_BALL: return (h >= c->radius) ? c->radius : sqrt(sq(c->radius) - sq(c->radius - h));
    case OC_BULL:
        return (h >= c->radius2) ? c->radius : c->radius1 + sqrt(sq(c->radius2) - sq(c->radius2 - h));
    case OC_CONE: return (h < c->center_height) ? h * tan(c->angle) : c->radius;
    default: {
        int idx = comp_height_to_index(c, h);
        return cutter_width(c->sub[idx], h - c->zoffset[idx]);
    }
    }
}

/* ------------------------------------------------ brent + ellipse (bull) */
/* numeric.cpp:36-51 */
static double xy_vector_to_diangle(double x, double y) {
    double diangle;
    if (y >= 0)
        diangle = (x >= 0 ? y / (x + y) : 1 - x / (-x + y));
    else
        diangle = (x < 0 ? 2 - y / (-x - y) : 3 + x / (x - y));
    return diangle;
}

typedef struct { double s, t, diangle; } EPos;

/* ellipseposition.cpp:47-70 */
static EPos epos_set(double dia) {
    EPos e;
    e.diangle = dia;
    double d = dia;
    while (d > 4.0) d -= 4.0;
    while (d < 0.0) d += 4.0;
    P3 p = p3((d < 2 ? 1 - d : d - 3), (d < 3 ? ((d > 1) ? 2 - d : d) : d - 4), 0.0);
    p = pnormalize(p);
    e.s = p.x;
    e.t = p.y;
    return e;
}

/* Ellipse (ellipse.cpp:38-46) with the fields it uses; `aligned` selects the
   AlignedEllipse overrides (ellipse.cpp:48-57, 75-78, 88-92, 100-114, 313-319) */
typedef struct {
    int aligned;
    P3 center;
    double a, b, offset;
    P3 target;
    P3 major_dir, minor_dir, error_dir;
    EPos pos1, pos2, pos_hi;
} Ell;

static P3 ell_epoint(const Ell* e, EPos pos) {
    if (e->aligned) {
        /* center + a*s*major + b*t*minor */
        return padd(padd(e->center, pmul(e->a * pos.s, e->major_dir)), pmul(e->b * pos.t, e->minor_dir));
    }
    P3 p = e->center;
    p.x += e->a * pos.s;
    p.y += e->b * pos.t;
    return p;
}
static P3 ell_normal(const Ell* e, EPos pos) {
    if (e->aligned) {
        P3 n = padd(pmul(pos.s * e->b, e->major_dir), pmul(pos.t * e->a, e->minor_dir));
        return pnormalize(n);
    }
    return pnormalize(p3(e->b * pos.s, e->a * pos.t, 0));
}
static P3 ell_oepoint(const Ell* e, EPos pos) {
    return padd(ell_epoint(e, pos), pmul(e->offset, ell_normal(e, pos)));
}
/* Ellipse::error(EllipsePosition&) ellipse.cpp:321-324 */
static double ell_error_pos(const Ell* e, EPos pos) { return ell_oepoint(e, pos).y; }
/* Ellipse::error(double) / AlignedEllipse::error(double), ellipse.cpp:313-330 */
static double ell_error(const Ell* e, double dia) {
    EPos tmp = epos_set(dia);
    if (e->aligned) {
        P3 p = ell_oepoint(e, tmp);
        P3 ev = psub(e->target, p);
        return pdot(ev, e->error_dir);
    }
    return ell_error_pos(e, tmp);
}

/* brent_zero.hpp:36-123 */
static double brent_zero(double a, double b, double eps, double t, const Ell* ell) {
    double c, d, e, fa, fb, fc, m, p, q, r, s, tol;
    fa = ell_error(ell, a);
    fb = ell_error(ell, b);
    c = a;
    fc = fa;
    e = b - a;
    d = e;
    while (1) {
        if (fabs(fc) < fabs(fb)) {
            a = b; b = c; c = a;
            fa = fb; fb = fc; fc = fa;
        }
        tol = 2.0 * eps * fabs(b) + t;
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
        fb = ell_error(ell, b);
        if (signbit(fb) == signbit(fc)) {
            c = a;
            fc = fa;
            e = b - a;
            d = e;
        }
    }
    return b;
}

#define OE_ERROR_TOLERANCE 1e-10

/* ellipse.cpp:120-181 */
static void ell_find_pos2(Ell* e) {
    double err1 = fabs(ell_error_pos(e, e->pos1));
    const double ss[4] = {1, -1, -1, 1}, ts[4] = {-1, 1, -1, 1};
    for (int k = 0; k < 4; ++k) {
        EPos p2 = e->pos1;
        p2.s = ss[k] * e->pos1.s;
        p2.t = ts[k] * e->pos1.t;
        e->pos2 = p2;
        if (fabs(ell_error_pos(e, p2)) < err1 + OE_ERROR_TOLERANCE) {
            if (k == 3) return; /* identical solutions accepted */
            if ((fabs(p2.s - e->pos1.s) > 1E-8) || (fabs(p2.t - e->pos1.t) > 1E-8)) return;
        }
    }
}

/* ellipse.cpp:185-208 */
static void ell_solver_brent(Ell* e) {
    EPos apos = epos_set(0.0), bpos = epos_set(3.0);
    if (fabs(ell_error_pos(e, apos)) < OE_ERROR_TOLERANCE) {
        e->pos1 = apos;
        ell_find_pos2(e);
        return;
    } else if (fabs(ell_error_pos(e, bpos)) < OE_ERROR_TOLERANCE) {
        e->pos1 = bpos;
        ell_find_pos2(e);
        return;
    }
    double dia = brent_zero(apos.diangle, bpos.diangle, 3E-16, OE_ERROR_TOLERANCE, e);
    e->pos1 = epos_set(dia);
    ell_find_pos2(e);
}

/* ellipse.cpp:335-356 */
static P3 ell_calc_ecenter(const Ell* e, P3 up1, P3 up2, int sln) {
    P3 cle = ell_oepoint(e, sln == 1 ? e->pos1 : e->pos2);
    double xoffset = -cle.x;
    double t = (e->center.x + xoffset - up1.x) / (up2.x - up1.x);
    return padd(up1, pmul(t, psub(up2, up1)));
}
static void ell_set_pos_hi(Ell* e, P3 u1, P3 u2) {
    P3 ec1 = ell_calc_ecenter(e, u1, u2, 1);
    P3 ec2 = ell_calc_ecenter(e, u1, u2, 2);
    if (ec1.z >= ec2.z) {
        e->pos_hi = e->pos1;
        e->center = ec1;
    } else {
        e->pos_hi = e->pos2;
        e->center = ec2;
    }
}

/* ellipse.cpp:212-311 (AlignedEllipse::aligned_solver) */
typedef struct { P3 p1, p2, dir; } Fib;

static int ell_aligned_solver(Ell* e, const Fib* f) {
    e->error_dir = pxyperp(f->dir);
    e->target = f->p1;
    double t1 = 0.0;
    if (f->p1.y == f->p2.y)
        t1 = sqrt(sq(e->b * e->minor_dir.y) / (sq(e->a * e->major_dir.y) + sq(e->b * e->minor_dir.y)));
    else if (f->p1.x == f->p2.x)
        t1 = sqrt(sq(e->b * e->minor_dir.x) / (sq(e->a * e->major_dir.x) + sq(e->b * e->minor_dir.x)));
    EPos tmp, apos = epos_set(0.0), bpos = epos_set(0.0);
    double s1 = sqrt(1.0 - sq(t1));
    int fpos = 0, fneg = 0;
    const double sx[4] = {1, 1, -1, -1}, ty[4] = {1, -1, 1, -1};
    for (int k = 0; k < 4; ++k) {
        tmp = epos_set(xy_vector_to_diangle(sx[k] * s1, ty[k] * t1));
        double err = ell_error(e, tmp.diangle);
        if (err > 0) { fpos = 1; apos = tmp; }
        else if (err < 0) { fneg = 1; bpos = tmp; }
    }
    if (fpos && fneg) {
        double lolim = 0.0, hilim = 0.0;
        if (apos.diangle > bpos.diangle) { lolim = bpos.diangle; hilim = apos.diangle; }
        else if (bpos.diangle > apos.diangle) { hilim = bpos.diangle; lolim = apos.diangle; }
        double d1 = brent_zero(lolim, hilim, 3E-16, OE_ERROR_TOLERANCE, e);
        double d2 = brent_zero(hilim - 4.0, lolim, 3E-16, OE_ERROR_TOLERANCE, e);
        e->pos1 = epos_set(d1);
        e->pos2 = epos_set(d2);
        return 1;
    }
    return 0;
}

/* -------------------------------------------------------------- drop */
/* feature codes: low 4 bits kind, bits 4..5 composite sub-cutter */
enum { F_FACET = 0, F_FACET_RIM = 1, F_VTX0 = 2, F_EDGE0 = 5 };

static int facet_drop(const oc_cutter* c, CL* cl, const Tri* t);
static int edge_drop(const oc_cutter* c, CL* cl, const Tri* t);

/* millingcutter.cpp:51-75 */
static int base_facet_drop(const oc_cutter* c, CL* cl, const Tri* t) {
    P3 normal = up_normal(t);
    if (iszero_tol(normal.z)) return 0;
    if (iszero_tol(normal.x) && iszero_tol(normal.y)) {
        P3 cc = p3(cl->x, cl->y, t->p[0].z);
        if (inside_tri(cc, t)) return cl_lift(cl, cc.z, cc, OC_FACET);
        return 0;
    }
    double d = -pdot(normal, t->p[0]);
    normal = pnormalize(normal);
    P3 xyn = pxynormalize(p3(normal.x, normal.y, 0.0));
    P3 rv = padd(pmul(c->xy_normal_length, xyn), pmul(c->normal_length, normal));
    P3 cc = psub(p3(cl->x, cl->y, cl->z), rv);
    cc.z = (1.0 / normal.z) * (-d - normal.x * cc.x - normal.y * cc.y);
    double tip_z = cc.z + rv.z - c->center_height;
    if (inside_tri(cc, t)) return cl_lift(cl, tip_z, cc, OC_FACET);
    return 0;
}

/* conecutter.cpp:70-105 */
static int cone_facet_drop(const oc_cutter* c, CL* cl, const Tri* t) {
    int result = 0;
    P3 normal = up_normal(t);
    if (iszero_tol(normal.z)) return 0;
    if (iszero_tol(normal.x) && iszero_tol(normal.y)) {
        P3 cc = p3(cl->x, cl->y, t->p[0].z);
        if (inside_tri(cc, t)) return cl_lift(cl, cc.z, cc, OC_FACET_TIP);
        return 0;
    }
    double a = normal.x, b = normal.y, cz = normal.z;
    double d = -pdot(normal, t->p[0]);
    normal = pxynormalize(normal);
    P3 cyl = psub(p3(cl->x, cl->y, cl->z), pmul(c->radius, normal));
    cyl.z = (1.0 / cz) * (-d - a * cyl.x - b * cyl.y);
    double cyl_cl_z = cyl.z - c->length;
    P3 tip = p3(cl->x, cl->y, 0.0);
    tip.z = (1.0 / cz) * (-d - a * tip.x - b * tip.y);
    double tip_cl_z = tip.z;
    int base_feat = cl->cur_feat;
    if (!result) result = inside_tri(tip, t) ? cl_lift(cl, tip_cl_z, tip, OC_FACET_TIP) : 0;
    if (!result) {
        cl->cur_feat = base_feat + F_FACET_RIM;
        result = inside_tri(cyl, t) ? cl_lift(cl, cyl_cl_z, cyl, OC_FACET_CYL) : 0;
        cl->cur_feat = base_feat;
    }
    return result;
}

/* compositecutter.cpp:124-147 (ccValidRadius) */
static int comp_cc_valid_radius(const oc_cutter* c, int n, const CL* cl) {
    if (cl->type == OC_NONE) return 0;
    double d = pxydist(p3(cl->x, cl->y, cl->z), cl->cc);
    double lo = (n == 0) ? -1E-6 : c->radiusvec[n - 1] - 1E-6;
    double hi = c->radiusvec[n] + 1e-6;
    if (d < lo) return 0;
    if (d > hi) return 0;
    return 1;
}

/* ---- parity diagnostics (test infrastructure; no reference counterpart).
   They record, in CL.quirk, where one of the reference's rules discards a
   contact that the geometric supremum of p.z - h(|p - cl|) over the
   triangle would keep, so that tests compare with the supremum everywhere
   else.  None of them changes a result. */
static int single_edge_drop(const oc_cutter* c, CL* cl, P3 p1, P3 p2, double d);
static int vertex_drop(const oc_cutter* c, CL* cl, const Tri* t);
static int edge_drop(const oc_cutter* c, CL* cl, const Tri* t);

/* A composite sub-cutter's edge drop keeps only its highest edge contact
   and the window test rejected it (compositecutter.cpp:175-194): did one of
   the triangle's edges on its own give sub-cutter n an in-window contact
   above the running height?  (Edge selection as millingcutter.cpp:79-94.) */
static int diag_comp_edge_lost(const oc_cutter* c, int n, const CL* cl, const Tri* t) {
    const oc_cutter* s = c->sub[n];
    for (int k = 0; k < 3; ++k) {
        P3 p1 = t->p[k], p2 = t->p[(k + 1) % 3];
        if (iszero_tol(p1.x - p2.x) && iszero_tol(p1.y - p2.y)) continue;
        CL tmp;
        memset(&tmp, 0, sizeof(tmp));
        tmp.x = cl->x; tmp.y = cl->y; tmp.z = cl->z + c->zoffset[n];
        tmp.type = OC_NONE; tmp.tri = -1; tmp.feat = -1;
        double d = xy_distance_to_line(p3(tmp.x, tmp.y, tmp.z), p1, p2);
        if (d <= s->radius && single_edge_drop(s, &tmp, p1, p2, d) && comp_cc_valid_radius(c, n, &tmp)) return 1;
    }
    return 0;
}

/* A cone sub-cutter's tip contact lifted, so its rim contact was not tried
   (conecutter.cpp:100-102): would the rim contact have been inside the
   facet, in the window and above the composite's height after the tip (the
   tip itself may have passed or failed the window)? */
static int diag_cone_rim_lost(const oc_cutter* c, int n, const CL* cl, const Tri* t) {
    const oc_cutter* s = c->sub[n];
    P3 normal = up_normal(t);
    if (iszero_tol(normal.x) && iszero_tol(normal.y)) return 0; /* horizontal: no rim case */
    double a = normal.x, b = normal.y, cz = normal.z;
    double d = -pdot(normal, t->p[0]);
    P3 xyn = pxynormalize(normal);
    CL tmp;
    memset(&tmp, 0, sizeof(tmp));
    tmp.x = cl->x; tmp.y = cl->y; tmp.z = cl->z + c->zoffset[n];
    P3 cyl = psub(p3(tmp.x, tmp.y, tmp.z), pmul(s->radius, xyn));
    cyl.z = (1.0 / cz) * (-d - a * cyl.x - b * cyl.y);
    if (!inside_tri(cyl, t)) return 0;
    if (!(cyl.z - s->length - c->zoffset[n] > cl->z)) return 0;
    tmp.z = cyl.z - s->length;
    tmp.cc = cyl;
    tmp.type = OC_FACET_CYL;
    return comp_cc_valid_radius(c, n, &tmp);
}

/* The facet lifted, so the vertex and edge tests of the triangle were
   skipped (millingcutter.cpp:398-405): would one of them have lifted
   higher? */
static int diag_facet_first_hid(const oc_cutter* c, const CL* cl, const Tri* t) {
    CL tmp = *cl;
    tmp.diag = 0;
    vertex_drop(c, &tmp, t);
    edge_drop(c, &tmp, t);
    return tmp.z > cl->z;
}

/* compositecutter.cpp:152-172 */
static int comp_facet_drop(const oc_cutter* c, CL* cl, const Tri* t) {
    int result = 0;
    for (int n = 0; n < c->nsub; ++n) {
        CL tmp;
        memset(&tmp, 0, sizeof(tmp));
        tmp.x = cl->x + 0; tmp.y = cl->y + 0; tmp.z = cl->z + c->zoffset[n];
        tmp.type = OC_NONE; tmp.tri = -1; tmp.feat = -1;
        tmp.cur_tri = cl->cur_tri; tmp.cur_feat = (n << 4) | F_FACET;
        if (facet_drop(c->sub[n], &tmp, t)) {
            if (comp_cc_valid_radius(c, n, &tmp)) {
                int save = cl->cur_feat;
                cl->cur_feat = tmp.feat;
                if (cl_lift(cl, tmp.z - c->zoffset[n], tmp.cc, OC_FACET)) result = 1;
                cl->cur_feat = save;
            }
            if (cl->diag && c->sub[n]->kind == OC_CONE && tmp.type == OC_FACET_TIP &&
                diag_cone_rim_lost(c, n, cl, t))
                cl->quirk |= OC_Q_CONE_TIP;
        }
    }
    return result;
}

static int facet_drop(const oc_cutter* c, CL* cl, const Tri* t) {
    if (c->kind == OC_CONE) return cone_facet_drop(c, cl, t);
    if (c->kind == OC_COMPOSITE) return comp_facet_drop(c, cl, t);
    return base_facet_drop(c, cl, t);
}

/* millingcutter.cpp:36-47 */
static int vertex_drop(const oc_cutter* c, CL* cl, const Tri* t) {
    int result = 0;
    P3 clp = p3(cl->x, cl->y, cl->z);
    int base_feat = cl->cur_feat;
    for (int k = 0; k < 3; ++k) {
        P3 p = t->p[k];
        double q = pxydist(clp, p);
        if (q <= c->radius) {
            cl->cur_feat = base_feat + F_VTX0 + k;
            if (cl_lift(cl, p.z - cutter_height(c, q), p, OC_VERTEX)) result = 1;
        }
    }
    cl->cur_feat = base_feat;
    return result;
}

typedef struct { double cc_u, cl_z; } CCZ;

/* cylcutter.cpp:52-70 */
static CCZ cyl_edge_canonical(const oc_cutter* c, P3 u1, P3 u2) {
    double s = sqrt(sq(c->radius) - sq(u1.y));
    P3 cc1 = p3(s, u1.y, 0), cc2 = p3(-s, u1.y, 0);
    cc1.z = z_project_onto_edge(cc1, u1, u2);
    cc2.z = z_project_onto_edge(cc2, u1, u2);
    CCZ r;
    if (cc1.z > cc2.z) { r.cc_u = cc1.x; r.cl_z = cc1.z; }
    else { r.cc_u = cc2.x; r.cl_z = cc2.z; }
    return r;
}
/* ballcutter.cpp:50-63 */
static CCZ ball_edge_canonical(const oc_cutter* c, P3 u1, P3 u2) {
    double s = sqrt(sq(c->radius) - sq(u1.y));
    P3 normal = pxynormalize(p3(u2.z - u1.z, -(u2.x - u1.x), 0));
    if (normal.y < 0) normal = pmul(-1, normal);
    P3 cc = p3(-s * normal.x, u1.y, 0);
    cc.z = z_project_onto_edge(cc, u1, u2);
    CCZ r;
    r.cc_u = cc.x;
    r.cl_z = cc.z + s * normal.y - c->radius;
    return r;
}
/* bullcutter.cpp:82-105 */
static CCZ bull_edge_canonical(const oc_cutter* c, P3 u1, P3 u2) {
    CCZ r;
    if (iszero_tol(u1.z - u2.z)) {
        r.cc_u = 0;
        r.cl_z = u1.z - cutter_height(c, u1.y);
        return r;
    }
    double b_axis = c->radius2;
    double theta = atan((u2.z - u1.z) / (u2.x - u1.x));
    double a_axis = fabs(c->radius2 / sin(theta));
    Ell e;
    memset(&e, 0, sizeof(e));
    e.center = p3(0, u1.y, 0);
    e.a = a_axis; e.b = b_axis; e.offset = c->radius1;
    e.target = p3(0, 0, 0);
    ell_solver_brent(&e);
    ell_set_pos_hi(&e, u1, u2);
    P3 ell_ccp = ell_epoint(&e, e.pos_hi);
    P3 cc_u = closest_point(ell_ccp, u1, u2);
    r.cc_u = cc_u.x;
    r.cl_z = e.center.z - c->radius2;
    return r;
}
/* conecutter.cpp:109-135 */
static CCZ cone_edge_canonical(const oc_cutter* c, P3 u1, P3 u2) {
    double d = u1.y;
    double m = (u2.z - u1.z) / (u2.x - u1.x);
    double xu = sqrt(sq(c->radius) - sq(u1.y));
    double mu = (c->center_height / c->radius) * xu / sqrt(sq(xu) + sq(d));
    int hyper = (fabs(m) <= fabs(mu));
    double ccu;
    if (hyper)
        ccu = sgn(m) * sqrt(sq(c->radius) * sq(m) * sq(d) / (sq(c->length) - sq(c->radius) * sq(m)));
    else
        ccu = sgn(m) * xu;
    P3 cc = p3(ccu, d, 0.0);
    cc.z = z_project_onto_edge(cc, u1, u2);
    CCZ r;
    r.cc_u = ccu;
    if (hyper)
        r.cl_z = cc.z - c->center_height + (c->radius - sqrt(sq(ccu) + sq(d))) / tan(c->angle);
    else
        r.cl_z = cc.z - c->center_height;
    return r;
}

/* millingcutter.cpp:121-136 */
static int single_edge_drop(const oc_cutter* c, CL* cl, P3 p1, P3 p2, double d) {
    P3 v = psub(p2, p1);
    P3 vxy = pxynormalize(p3(v.x, v.y, 0.0));
    P3 clp = p3(cl->x, cl->y, cl->z);
    P3 sc = xy_closest_point(clp, p1, p2);
    P3 u1 = p3(pdot(psub(p1, sc), vxy), d, p1.z);
    P3 u2 = p3(pdot(psub(p2, sc), vxy), d, p2.z);
    CCZ con;
    switch (c->kind) {
    case OC_CYL: con = cyl_edge_canonical(c, u1, u2); break;
    case OC_BALL: con = ball_edge_canonical(c, u1, u2); break;
    case OC_BULL: con = bull_edge_canonical(c, u1, u2); break;
    default: con = cone_edge_canonical(c, u1, u2); break;
    }
    P3 cc = padd(sc, pmul(con.cc_u, vxy));
    cc.z = z_project_onto_edge(cc, p1, p2);
    if (inside_points(cc, p1, p2)) return cl_lift(cl, con.cl_z, cc, OC_EDGE);
    return 0;
}

/* One edge of MillingCutter::edgeDrop (millingcutter.cpp:84-91) for a
   simple cutter: the XY-degenerate skip, the radius test, singleEdgeDrop
   (millingcutter.cpp:121-136).  p1p2 = 6 doubles; z_io = running height in
   and out; cc_out (3) = CC point when lifted.  Returns 1 if lifted. */
int oc_edge_drop(const oc_cutter* c, const double* p1p2, double x, double y, double* z_io, double* cc_out) {
    if (c->kind == OC_COMPOSITE) return 0;
    P3 p1 = p3(p1p2[0], p1p2[1], p1p2[2]), p2 = p3(p1p2[3], p1p2[4], p1p2[5]);
    if (iszero_tol(p1.x - p2.x) && iszero_tol(p1.y - p2.y)) return 0;
    CL cl;
    memset(&cl, 0, sizeof(cl));
    cl.x = x; cl.y = y; cl.z = *z_io;
    cl.type = OC_NONE; cl.tri = -1; cl.feat = -1;
    double d = xy_distance_to_line(p3(cl.x, cl.y, cl.z), p1, p2);
    if (!(d <= c->radius)) return 0;
    int r = single_edge_drop(c, &cl, p1, p2, d);
    *z_io = cl.z;
    if (r && cc_out) { cc_out[0] = cl.cc.x; cc_out[1] = cl.cc.y; cc_out[2] = cl.cc.z; }
    return r;
}

/* millingcutter.cpp:79-94 */
static int base_edge_drop(const oc_cutter* c, CL* cl, const Tri* t) {
    int result = 0;
    P3 clp = p3(cl->x, cl->y, cl->z);
    int base_feat = cl->cur_feat;
    for (int n = 0; n < 3; n++) {
        P3 p1 = t->p[n], p2 = t->p[(n + 1) % 3];
        if (!iszero_tol(p1.x - p2.x) || !iszero_tol(p1.y - p2.y)) {
            double d = xy_distance_to_line(clp, p1, p2);
            if (d <= c->radius) {
                cl->cur_feat = base_feat + F_EDGE0 + n;
                if (single_edge_drop(c, cl, p1, p2, d)) result = 1;
            }
        }
    }
    cl->cur_feat = base_feat;
    return result;
}

/* compositecutter.cpp:175-194 */
static int comp_edge_drop(const oc_cutter* c, CL* cl, const Tri* t) {
    int result = 0;
    for (int n = 0; n < c->nsub; ++n) {
        CL tmp;
        memset(&tmp, 0, sizeof(tmp));
        tmp.x = cl->x + 0; tmp.y = cl->y + 0; tmp.z = cl->z + c->zoffset[n];
        tmp.type = OC_NONE; tmp.tri = -1; tmp.feat = -1;
        tmp.cur_tri = cl->cur_tri; tmp.cur_feat = (n << 4);
        if (edge_drop(c->sub[n], &tmp, t)) {
            if (comp_cc_valid_radius(c, n, &tmp)) {
                int save = cl->cur_feat;
                cl->cur_feat = tmp.feat;
                if (cl_lift(cl, tmp.z - c->zoffset[n], tmp.cc, OC_EDGE)) result = 1;
                cl->cur_feat = save;
            } else if (cl->diag && diag_comp_edge_lost(c, n, cl, t)) {
                cl->quirk |= OC_Q_COMP_EDGE;
            }
        }
    }
    return result;
}

static int edge_drop(const oc_cutter* c, CL* cl, const Tri* t) {
    if (c->kind == OC_COMPOSITE) return comp_edge_drop(c, cl, t);
    return base_edge_drop(c, cl, t);
}

/* millingcutter.cpp:386-409 */
static int drop_cutter(const oc_cutter* c, CL* cl, const Tri* t) {
    int facet = 0, vertex = 0, edge = 0;
    if (cl_below(cl, t)) {
        cl->cur_feat = 0;
        facet = facet_drop(c, cl, t);
        if (!facet) {
            vertex = vertex_drop(c, cl, t);
            if (cl_below(cl, t)) edge = edge_drop(c, cl, t);
        } else if (cl->diag && diag_facet_first_hid(c, cl, t)) {
            cl->quirk |= OC_Q_FACET_FIRST;
        }
    }
    return facet || vertex || edge;
}

/* millingcutter.cpp:422-433 */
static inline int overlaps(const oc_cutter* c, double x, double y, const Tri* t) {
    double r = c->radius;
    if (t->bb[1] < x - r) return 0;
    if (t->bb[0] > x + r) return 0;
    if (t->bb[3] < y - r) return 0;
    if (t->bb[2] > y + r) return 0;
    return 1;
}

/* ----------------------------------------------------------------- kdtree */
typedef struct {
    int dim;
    double cutval;
    int hi, lo;        /* child node index or -1 */
    int leaf_begin, leaf_end; /* item range, leaf_begin < 0 for internal */
} KDN;

typedef struct kdtree {
    KDN* nodes;
    int nn, cap;
    int* items;   /* triangle indices in leaf (rank) order */
    int dims[4];
    unsigned bucket;
    int root;
} KDT;

static const int DIMSETS[3][4] = {{0, 1, 2, 3}, {2, 3, 4, 5}, {0, 1, 4, 5}};

static int kd_new_node(KDT* k) {
    if (k->nn == k->cap) {
        k->cap = k->cap ? 2 * k->cap : 1024;
        k->nodes = (KDN*)realloc(k->nodes, sizeof(KDN) * (size_t)k->cap);
    }
    return k->nn++;
}

/* kdtree.hpp:141-268 on the item range [b, e), with scratch buffer */
static int kd_build_node(KDT* k, const Tri* tris, int b, int e, int* scratch) {
    int n = e - b;
    double maxv[6] = {0}, minv[6] = {0};
    int first = 1;
    for (int i = b; i < e; ++i) {
        const double* bb = tris[k->items[i]].bb;
        for (int m = 0; m < 4; ++m) {
            int pt = k->dims[m];
            double v = bb[pt];
            if (first) {
                maxv[pt] = v;
                minv[pt] = v;
                if (m == 3) first = 0;
            } else {
                if (maxv[pt] < v) maxv[pt] = v;
                if (minv[pt] > v) minv[pt] = v;
            }
        }
    }
    double mx = 0;
    int maxM = 0;
    for (int m = 0; m < 4; ++m) {
        double val = maxv[k->dims[m]] - minv[k->dims[m]];
        if (val > mx) { mx = val; maxM = m; }
    }
    int d = k->dims[maxM];
    double val = maxv[d] - minv[d];
    double start = minv[d];
    double cutvalue = start + val / 2;
    int id = kd_new_node(k);
    k->nodes[id].dim = d;
    k->nodes[id].cutval = cutvalue;
    k->nodes[id].hi = k->nodes[id].lo = -1;
    if (((unsigned)n <= k->bucket) || iszero_tol(val)) {
        k->nodes[id].leaf_begin = b;
        k->nodes[id].leaf_end = e;
        return id;
    }
    k->nodes[id].leaf_begin = -1;
    k->nodes[id].leaf_end = -1;
    /* stable partition: hi (bb[d] > cutvalue) first, then lo */
    int nh = 0, nl = 0;
    for (int i = b; i < e; ++i)
        if (tris[k->items[i]].bb[d] > cutvalue) scratch[nh++] = k->items[i];
    for (int i = b; i < e; ++i)
        if (!(tris[k->items[i]].bb[d] > cutvalue)) scratch[nh + nl++] = k->items[i];
    memcpy(k->items + b, scratch, sizeof(int) * (size_t)n);
    int hi = -1, lo = -1;
    if (nh > 0) hi = kd_build_node(k, tris, b, b + nh, scratch);
    if (nl > 0) lo = kd_build_node(k, tris, b + nh, e, scratch);
    k->nodes[id].hi = hi;
    k->nodes[id].lo = lo;
    return id;
}

static KDT* kd_build(const oc_mesh* m, int dimset, unsigned bucket) {
    KDT* k = (KDT*)calloc(1, sizeof(KDT));
    for (int i = 0; i < 4; ++i) k->dims[i] = DIMSETS[dimset][i];
    k->bucket = bucket;
    k->items = (int*)malloc(sizeof(int) * (size_t)(m->n > 0 ? m->n : 1));
    for (long i = 0; i < m->n; ++i) k->items[i] = (int)i;
    int* scratch = (int*)malloc(sizeof(int) * (size_t)(m->n > 0 ? m->n : 1));
    k->root = m->n > 0 ? kd_build_node(k, m->t, 0, (int)m->n, scratch) : -1;
    free(scratch);
    return k;
}

static void kd_free(KDT* k) {
    if (!k) return;
    free(k->nodes);
    free(k->items);
    free(k);
}

static KDT* mesh_kd(oc_mesh* m, int dimset, unsigned bucket) {
    KDT* k;
#pragma omp critical(oc_kd)
    {
        if (!m->kd[dimset] || m->kd_bucket[dimset] != bucket) {
            kd_free(m->kd[dimset]);
            m->kd[dimset] = kd_build(m, dimset, bucket);
            m->kd_bucket[dimset] = bucket;
        }
        k = m->kd[dimset];
    }
    return k;
}

typedef struct { int* v; long n, cap; } IVec;
static void ivec_push(IVec* a, int x) {
    if (a->n == a->cap) {
        a->cap = a->cap ? 2 * a->cap : 256;
        a->v = (int*)realloc(a->v, sizeof(int) * (size_t)a->cap);
    }
    a->v[a->n++] = x;
}

/* kdtree.hpp:273-304 */
static void kd_search_node(const KDT* k, const double* bb, int id, IVec* out) {
    const KDN* node = &k->nodes[id];
    if (node->leaf_begin >= 0) {
        for (int i = node->leaf_begin; i < node->leaf_end; ++i) ivec_push(out, k->items[i]);
        return;
    } else if ((node->dim % 2) == 0) {
        int maxdim = node->dim + 1;
        if (node->cutval > bb[maxdim]) {
            if (node->lo >= 0) kd_search_node(k, bb, node->lo, out);
        } else {
            if (node->hi >= 0) kd_search_node(k, bb, node->hi, out);
            if (node->lo >= 0) kd_search_node(k, bb, node->lo, out);
        }
    } else {
        int mindim = node->dim - 1;
        if (node->cutval < bb[mindim]) {
            if (node->hi >= 0) kd_search_node(k, bb, node->hi, out);
        } else {
            if (node->hi >= 0) kd_search_node(k, bb, node->hi, out);
            if (node->lo >= 0) kd_search_node(k, bb, node->lo, out);
        }
    }
}

int oc_kd_rank(const oc_mesh* m, int dimset, unsigned bucket, int* rank_out) {
    if (dimset < 0 || dimset > 2) return -1;
    KDT* k = kd_build(m, dimset, bucket);
    for (long i = 0; i < m->n; ++i) rank_out[k->items[i]] = (int)i;
    kd_free(k);
    return 0;
}

/* ------------------------------------------------------------------- mesh */
oc_mesh* oc_mesh_create(const double* xyz9, long ntri) {
    oc_mesh* m = (oc_mesh*)calloc(1, sizeof(oc_mesh));
    m->n = ntri;
    m->t = (Tri*)malloc(sizeof(Tri) * (size_t)(ntri > 0 ? ntri : 1));
    for (long i = 0; i < ntri; ++i) {
        const double* q = xyz9 + 9 * i;
        tri_init(&m->t[i], p3(q[0], q[1], q[2]), p3(q[3], q[4], q[5]), p3(q[6], q[7], q[8]));
    }
    return m;
}
void oc_mesh_free(oc_mesh* m) {
    if (!m) return;
    for (int i = 0; i < 3; ++i) kd_free(m->kd[i]);
    free(m->t);
    free(m);
}
long oc_mesh_size(const oc_mesh* m) { return m->n; }

long oc_kd_search_drop(oc_mesh* m, const oc_cutter* c, double x, double y, double z,
                       unsigned bucket, int* out, long cap) {
    if (m->n == 0) return 0;
    KDT* k = mesh_kd(m, 0, bucket);
    double r = c->radius;
    double bb[6] = {x - r, x + r, y - r, y + r, z, z + c->length};
    IVec v = {0, 0, 0};
    kd_search_node(k, bb, k->root, &v);
    long n = v.n;
    if (n <= cap) memcpy(out, v.v, sizeof(int) * (size_t)n);
    free(v.v);
    return n <= cap ? n : -n;
}

/* KD search of one push fiber (batchpushcutter.cpp:160-170): dimset YZ for
   X fibers (dir 0), XZ for Y fibers (dir 1) */
long oc_kd_search_push(oc_mesh* m, const oc_cutter* c, int dir, const double* fib6, unsigned bucket, int* out,
                       long cap) {
    if (m->n == 0) return 0;
    KDT* k = mesh_kd(m, dir == 0 ? 1 : 2, bucket);
    double r = c->radius;
    double clx = dir == 0 ? 0 : fib6[0], cly = dir == 0 ? fib6[1] : 0, clz = fib6[2];
    double bb[6] = {clx - r, clx + r, cly - r, cly + r, clz, clz + c->length};
    IVec v = {0, 0, 0};
    kd_search_node(k, bb, k->root, &v);
    long n = v.n;
    if (n <= cap) memcpy(out, v.v, sizeof(int) * (size_t)n);
    free(v.v);
    return n <= cap ? n : -n;
}

/* ------------------------------------------------------------ batch drop */
typedef struct { int key; int tri; } KR;
static int kr_cmp(const void* a, const void* b) {
    int x = ((const KR*)a)->key, y = ((const KR*)b)->key;
    return (x > y) - (x < y);
}

/* uniform XY grid used by search_mode 1 */
typedef struct {
    double x0, y0, h;
    int nx, ny;
    long* start; /* nx*ny+1 */
    int* items;
} Grid;

static Grid* grid_build(const oc_mesh* m, double h) {
    Grid* g = (Grid*)calloc(1, sizeof(Grid));
    double mnx = 1e300, mny = 1e300, mxx = -1e300, mxy = -1e300;
    for (long i = 0; i < m->n; ++i) {
        const double* bb = m->t[i].bb;
        if (bb[0] < mnx) mnx = bb[0];
        if (bb[2] < mny) mny = bb[2];
        if (bb[1] > mxx) mxx = bb[1];
        if (bb[3] > mxy) mxy = bb[3];
    }
    g->x0 = mnx; g->y0 = mny;
    double ext = (mxx - mnx) > (mxy - mny) ? (mxx - mnx) : (mxy - mny);
    if (h < ext / 2048.0) h = ext / 2048.0;
    if (!(h > 0)) h = 1.0;
    g->h = h;
    g->nx = (int)((mxx - mnx) / g->h) + 1;
    g->ny = (int)((mxy - mny) / g->h) + 1;
    long nc = (long)g->nx * g->ny;
    g->start = (long*)calloc((size_t)nc + 1, sizeof(long));
    for (int pass = 0; pass < 2; ++pass) {
        for (long i = 0; i < m->n; ++i) {
            const double* bb = m->t[i].bb;
            int ix0 = (int)((bb[0] - g->x0) / g->h), ix1 = (int)((bb[1] - g->x0) / g->h);
            int iy0 = (int)((bb[2] - g->y0) / g->h), iy1 = (int)((bb[3] - g->y0) / g->h);
            if (ix1 >= g->nx) ix1 = g->nx - 1;
            if (iy1 >= g->ny) iy1 = g->ny - 1;
            for (int iy = iy0; iy <= iy1; ++iy)
                for (int ix = ix0; ix <= ix1; ++ix) {
                    long cidx = (long)iy * g->nx + ix;
                    if (pass == 0) g->start[cidx + 1]++;
                    else g->items[g->start[cidx]++] = (int)i;
                }
        }
        if (pass == 0) {
            for (long cidx = 0; cidx < nc; ++cidx) g->start[cidx + 1] += g->start[cidx];
            g->items = (int*)malloc(sizeof(int) * (size_t)(g->start[nc] > 0 ? g->start[nc] : 1));
        } else {
            for (long cidx = nc; cidx > 0; --cidx) g->start[cidx] = g->start[cidx - 1];
            g->start[0] = 0;
        }
    }
    return g;
}
static void grid_free(Grid* g) {
    if (!g) return;
    free(g->start);
    free(g->items);
    free(g);
}

int oc_drop_triangle_diag(const oc_cutter* c, const double* q, double x, double y,
                          double* z_io, double* cc_out, int* cctype_out, int* feat_out, int* quirk_out) {
    Tri t;
    tri_init(&t, p3(q[0], q[1], q[2]), p3(q[3], q[4], q[5]), p3(q[6], q[7], q[8]));
    CL cl;
    memset(&cl, 0, sizeof(cl));
    cl.x = x; cl.y = y; cl.z = *z_io;
    cl.type = OC_NONE; cl.tri = -1; cl.feat = -1; cl.cur_tri = 0;
    cl.diag = quirk_out != NULL;
    int r = drop_cutter(c, &cl, &t);
    *z_io = cl.z;
    if (cc_out) { cc_out[0] = cl.cc.x; cc_out[1] = cl.cc.y; cc_out[2] = cl.cc.z; }
    if (cctype_out) *cctype_out = cl.type;
    if (feat_out) *feat_out = cl.feat;
    if (quirk_out) *quirk_out = cl.quirk;
    return r;
}

int oc_drop_triangle(const oc_cutter* c, const double* q, double x, double y,
                     double* z_io, double* cc_out, int* cctype_out, int* feat_out) {
    return oc_drop_triangle_diag(c, q, x, y, z_io, cc_out, cctype_out, feat_out, NULL);
}

int oc_batch_drop_diag(oc_mesh* m, const oc_cutter* c, const double* cl_xyz, long n,
                       int search_mode, int nthreads, unsigned bucket,
                       double* z_out, double* cc_out, int* cctype_out,
                       int* tri_out, int* feat_out, long* ncalls_out, int* quirk_out) {
    if (bucket < 1) bucket = 1;
    KDT* k = NULL;
    Grid* g = NULL;
    int* rank = NULL;
    if (m->n > 0 && search_mode == 0) k = mesh_kd(m, 0, bucket);
    if (m->n > 0 && search_mode == 1) {
        k = mesh_kd(m, 0, bucket);
        rank = (int*)malloc(sizeof(int) * (size_t)m->n);
        for (long i = 0; i < m->n; ++i) rank[k->items[i]] = (int)i;
        double h = c->radius > 0 ? c->radius : 1.0;
        g = grid_build(m, h);
    }
    long calls = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel reduction(+ : calls)
    {
        IVec cand = {0, 0, 0};
        KR* kr = NULL;
        long krcap = 0;
        int* stamp = NULL;
        if (g) stamp = (int*)calloc((size_t)m->n, sizeof(int));
        int stampv = 0;
#pragma omp for schedule(dynamic)
        for (long i = 0; i < n; ++i) {
            CL cl;
            memset(&cl, 0, sizeof(cl));
            cl.x = cl_xyz[3 * i]; cl.y = cl_xyz[3 * i + 1]; cl.z = cl_xyz[3 * i + 2];
            cl.type = OC_NONE; cl.tri = -1; cl.feat = -1;
            cl.cc = p3(0, 0, 0);
            cl.diag = quirk_out != NULL;
            if (m->n > 0) {
                cand.n = 0;
                if (search_mode == 0) {
                    double r = c->radius;
                    double bb[6] = {cl.x - r, cl.x + r, cl.y - r, cl.y + r, cl.z, cl.z + c->length};
                    kd_search_node(k, bb, k->root, &cand);
                } else if (search_mode == 1) {
                    double r = c->radius;
                    int ix0 = (int)floor((cl.x - r - g->x0) / g->h), ix1 = (int)floor((cl.x + r - g->x0) / g->h);
                    int iy0 = (int)floor((cl.y - r - g->y0) / g->h), iy1 = (int)floor((cl.y + r - g->y0) / g->h);
                    if (ix0 < 0) ix0 = 0;
                    if (iy0 < 0) iy0 = 0;
                    if (ix1 >= g->nx) ix1 = g->nx - 1;
                    if (iy1 >= g->ny) iy1 = g->ny - 1;
                    ++stampv;
                    long nk = 0;
                    for (int iy = iy0; iy <= iy1; ++iy)
                        for (int ix = ix0; ix <= ix1; ++ix) {
                            long cidx = (long)iy * g->nx + ix;
                            for (long j = g->start[cidx]; j < g->start[cidx + 1]; ++j) {
                                int ti = g->items[j];
                                if (stamp[ti] == stampv) continue;
                                stamp[ti] = stampv;
                                if (!overlaps(c, cl.x, cl.y, &m->t[ti])) continue;
                                if (nk == krcap) {
                                    krcap = krcap ? 2 * krcap : 1024;
                                    kr = (KR*)realloc(kr, sizeof(KR) * (size_t)krcap);
                                }
                                kr[nk].key = rank[ti];
                                kr[nk].tri = ti;
                                ++nk;
                            }
                        }
                    qsort(kr, (size_t)nk, sizeof(KR), kr_cmp);
                    for (long j = 0; j < nk; ++j) ivec_push(&cand, kr[j].tri);
                } else {
                    for (long j = 0; j < m->n; ++j) ivec_push(&cand, (int)j);
                }
                for (long j = 0; j < cand.n; ++j) {
                    const Tri* t = &m->t[cand.v[j]];
                    cl.cur_tri = cand.v[j];
                    if (search_mode == 2) {
                        drop_cutter(c, &cl, t);
                        ++calls;
                    } else if (overlaps(c, cl.x, cl.y, t)) {
                        if (cl_below(&cl, t)) {
                            drop_cutter(c, &cl, t);
                            ++calls;
                        }
                    }
                }
            }
            z_out[i] = cl.z;
            if (cc_out) { cc_out[3 * i] = cl.cc.x; cc_out[3 * i + 1] = cl.cc.y; cc_out[3 * i + 2] = cl.cc.z; }
            if (cctype_out) cctype_out[i] = cl.type;
            if (tri_out) tri_out[i] = cl.tri;
            if (feat_out) feat_out[i] = cl.feat;
            if (quirk_out) quirk_out[i] = cl.quirk;
        }
        free(cand.v);
        free(kr);
        free(stamp);
    }
    grid_free(g);
    free(rank);
    if (ncalls_out) *ncalls_out = calls;
    return 0;
}

int oc_batch_drop(oc_mesh* m, const oc_cutter* c, const double* cl_xyz, long n,
                  int search_mode, int nthreads, unsigned bucket,
                  double* z_out, double* cc_out, int* cctype_out,
                  int* tri_out, int* feat_out, long* ncalls_out) {
    return oc_batch_drop_diag(m, c, cl_xyz, n, search_mode, nthreads, bucket, z_out, cc_out, cctype_out, tri_out,
                              feat_out, ncalls_out, NULL);
}

/* ------------------------------------------------------------------- push */
typedef struct {
    double lower, upper;
    P3 lcc, ucc;
    int ltype, utype;
} Ival;

static void ival_init(Ival* i) {
    i->lower = 0.0; i->upper = 0.0;
    i->lcc = p3(0, 0, 0); i->ucc = p3(0, 0, 0);
    i->ltype = OC_NONE; i->utype = OC_NONE;
}
/* interval.cpp:61-76 */
static void ival_update_upper(Ival* i, double t, P3 p, int type) {
    if (i->utype == OC_NONE) {
        i->upper = t; i->lower = t;
        i->ucc = p; i->utype = type;
        i->lcc = p; i->ltype = type;
    }
    if (t > i->upper) { i->upper = t; i->ucc = p; i->utype = type; }
}
/* interval.cpp:78-93 */
static void ival_update_lower(Ival* i, double t, P3 p, int type) {
    if (i->ltype == OC_NONE) {
        i->lower = t; i->upper = t;
        i->lcc = p; i->ltype = type;
        i->ucc = p; i->utype = type;
    }
    if (t < i->lower) { i->lower = t; i->lcc = p; i->ltype = type; }
}
/* interval.cpp:46-49 */
static void ival_update(Ival* i, double t, P3 p, int type) {
    ival_update_upper(i, t, p, type);
    ival_update_lower(i, t, p, type);
}
/* interval.cpp:51-59 */
static int ival_update_if(Ival* i, double t, P3 cc, int type, P3 p1, P3 p2, int cond) {
    if (inside_points(cc, p1, p2) && cond) {
        ival_update(i, t, cc, type);
        return 1;
    }
    return 0;
}
static int ival_outside(const Ival* a, const Ival* b) { /* interval.cpp:96-103 */
    if (a->lower > b->upper) return 1;
    if (a->upper < b->lower) return 1;
    return 0;
}
static int ival_inside(const Ival* a, const Ival* b) { /* interval.cpp:106-111 */
    return (a->lower > b->lower) && (a->upper < b->upper);
}
static int ival_empty(const Ival* a) { return a->lower == 0.0 && a->upper == 0.0; }

/* fiber.cpp:96-101 */
static double fib_tval(const Fib* f, P3 p) {
    P3 d = psub(f->p2, f->p1);
    return pdot(psub(p, f->p1), d) / pdot(d, d);
}
/* fiber.cpp:103-106 */
static P3 fib_point(const Fib* f, double t) { return padd(f->p1, pmul(t, psub(f->p2, f->p1))); }

/* numeric.cpp:118-140 */
static int two_by_two(double a, double b, double c, double d, double e, double f, double* u, double* v) {
    double det = a * d - c * b;
    if (iszero_tol(det)) return 0;
    *u = (1.0 / det) * (d * e - b * f);
    *v = (1.0 / det) * (-c * e + a * f);
    return 1;
}
/* numeric.cpp:147-154 */
static int xy_line_line(P3 p1, P3 p2, double* v, P3 p3_, P3 p4, double* t) {
    return two_by_two(p2.x - p1.x, -(p4.x - p3_.x), p2.y - p1.y, -(p4.y - p3_.y),
                      (p3_.x - p1.x), p3_.y - p1.y, v, t);
}

/* millingcutter.cpp:148-167 */
static int single_vertex_push(const oc_cutter* c, const Fib* f, Ival* i, P3 p, int type) {
    if ((p.z >= f->p1.z) && (p.z <= (f->p1.z + c->length))) {
        P3 pq = xy_closest_point(p, f->p1, f->p2);
        double q = pxynorm(psub(p, pq));
        double h = p.z - f->p1.z;
        double cw = cutter_width(c, h);
        if (q <= cw) {
            double ofs = sqrt(sq(cw) - sq(q));
            P3 start = psub(pq, pmul(ofs, f->dir));
            P3 stop = padd(pq, pmul(ofs, f->dir));
            ival_update_upper(i, fib_tval(f, stop), p, type);
            ival_update_lower(i, fib_tval(f, start), p, type);
            return 1;
        }
    }
    return 0;
}

/* triangle.cpp:77-128 */
static int zslice_verts(const Tri* t, P3* p1, P3* p2, double zcut) {
    if ((zcut <= t->bb[4]) || (zcut >= t->bb[5])) return 0;
    P3 below[3], above[3];
    int nb = 0, na = 0;
    for (int m = 0; m < 3; ++m) {
        if (t->p[m].z <= zcut) below[nb++] = t->p[m];
        else above[na++] = t->p[m];
    }
    if (nb == 2) {
        double t1 = (zcut - above[0].z) / (below[0].z - above[0].z);
        double t2 = (zcut - above[0].z) / (below[1].z - above[0].z);
        *p1 = padd(above[0], pmul(t1, psub(below[0], above[0])));
        *p2 = padd(above[0], pmul(t2, psub(below[1], above[0])));
        return 1;
    } else if (nb == 1) {
        double t1 = (zcut - above[0].z) / (below[0].z - above[0].z);
        double t2 = (zcut - above[1].z) / (below[0].z - above[1].z);
        *p1 = padd(above[0], pmul(t1, psub(below[0], above[0])));
        *p2 = padd(above[1], pmul(t2, psub(below[0], above[1])));
        return 1;
    }
    return 0;
}

static int vertex_push(const oc_cutter* c, const Fib* f, Ival* i, const Tri* t);
static int facet_push(const oc_cutter* c, const Fib* f, Ival* i, const Tri* t);
static int edge_push(const oc_cutter* c, const Fib* f, Ival* i, const Tri* t);

/* millingcutter.cpp:139-146, cylcutter.cpp:73-91 */
static int base_vertex_push(const oc_cutter* c, const Fib* f, Ival* i, const Tri* t) {
    int r = 0;
    for (int k = 0; k < 3; ++k)
        if (single_vertex_push(c, f, i, t->p[k], OC_VERTEX)) r = 1;
    if (c->kind == OC_CYL) {
        P3 a, b;
        if (zslice_verts(t, &a, &b, f->p1.z)) {
            a.z = f->p1.z;
            b.z = f->p1.z;
            if (single_vertex_push(c, f, i, a, OC_VERTEX_CYL)) r = 1;
            if (single_vertex_push(c, f, i, b, OC_VERTEX_CYL)) r = 1;
        }
    }
    return r;
}

/* millingcutter.cpp:177-275 */
static int general_facet_push(double normal_length, double center_height, double xy_normal_length,
                              const Fib* fib, Ival* i, const Tri* t) {
    P3 normal = up_normal(t);
    if (normal.x == 0.0 && normal.y == 0.0) return 0; /* zParallel, point.cpp:316-323 */
    normal = pnormalize(normal);
    P3 xyn = normal;
    xyn.z = 0;
    xyn = pxynormalize(xyn);
    double a, b, e, u, v;
    double c = t->p[1].z - t->p[0].z;
    double d = t->p[2].z - t->p[0].z;
    double f = -t->p[0].z - normal_length * normal.z + fib->p1.z + center_height;
    if (fib->p1.y == fib->p2.y) {
        a = t->p[1].y - t->p[0].y;
        b = t->p[2].y - t->p[0].y;
        e = -t->p[0].y - normal_length * normal.y - xy_normal_length * xyn.y + fib->p1.y;
        if (!two_by_two(a, b, c, d, e, f, &u, &v)) return 0;
        P3 cc = padd(padd(t->p[0], pmul(u, psub(t->p[1], t->p[0]))), pmul(v, psub(t->p[2], t->p[0])));
        if (!inside_tri(cc, t)) return 0;
        double tval = (1.0 / (fib->p2.x - fib->p1.x)) *
                      (t->p[0].x + normal_length * normal.x + xy_normal_length * xyn.x - fib->p1.x +
                       u * (t->p[1].x - t->p[0].x) + v * (t->p[2].x - t->p[0].x));
        ival_update(i, tval, cc, OC_FACET);
        return 1;
    } else if (fib->p1.x == fib->p2.x) {
        a = t->p[1].x - t->p[0].x;
        b = t->p[2].x - t->p[0].x;
        e = -t->p[0].x - normal_length * normal.x - xy_normal_length * xyn.x + fib->p1.x;
        if (!two_by_two(a, b, c, d, e, f, &u, &v)) return 0;
        P3 cc = padd(padd(t->p[0], pmul(u, psub(t->p[1], t->p[0]))), pmul(v, psub(t->p[2], t->p[0])));
        if (!inside_tri(cc, t)) return 0;
        double tval = (1.0 / (fib->p2.y - fib->p1.y)) *
                      (t->p[0].y + normal_length * normal.y + xy_normal_length * xyn.y - fib->p1.y +
                       u * (t->p[1].y - t->p[0].y) + v * (t->p[2].y - t->p[0].y));
        ival_update(i, tval, cc, OC_FACET);
        return 1;
    }
    return 0;
}

/* millingcutter.cpp:370-375 */
static int calc_cc_update(double t, double u, P3 q, P3 p1, P3 p2, const Fib* f, Ival* i,
                          double height, int type) {
    P3 cc = padd(q, pmul(u, psub(p2, p1)));
    return ival_update_if(i, t, cc, type, p1, p2, cc.z >= height);
}

/* millingcutter.cpp:335-368 */
static int horiz_edge_push(const oc_cutter* c, const Fib* f, Ival* i, P3 p1, P3 p2) {
    int r = 0;
    double h = p1.z - f->p1.z;
    if (h > 0.0) {
        if (iszero_tol(p2.z - p1.z)) {
            double eff = cutter_width(c, h);
            double qt, qv;
            if (xy_line_line(p1, p2, &qv, f->p1, f->p2, &qt)) {
                P3 q = padd(p1, pmul(qv, psub(p2, p1)));
                double ccv, clt;
                P3 tang = psub(p2, p1);
                tang.z = 0;
                tang = pxynormalize(tang);
                P3 nrm = pxyperp(tang);
                P3 q1 = padd(q, pmul(eff, nrm));
                P3 q2 = padd(q1, psub(p2, p1));
                if (xy_line_line(q1, q2, &ccv, f->p1, f->p2, &clt)) {
                    double t1 = clt;
                    double t2 = qt + (qt - clt);
                    if (calc_cc_update(t1, ccv, q, p1, p2, f, i, f->p1.z, OC_EDGE_HORIZ)) r = 1;
                    if (calc_cc_update(t2, -ccv, q, p1, p2, f, i, f->p1.z, OC_EDGE_HORIZ)) r = 1;
                }
            }
        }
    }
    return r;
}

/* millingcutter.cpp:304-332 */
static int shaft_edge_push(const oc_cutter* c, const Fib* f, Ival* i, P3 p1, P3 p2) {
    int r = 0;
    double u, v;
    if (xy_line_line(p1, p2, &u, f->p1, f->p2, &v)) {
        P3 q = padd(p1, pmul(u, psub(p2, p1)));
        P3 tang = psub(p2, p1);
        tang.z = 0;
        tang = pxynormalize(tang);
        P3 nrm = pxyperp(tang);
        P3 q1 = padd(q, pmul(c->radius, nrm));
        P3 q2 = padd(q1, psub(p2, p1));
        double u_cc, t_cl;
        if (xy_line_line(q1, q2, &u_cc, f->p1, f->p2, &t_cl)) {
            double t1 = t_cl;
            double t2 = v + (v - t_cl);
            if (calc_cc_update(t1, u_cc, q, p1, p2, f, i, f->p1.z + c->center_height, OC_EDGE_SHAFT)) r = 1;
            if (calc_cc_update(t2, -u_cc, q, p1, p2, f, i, f->p1.z + c->center_height, OC_EDGE_SHAFT)) r = 1;
        }
    }
    return r;
}

/* ballcutter.cpp:127-134 */
static int ball_cc_update(const oc_cutter* c, double t, P3 p1, P3 p2, const Fib* f, Ival* i) {
    P3 center = padd(fib_point(f, t), p3(0, 0, c->radius));
    P3 cc = closest_point(center, p1, p2);
    return ival_update_if(i, t, cc, OC_EDGE_BALL, p1, p2, (center.z - cc.z >= 0));
}

/* ballcutter.cpp:66-124 */
static int ball_general_edge_push(const oc_cutter* c, const Fib* f, Ival* i, P3 p1, P3 p2) {
    int r = 0;
    if (iszero_tol(pxynorm(psub(p2, p1)))) return 0;
    P3 ufp1 = padd(f->p1, p3(0, 0, c->radius));
    P3 ufp2 = padd(f->p2, p3(0, 0, c->radius));
    P3 ab = psub(p2, p1);
    P3 ao = psub(ufp1, p1);
    P3 ao_x_ab = pcross(ao, ab);
    P3 v_x_ab = pcross(psub(ufp2, ufp1), ab);
    double ab2 = pdot(ab, ab);
    double a = pdot(v_x_ab, v_x_ab);
    double b = 2 * (pdot(v_x_ab, ao_x_ab));
    double cc = pdot(ao_x_ab, ao_x_ab) - (c->radius * c->radius * ab2);
    double discr = b * b - 4 * a * cc;
    if (iszero_tol(discr)) {
        double t1 = -b / (2 * a);
        if (ball_cc_update(c, t1, p1, p2, f, i)) r = 1;
    } else if (discr > 0.0) {
        double t1 = (-b + sqrt(discr)) / (2 * a);
        double t2 = (-b - sqrt(discr)) / (2 * a);
        if (ball_cc_update(c, t1, p1, p2, f, i)) r = 1;
        if (ball_cc_update(c, t2, p1, p2, f, i)) r = 1;
    }
    return r;
}

/* bullcutter.cpp:109-155 */
static int bull_general_edge_push(const oc_cutter* c, const Fib* f, Ival* i, P3 p1, P3 p2) {
    int r = 0;
    if (iszero_tol(pxynorm(psub(p2, p1)))) return 0;
    if (iszero_tol(p2.z - p1.z)) return 0;
    double tplane = (f->p1.z + c->radius2 - p1.z) / (p2.z - p1.z);
    P3 ell_center = padd(p1, pmul(tplane, psub(p2, p1)));
    P3 major = psub(p2, p1);
    major.z = 0;
    major = pxynormalize(major);
    P3 minor = pxyperp(major);
    double theta = atan((p2.z - p1.z) / pxynorm(psub(p2, p1)));
    double major_length = fabs(c->radius2 / sin(theta));
    double minor_length = c->radius2;
    Ell e;
    memset(&e, 0, sizeof(e));
    e.aligned = 1;
    e.center = ell_center;
    e.a = major_length; e.b = minor_length; e.offset = c->radius1;
    e.major_dir = major; e.minor_dir = minor;
    if (ell_aligned_solver(&e, f)) {
        P3 pcc = ell_epoint(&e, e.pos1);
        P3 pcc2 = ell_epoint(&e, e.pos2);
        P3 cc = closest_point(pcc, p1, p2);
        P3 cc2 = closest_point(pcc2, p1, p2);
        P3 cl = psub(ell_oepoint(&e, e.pos1), p3(0, 0, c->center_height));
        P3 cl2 = psub(ell_oepoint(&e, e.pos2), p3(0, 0, c->center_height));
        double cl_t = fib_tval(f, cl);
        double cl_t2 = fib_tval(f, cl2);
        if (ival_update_if(i, cl_t, cc, OC_EDGE_POS, p1, p2, 1)) r = 1;
        if (ival_update_if(i, cl_t2, cc2, OC_EDGE_POS, p1, p2, 1)) r = 1;
    }
    return r;
}

/* conecutter.cpp:300-308 */
static int cone_circle_cc(const oc_cutter* c, double t, P3 p1, P3 p2, const Fib* f, Ival* i) {
    double t_cc = (f->p1.z + c->center_height - p1.z) / (p2.z - p1.z);
    P3 cc = padd(p1, pmul(t_cc, psub(p2, p1)));
    return ival_update_if(i, t, cc, OC_EDGE_CONE_BASE, p1, p2, 1);
}
/* conecutter.cpp:279-295 */
static int cone_cone_cc(P3 tang, P3 tip, P3 base, P3 p1, P3 p2, const Fib* f, Ival* i) {
    double u, t;
    if (xy_line_line(f->p1, f->p2, &u, tang, tip, &t)) {
        if ((t >= 0.0) && (t <= 1.0)) {
            P3 cc = padd(base, pmul(t, psub(tip, base)));
            cc.z = z_project_onto_edge(cc, p1, p2);
            return ival_update_if(i, u, cc, OC_EDGE_CONE, p1, p2, 1);
        }
    }
    return 0;
}
/* conecutter.cpp:150-275 */
static int cone_general_edge_push(const oc_cutter* c, const Fib* f, Ival* i, P3 p1, P3 p2) {
    int r = 0;
    if (iszero_tol(p2.z - p1.z)) return 0;
    double t_tip = (f->p1.z - p1.z) / (p2.z - p1.z);
    P3 p_tip = padd(p1, pmul(t_tip, psub(p2, p1)));
    double t_base = (f->p1.z + c->center_height - p1.z) / (p2.z - p1.z);
    P3 p_base = padd(p1, pmul(t_base, psub(p2, p1)));
    p_base.z = f->p1.z;
    double L = pxynorm(psub(p_base, p_tip));
    double d = xy_distance_to_line(p_base, f->p1, f->p2);
    if (d <= c->radius) {
        double dx = f->p2.x - f->p1.x;
        double dy = f->p2.y - f->p1.y;
        double dr = sqrt(sq(dx) + sq(dy));
        double det = (f->p1.x - p_base.x) * (f->p2.y - p_base.y) - (f->p2.x - p_base.x) * (f->p1.y - p_base.y);
        double discr = sq(c->radius) * sq(dr) - sq(det);
        if (discr >= 0.0) {
            if (discr == 0.0) {
                double xt = (det * dy) / sq(dr);
                double yt = -(det * dx) / sq(dr);
                P3 pt = p3(xt + p_base.x, yt + p_base.y, 0);
                if (cone_circle_cc(c, fib_tval(f, pt), p1, p2, f, i)) r = 1;
            } else {
                double xp = (det * dy + sgn(dy) * dx * sqrt(discr)) / sq(dr);
                double yp = (-det * dx + fabs(dy) * sqrt(discr)) / sq(dr);
                P3 clp = p3(xp + p_base.x, yp + p_base.y, 0);
                double tp = fib_tval(f, clp);
                double xn = (det * dy - sgn(dy) * dx * sqrt(discr)) / sq(dr);
                double yn = (-det * dx - fabs(dy) * sqrt(discr)) / sq(dr);
                P3 cln = p3(xn + p_base.x, yn + p_base.y, 0);
                double tn = fib_tval(f, cln);
                if (cone_circle_cc(c, tp, p1, p2, f, i)) r = 1;
                if (cone_circle_cc(c, tn, p1, p2, f, i)) r = 1;
            }
        }
    }
    if (L > c->radius) {
        P3 p_mid = pmul(0.5, padd(p_base, p_tip));
        p_mid.z = f->p1.z;
        double r_tang = L / 2;
        P3 pd = psub(p_mid, p_base);
        pd.z = 0;
        double dist = pxynorm(pd);
        int case1 = (iszero_tol(dist) && iszero_tol(fabs(c->radius - r_tang)));
        int case2 = (dist > (c->radius + r_tang));
        int case3 = (dist < fabs(c->radius - r_tang));
        int case4 = (iszero_tol(dist - (c->radius + r_tang)));
        if (!(case1 || case2 || case3 || case4)) {
            double a = (-sq(r_tang) + sq(c->radius) + sq(dist)) / (2.0 * dist);
            P3 v2 = padd(p_base, pmul(a / dist, pd));
            double h = sqrt(sq(c->radius) - sq(a));
            P3 ofs = p3(-pd.y * (h / dist), pd.x * (h / dist), 0);
            P3 tang1 = padd(v2, ofs);
            P3 tang2 = psub(v2, ofs);
            if (cone_cone_cc(tang1, p_tip, p_base, p1, p2, f, i)) r = 1;
            if (cone_cone_cc(tang2, p_tip, p_base, p1, p2, f, i)) r = 1;
        }
    }
    return r;
}

static int general_edge_push(const oc_cutter* c, const Fib* f, Ival* i, P3 p1, P3 p2) {
    switch (c->kind) {
    case OC_BALL: return ball_general_edge_push(c, f, i, p1, p2);
    case OC_BULL: return bull_general_edge_push(c, f, i, p1, p2);
    case OC_CONE: return cone_general_edge_push(c, f, i, p1, p2);
    default: return 0;
    }
}

/* millingcutter.cpp:277-301 */
static int base_edge_push(const oc_cutter* c, const Fib* f, Ival* i, const Tri* t) {
    int r = 0;
    for (int n = 0; n < 3; n++) {
        P3 p1 = t->p[n], p2 = t->p[(n + 1) % 3];
        if (horiz_edge_push(c, f, i, p1, p2)) r = 1;
        else {
            if (shaft_edge_push(c, f, i, p1, p2)) r = 1;
            if (general_edge_push(c, f, i, p1, p2)) r = 1;
        }
    }
    return r;
}

/* compositecutter.cpp:196-270: run the sub-cutter method on a raised fiber,
   keep contacts whose cc height maps back to that sub-cutter */
/* parity diagnostic of the push (OC_Q_COMP_BELOW), per thread */
static __thread int g_push_quirk;

static int comp_push(const oc_cutter* c, const Fib* f, Ival* i, const Tri* t, int which) {
    int result = 0;
    double ct[2 * OC_MAX_SUB];
    P3 cp[2 * OC_MAX_SUB];
    int cty[2 * OC_MAX_SUB];
    int nc = 0;
    for (int n = 0; n < c->nsub; ++n) {
        Ival ci;
        ival_init(&ci);
        Fib cf = *f;
        cf.p1.z = f->p1.z + c->zoffset[n];
        cf.p2.z = f->p2.z + c->zoffset[n];
        int ok;
        if (which == 0) ok = vertex_push(c->sub[n], &cf, &ci, t);
        else if (which == 1) ok = facet_push(c->sub[n], &cf, &ci, t);
        else ok = edge_push(c->sub[n], &cf, &ci, t);
        if (ok) {
            if (n == comp_height_to_index(c, ci.ucc.z - f->p1.z)) {
                ct[nc] = ci.upper; cp[nc] = ci.ucc; cty[nc] = ci.utype; ++nc;
                if (ci.ucc.z - f->p1.z < -1E-6) g_push_quirk |= OC_Q_COMP_BELOW;
            }
            if (n == comp_height_to_index(c, ci.lcc.z - f->p1.z)) {
                ct[nc] = ci.lower; cp[nc] = ci.lcc; cty[nc] = ci.ltype; ++nc;
                if (ci.lcc.z - f->p1.z < -1E-6) g_push_quirk |= OC_Q_COMP_BELOW;
            }
        }
    }
    for (int n = 0; n < nc; ++n) {
        ival_update(i, ct[n], cp[n], cty[n]);
        result = 1;
    }
    return result;
}

static int vertex_push(const oc_cutter* c, const Fib* f, Ival* i, const Tri* t) {
    if (c->kind == OC_COMPOSITE) return comp_push(c, f, i, t, 0);
    return base_vertex_push(c, f, i, t);
}
/* millingcutter.cpp:169-175, conecutter.cpp:137-147 */
static int facet_push(const oc_cutter* c, const Fib* f, Ival* i, const Tri* t) {
    if (c->kind == OC_COMPOSITE) return comp_push(c, f, i, t, 1);
    if (c->kind == OC_CONE) {
        int r = 0;
        if (general_facet_push(0, 0, 0, f, i, t)) r = 1;
        if (general_facet_push(0, c->center_height, c->xy_normal_length, f, i, t)) r = 1;
        return r;
    }
    return general_facet_push(c->normal_length, c->center_height, c->xy_normal_length, f, i, t);
}
static int edge_push(const oc_cutter* c, const Fib* f, Ival* i, const Tri* t) {
    if (c->kind == OC_COMPOSITE) return comp_push(c, f, i, t, 2);
    return base_edge_push(c, f, i, t);
}
/* millingcutter.cpp:377-382 */
static int push_cutter(const oc_cutter* c, const Fib* f, Ival* i, const Tri* t) {
    int v = vertex_push(c, f, i, t);
    int fa = facet_push(c, f, i, t);
    int e = edge_push(c, f, i, t);
    return v || fa || e;
}

/* fiber.cpp:58-93 on a growable list */
typedef struct { Ival* v; long n, cap; } IvList;
static void ivl_push(IvList* l, const Ival* x) {
    if (l->n == l->cap) {
        l->cap = l->cap ? 2 * l->cap : 16;
        l->v = (Ival*)realloc(l->v, sizeof(Ival) * (size_t)l->cap);
    }
    l->v[l->n++] = *x;
}
static void fiber_add_interval(IvList* ints, const Ival* i) {
    if (ival_empty(i)) return;
    if (ints->n == 0) { ivl_push(ints, i); return; }
    for (long k = 0; k < ints->n; ++k) /* contains */
        if (ival_inside(i, &ints->v[k])) return;
    int missing = 1;
    for (long k = 0; k < ints->n; ++k)
        if (!ival_outside(i, &ints->v[k])) missing = 0;
    if (missing) { ivl_push(ints, i); return; }
    Ival ov[64];
    Ival* over = ov;
    long no = 0;
    if (ints->n + 1 > 64) over = (Ival*)malloc(sizeof(Ival) * (size_t)(ints->n + 1));
    long w = 0;
    for (long k = 0; k < ints->n; ++k) {
        if (!ival_outside(&ints->v[k], i)) over[no++] = ints->v[k];
        else ints->v[w++] = ints->v[k];
    }
    ints->n = w;
    over[no++] = *i;
    Ival sum;
    ival_init(&sum);
    for (long k = 0; k < no; ++k) {
        ival_update_lower(&sum, over[k].lower, over[k].lcc, over[k].ltype);
        ival_update_upper(&sum, over[k].upper, over[k].ucc, over[k].utype);
    }
    ivl_push(ints, &sum);
    if (over != ov) free(over);
}

static void fib_init(Fib* f, const double* q) {
    f->p1 = p3(q[0], q[1], q[2]);
    f->p2 = p3(q[3], q[4], q[5]);
    f->dir = pnormalize(psub(f->p2, f->p1));
}

static void ival_out(const Ival* i, double* o) {
    o[0] = i->lower; o[1] = i->upper;
    o[2] = i->lcc.x; o[3] = i->lcc.y; o[4] = i->lcc.z; o[5] = i->ltype;
    o[6] = i->ucc.x; o[7] = i->ucc.y; o[8] = i->ucc.z; o[9] = i->utype;
}
static void ival_in(Ival* i, const double* o) {
    i->lower = o[0]; i->upper = o[1];
    i->lcc = p3(o[2], o[3], o[4]); i->ltype = (int)o[5];
    i->ucc = p3(o[6], o[7], o[8]); i->utype = (int)o[9];
}

int oc_push_triangle_diag(const oc_cutter* c, const double* fib6, const double* q, double* out10, int* quirk_out) {
    Tri t;
    tri_init(&t, p3(q[0], q[1], q[2]), p3(q[3], q[4], q[5]), p3(q[6], q[7], q[8]));
    Fib f;
    fib_init(&f, fib6);
    Ival i;
    ival_init(&i);
    g_push_quirk = 0;
    push_cutter(c, &f, &i, &t);
    ival_out(&i, out10);
    if (quirk_out) *quirk_out = g_push_quirk;
    return !ival_empty(&i);
}

int oc_push_triangle(const oc_cutter* c, const double* fib6, const double* q, double* out10) {
    return oc_push_triangle_diag(c, fib6, q, out10, NULL);
}

long oc_fiber_union(const double* ints10, long n, long cap, double* out10) {
    IvList l = {0, 0, 0};
    for (long k = 0; k < n; ++k) {
        Ival i;
        ival_in(&i, ints10 + 10 * k);
        fiber_add_interval(&l, &i);
    }
    long r = l.n;
    if (r <= cap)
        for (long k = 0; k < r; ++k) ival_out(&l.v[k], out10 + 10 * k);
    free(l.v);
    return r <= cap ? r : -r;
}

long oc_batch_push(oc_mesh* m, const oc_cutter* c, int dir, const double* fib, long nf,
                   int search_mode, int nthreads, unsigned bucket,
                   long cap, long* offsets, double* lohi, double* cc, long* ncalls_out) {
    if (bucket < 1) bucket = 1;
    int dimset = dir == 0 ? 1 : 2;
    KDT* k = m->n > 0 ? mesh_kd(m, dimset, bucket) : NULL;
    IvList* res = (IvList*)calloc((size_t)(nf > 0 ? nf : 1), sizeof(IvList));
    long calls = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel reduction(+ : calls)
    {
        IVec cand = {0, 0, 0};
#pragma omp for schedule(dynamic)
        for (long n = 0; n < nf; ++n) {
            Fib f;
            fib_init(&f, fib + 6 * n);
            if (!k) continue;
            double clx, cly, clz = f.p1.z;
            if (dir == 0) { clx = 0; cly = f.p1.y; }
            else { clx = f.p1.x; cly = 0; }
            double r = c->radius;
            double bb[6] = {clx - r, clx + r, cly - r, cly + r, clz, clz + c->length};
            cand.n = 0;
            kd_search_node(k, bb, k->root, &cand);
            for (long j = 0; j < cand.n; ++j) {
                const Tri* t = &m->t[cand.v[j]];
                if (search_mode == 2) {
                    int d0 = k->dims[0], d2 = k->dims[2];
                    if (t->bb[d0 + 1] < bb[d0] || t->bb[d0] > bb[d0 + 1]) continue;
                    if (t->bb[d2 + 1] < bb[d2] || t->bb[d2] > bb[d2 + 1]) continue;
                }
                Ival i;
                ival_init(&i);
                push_cutter(c, &f, &i, t);
                fiber_add_interval(&res[n], &i);
                ++calls;
            }
        }
        free(cand.v);
    }
    long total = 0;
    for (long n = 0; n < nf; ++n) total += res[n].n;
    long ret = total;
    if (total <= cap) {
        long w = 0;
        for (long n = 0; n < nf; ++n) {
            offsets[n] = w;
            for (long j = 0; j < res[n].n; ++j) {
                double o[10];
                ival_out(&res[n].v[j], o);
                lohi[2 * w] = o[0];
                lohi[2 * w + 1] = o[1];
                for (int q = 0; q < 8; ++q) cc[8 * w + q] = o[2 + q];
                ++w;
            }
        }
        offsets[nf] = w;
    } else {
        ret = -total;
    }
    for (long n = 0; n < nf; ++n) free(res[n].v);
    free(res);
    if (ncalls_out) *ncalls_out = calls;
    return ret;
}

/* ============================================================ gradient */
/* Analytic derivative of the drop height at one CL point, given the active
   contact the forward pass recorded (triangle, feature code, CC point).  The
   reference has no gradient; this is the FP64 CPU check of the device
   backward kernel, and it is itself checked against central finite
   differences of oc_drop_triangle / oc_batch_drop (tests/test_drop_grad.py).

   With P(s) the contact on the feature (s: none for a vertex, the segment
   parameter for an edge, the barycentrics for a facet), the drop height is
   z = max_s P(s).z - h(q(s)) with q = |P_xy - cl_xy|, h the profile height:
     - smooth profile at q (ball, bull torus and flat, cone flank, cyl inside
       the disc): envelope theorem, dz/dtheta = partial of P.z - h(q) at the
       optimum, so dz/dx = h'(q) (P.x - x)/q, dz/dP.z = 1, dz/dP.xy = -dz/dxy,
       spread over the vertices by the contact's weights, dz/dr = -dh/dr(q);
     - contact on the rim q = r (cyl edges and facets, cone base): the
       constraint q^2 = r^2 is active, with multiplier
       lambda = (grad_s P.z . grad_s c)/|grad_s c|^2, c = q^2 - r^2, so
       dz/dx = 2 lambda (P.x - x), dz/dP.xy = -2 lambda (P - cl)_xy,
       dz/dr = -d h(r)/dr + 2 lambda r;
     - cone tip on a facet (q = 0): z is the plane height under the CL point;
     - cone edge in the reference's hyperbola case: its formula as written
       (cone_hyper_edge_grad), which for a cone with a shank is not the
       envelope of the cone.
   out = dz/dx, dz/dy, dz/dr, dz/dp0.xyz, dz/dp1.xyz, dz/dp2.xyz.
   Composite cutters: the derivatives of the active sub-cutter; dz/dr is NaN
   (no single radius).  Returns 1 when there is no contact. */
static void prof_slope(const oc_cutter* s, double q, double* hp, double* hr) {
    *hp = 0.0;
    *hr = 0.0;
    switch (s->kind) {
    case OC_BALL: {
        double w = sqrt(sq(s->radius) - sq(q));
        *hp = q / w;
        *hr = 1.0 - s->radius / w;
        break;
    }
    case OC_BULL:
        if (q > s->radius1) {
            double w = sqrt(sq(s->radius2) - sq(q - s->radius1));
            *hp = (q - s->radius1) / w;
            *hr = -*hp; /* d/d(radius) with the corner radius fixed */
        }
        break;
    case OC_CONE: *hp = 1.0 / tan(s->angle); break;
    default: break;
    }
}

static double rim_height_slope(const oc_cutter* s) {
    switch (s->kind) {
    case OC_BALL: return 1.0;
    case OC_CONE: return 1.0 / tan(s->angle);
    default: return 0.0;
    }
}

/* Forward-mode derivatives in 9 directions (x, y, r, edge start xyz, edge
   end xyz) of the reference's cone edge height in its hyperbola case
   (conecutter.cpp:109-135).  There z = a.z + M s + G d, with s and d the CL
   point's offset along and distance from the edge's XY line, M the edge's
   slope and G = K M^2 - sqrt(K^2 M^2 + 1) / tan(angle),
   K = r / sqrt(length^2 - r^2 M^2).  The formula takes the cutter's length
   (center height + shank l) where the cone's true contact takes its center
   height, so for l > 0 its CC point is not where the slopes match and z is
   not stationary in the contact parameter: the envelope theorem does not
   hold, and z is differentiated as written.  (For l = 0, G = -sqrt(ch^2 -
   r^2 M^2)/r and the result equals the envelope form.) */
typedef struct {
    double v, g[9];
} D9;
static D9 d9(double v, int i) {
    D9 r;
    r.v = v;
    for (int k = 0; k < 9; ++k) r.g[k] = (k == i) ? 1.0 : 0.0;
    return r;
}
static D9 d9_add(D9 a, D9 b) { for (int k = 0; k < 9; ++k) a.g[k] += b.g[k]; a.v += b.v; return a; }
static D9 d9_sub(D9 a, D9 b) { for (int k = 0; k < 9; ++k) a.g[k] -= b.g[k]; a.v -= b.v; return a; }
static D9 d9_mul(D9 a, D9 b) {
    D9 r;
    r.v = a.v * b.v;
    for (int k = 0; k < 9; ++k) r.g[k] = a.g[k] * b.v + a.v * b.g[k];
    return r;
}
static D9 d9_div(D9 a, D9 b) {
    D9 r;
    r.v = a.v / b.v;
    for (int k = 0; k < 9; ++k) r.g[k] = (a.g[k] * b.v - a.v * b.g[k]) / (b.v * b.v);
    return r;
}
static D9 d9_sqrt(D9 a) {
    D9 r;
    r.v = sqrt(a.v);
    for (int k = 0; k < 9; ++k) r.g[k] = a.g[k] / (2.0 * r.v);
    return r;
}
static D9 d9_scale(D9 a, double s) { for (int k = 0; k < 9; ++k) a.g[k] *= s; a.v *= s; return a; }

/* 1 when the reference's cone edge drop at (x, y) takes its hyperbola case
   for edge a -> b (conecutter.cpp:115-117) */
static int cone_edge_hyper(const oc_cutter* s, P3 a, P3 b, double x, double y) {
    const double ex = b.x - a.x, ey = b.y - a.y;
    const double lxy = sqrt(ex * ex + ey * ey);
    const double d = fabs((ex * (y - a.y) - ey * (x - a.x)) / lxy);
    const double m = (b.z - a.z) / lxy;
    const double xu = sqrt(sq(s->radius) - sq(d));
    const double mu = (s->center_height / s->radius) * xu / sqrt(sq(xu) + sq(d));
    return fabs(m) <= fabs(mu);
}

static void cone_hyper_edge_grad(const oc_cutter* s, P3 a, P3 b, double x, double y, double* g9) {
    const double tana = tan(s->angle);
    const double shank = s->length - s->radius / tana; /* l: fixed when r varies */
    const D9 X = d9(x, 0), Y = d9(y, 1), R = d9(s->radius, 2);
    const D9 ax = d9(a.x, 3), ay = d9(a.y, 4), az = d9(a.z, 5);
    const D9 bx = d9(b.x, 6), by = d9(b.y, 7), bz = d9(b.z, 8);
    const D9 ex = d9_sub(bx, ax), ey = d9_sub(by, ay);
    const D9 lxy = d9_sqrt(d9_add(d9_mul(ex, ex), d9_mul(ey, ey)));
    const D9 vx = d9_div(ex, lxy), vy = d9_div(ey, lxy);
    const D9 M = d9_div(d9_sub(bz, az), lxy);
    const D9 wx = d9_sub(X, ax), wy = d9_sub(Y, ay);
    const D9 sa = d9_add(d9_mul(wx, vx), d9_mul(wy, vy));
    D9 d = d9_sub(d9_mul(vx, wy), d9_mul(vy, wx));
    if (d.v < 0.0) d = d9_scale(d, -1.0);
    const D9 L = d9_add(d9_scale(R, 1.0 / tana), d9(shank, -1));
    const D9 M2 = d9_mul(M, M);
    const D9 K = d9_div(R, d9_sqrt(d9_sub(d9_mul(L, L), d9_mul(d9_mul(R, R), M2))));
    const D9 G = d9_sub(d9_mul(K, M2), d9_scale(d9_sqrt(d9_add(d9_mul(d9_mul(K, K), M2), d9(1.0, -1))), 1.0 / tana));
    const D9 z = d9_add(az, d9_add(d9_mul(M, sa), d9_mul(G, d)));
    for (int k = 0; k < 9; ++k) g9[k] = z.g[k];
}

int oc_drop_grad(const oc_cutter* c, const double* tri9, int feat, int cctype, double x, double y,
                 const double* cc3, double* out) {
    for (int k = 0; k < 12; ++k) out[k] = 0.0;
    if (cctype == OC_NONE || feat < 0) return 1;
    const int k = feat & 15;
    P3 V[3];
    for (int j = 0; j < 3; ++j) V[j] = p3(tri9[3 * j], tri9[3 * j + 1], tri9[3 * j + 2]);
    const P3 P = p3(cc3[0], cc3[1], cc3[2]);
    const double dx = P.x - x, dy = P.y - y;
    const double q = sqrt(dx * dx + dy * dy);
    /* composite: facets and edges carry their sub-cutter; a vertex contact
       uses the profile of the sub-cutter at its radius (compositecutter.cpp:58-61) */
    const oc_cutter* s = c;
    if (c->kind == OC_COMPOSITE)
        s = c->sub[(k >= F_VTX0 && k < F_EDGE0) ? comp_radius_to_index(c, q) : ((feat >> 4) & 3)];
    double w[3] = {0, 0, 0};
    double fs[2] = {0, 0}, cs[2] = {0, 0};
    int m = 0;
    if (k >= F_EDGE0) {
        const int a = k - F_EDGE0, b = (a + 1) % 3;
        const P3 e = psub(V[b], V[a]);
        const double t = ((P.x - V[a].x) * e.x + (P.y - V[a].y) * e.y) / (e.x * e.x + e.y * e.y);
        w[a] = 1.0 - t;
        w[b] = t;
        m = 1;
        fs[0] = e.z;
        cs[0] = 2.0 * (dx * e.x + dy * e.y);
    } else if (k >= F_VTX0) {
        w[k - F_VTX0] = 1.0;
    } else {
        const P3 e1 = psub(V[1], V[0]), e2 = psub(V[2], V[0]);
        const double det = e1.x * e2.y - e1.y * e2.x;
        const double px = P.x - V[0].x, py = P.y - V[0].y;
        const double u = (px * e2.y - py * e2.x) / det, v = (e1.x * py - e1.y * px) / det;
        w[0] = 1.0 - u - v;
        w[1] = u;
        w[2] = v;
        m = 2;
        fs[0] = e1.z;
        fs[1] = e2.z;
        cs[0] = 2.0 * (dx * e1.x + dy * e1.y);
        cs[1] = 2.0 * (dx * e2.x + dy * e2.y);
    }
    if (m == 1 && s->kind == OC_CONE) {
        const int a = k - F_EDGE0, b = (a + 1) % 3;
        if (cone_edge_hyper(s, V[a], V[b], x, y)) {
            double g9[9];
            cone_hyper_edge_grad(s, V[a], V[b], x, y, g9);
            out[0] = g9[0];
            out[1] = g9[1];
            out[2] = (c->kind == OC_COMPOSITE) ? NAN : g9[2];
            for (int j = 0; j < 3; ++j) {
                out[3 + 3 * a + j] = g9[3 + j];
                out[3 + 3 * b + j] = g9[6 + j];
            }
            return 0;
        }
    }
    double gx, gy, gr;
    double gxy_v[2]; /* dz/dP.xy, spread by w */
    const int tip = (s->kind == OC_CONE) && m == 2 && cctype == OC_FACET_TIP;
    /* rim contacts: every cyl edge or facet contact; the cone's base facet
       contact (its own feature code) and cone edge contacts on the base circle */
    const int rim = m > 0 && (s->kind == OC_CYL ||
                              (s->kind == OC_CONE && (m == 2 ? k == F_FACET_RIM : q >= s->radius * (1.0 - 1e-9))));
    if (tip) {
        /* plane height under the CL point: slope from the facet normal */
        const P3 n = pcross(psub(V[0], V[1]), psub(V[0], V[2]));
        gx = -n.x / n.z;
        gy = -n.y / n.z;
        gxy_v[0] = -gx;
        gxy_v[1] = -gy;
        gr = 0.0;
    } else if (rim) {
        const double den = cs[0] * cs[0] + cs[1] * cs[1];
        const double lam = den > 0.0 ? (fs[0] * cs[0] + fs[1] * cs[1]) / den : 0.0;
        gx = 2.0 * lam * dx;
        gy = 2.0 * lam * dy;
        gxy_v[0] = -gx;
        gxy_v[1] = -gy;
        gr = -rim_height_slope(s) + 2.0 * lam * s->radius;
    } else {
        double hp, hr;
        prof_slope(s, q, &hp, &hr);
        gx = q > 0.0 ? hp * dx / q : 0.0;
        gy = q > 0.0 ? hp * dy / q : 0.0;
        gxy_v[0] = -gx;
        gxy_v[1] = -gy;
        gr = -hr;
    }
    out[0] = gx;
    out[1] = gy;
    out[2] = (c->kind == OC_COMPOSITE) ? NAN : gr;
    for (int j = 0; j < 3; ++j) {
        out[3 + 3 * j] = gxy_v[0] * w[j];
        out[4 + 3 * j] = gxy_v[1] * w[j];
        out[5 + 3 * j] = w[j];
    }
    return 0;
}

/* oc_drop_grad over a batch (OpenMP): tri = input triangle index (-1: none). */
void oc_batch_drop_grad(const oc_mesh* m, const oc_cutter* c, const double* cl_xyz, long n, const int* tri,
                        const int* feat, const int* cctype, const double* cc, int nthreads, double* out12) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
    for (long i = 0; i < n; ++i) {
        double* o = out12 + 12 * i;
        if (tri[i] < 0) {
            for (int k = 0; k < 12; ++k) o[k] = 0.0;
            continue;
        }
        const Tri* t = &m->t[tri[i]];
        double q[9];
        for (int j = 0; j < 3; ++j) {
            q[3 * j] = t->p[j].x;
            q[3 * j + 1] = t->p[j].y;
            q[3 * j + 2] = t->p[j].z;
        }
        oc_drop_grad(c, q, feat[i], cctype[i], cl_xyz[3 * i], cl_xyz[3 * i + 1], cc + 3 * i, o);
    }
}
