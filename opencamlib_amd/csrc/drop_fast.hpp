// Selection pass of the drop-cutter: per-pair contact heights from
// precomputed per-triangle constants, in T = float (mixed / FP32 modes) or
// T = double (FP64 mode).
//
// The reference contact geometry (millingcutter.cpp:36-136 and the cutter
// overrides) is rewritten as affine functions of the CL position plus one
// square root per vertex/edge:
//   facet   tip_z = F0 + GX*x + GY*y, inside iff u,v,w > 0 with
//           u = U0 + U1*x + U2*y, v = V0 + V1*x + V2*y, w = 1-u-v
//           (the cutter offset of facetDrop is folded into F0, U0, V0);
//   vertex  m = R^2 - |p - cl|^2_xy >= 0, z = p.z - height(q);
//   edge    signed distance sd = C - EY*x + EX*y, m = R^2 - sd^2 >= 0,
//           canonical u1 = A - EX*x - EY*y, edge height Z1 + (u - u1)*M, the
//           per-cutter canonical contact (cylcutter.cpp:52-70,
//           ballcutter.cpp:50-63, bullcutter.cpp:82-105, conecutter.cpp:
//           109-135) and the segment test t = (cc_u - u1)*INVL in [0,1].
// x, y are relative to the level-0 box centre and z to the mesh mid height,
// so FP32 magnitudes stay near the cutter size.
//
// BullCutter edges: the reference solves the offset ellipse with Brent's
// method on a diamond angle.  With Q = sqrt(1+M^2) the same contact satisfies
//   d = tau * (r2 + r1*Q / sqrt(M^2 + tau^2)),   tau in [0, 1],
// (tau = -sin of the ellipse parameter), an increasing concave function of
// tau, solved here by Newton from a lower bound (monotone convergence; a
// truncated solve overestimates the contact height, so it can never hide the
// true winner).  The winning contact is always recomputed by the exact FP64
// Brent path.
//
// The selection keeps the reference's sequential semantics (strict '>' in
// rank order; a lifting facet skips the triangle's vertices and edges).  To
// decide when T cannot reproduce the FP64 outcome, each lane also tracks
//   - the highest contact of each CC type (near-ties between types),
//   - in FP32, the highest upper value (z + error bound) of every valid
//     contact other than the winner, of any type and triangle: a contact
//     within the winner's band may be the FP64 winner (an exact tie of two
//     contacts on a flat face, which the reference breaks by rank, or an
//     ill-conditioned edge whose error bound is large),
//   - the highest contact whose validity test lies within the error margin of
//     its threshold or whose height is ill-conditioned ("ambiguous"),
//   - whether the winner itself is ambiguous.
// Points where any of these can change the FP64 winner are escalated.
#pragma once

#include <cfloat>
#include <cmath>

#include "cutter.hpp"
#include "drop_exact.hpp"
#include "geom64.hpp"
#include "feature_ids.hpp"

namespace ocl_amd {

// ------------------------------------------------------------- layout
enum FastRec : int {
    R_BX0 = 0, R_BX1, R_BY0, R_BY1, R_BZ1,                        // xy box, max z (+ margin)
    R_F0, R_GX, R_GY, R_U0, R_U1, R_U2, R_V0, R_V1, R_V2,         // facet (rim for cone: below)
    R_F0B, R_U0B, R_V0B,                                          // cone rim facet
    R_P0X, R_P0Y, R_P0Z, R_P1X, R_P1Y, R_P1Z, R_P2X, R_P2Y, R_P2Z, // vertices
    R_E0 = 26,                                                    // 3 edges x kEdge values
    kEdge = 13,
    R_EBF = R_E0 + 3 * kEdge,  // facet barycentric margin
    R_EZF,                     // facet height error bound
    R_FLAGS,                   // see FL_* bits
    R_ZCAP,                    // cull cap: max z + largest error bound / slack of the contacts
    R_EMAX,                    // largest error bound of the triangle's contacts
    R_PMAR,                    // plane cap margin (see tri_cap), +inf: no plane cap
    kFastRec = 72,             // FP64 record: 576 B
    R_FID0 = 72,               // FP32 records only: feature IDs (bit patterns) of
                               // vertices 0..2 and edges 0..2 (feature_ids.hpp)
    kFastRec32 = 80            // FP32 record: 320 B (16-byte aligned)
};
// record length in T values
template <class T>
OCL_HD constexpr int rec_len() {
    return sizeof(T) == 4 ? (int)kFastRec32 : (int)kFastRec;
}
// a feature ID stored in, and read from, an FP32 record slot (bit pattern)
OCL_HD float id_bits(uint32_t v) {
    float f;
    __builtin_memcpy(&f, &v, 4);
    return f;
}
OCL_HD uint32_t id_of(float f) {
    uint32_t v;
    __builtin_memcpy(&v, &f, 4);
    return v;
}
// ID of a facet contact (tip/plane facet, or the cone's rim facet) of the
// triangle of rank k: unique to the triangle
OCL_HD uint32_t facet_id(uint32_t k, bool rim) { return 0x80000000u | (k << 1) | (rim ? 1u : 0u); }
// per-edge values: direction (EX,EY), line constants C and A, 1/length, slope
// M, start height Z1, three cutter constants K1..K3, the s^2 band below which
// the contact height is ill-conditioned (S2E), the height slack of an
// ambiguous contact (EAE) and the height error bound of the contact (EZE).
enum EdgeVal : int { E_EX = 0, E_EY, E_C, E_A, E_INVL, E_M, E_Z1, E_K1, E_K2, E_K3, E_S2E, E_EAE, E_EZE };
static_assert(R_PMAR < kFastRec && R_FID0 + 6 <= kFastRec32, "record overflow");

enum : int { FL_VERT = 1, FL_HORIZ = 2, FL_EDEG0 = 8, FL_EHOR0 = 64 };

// Local frame of a level-0 box: xy centre and mid height of its triangles.
struct BoxOrigin {
    double ox, oy, oz;
    float hx, hy;  // half extents of the level-0 box around (ox, oy), rounded up
};

// Cutter constants of the selection pass (T precision)
template <class T>
struct SelParams {
    int kind;
    T R, R2;        // radius, radius^2
    T r1, r2, r22, ir2;  // bull
    T cot;          // cone 1/tan(angle)
    T ch;           // cone centre height
    T LEN;          // cone length (its rim facet contact sits this far below the contact point)
    T DM;           // error bound of m = R^2 - q^2 (validity band |m| <= DM)
    T S2V;          // vertex: m below this -> height ill-conditioned
    T EAV;          // vertex: height slack of an ambiguous contact
    T EPOS;         // edge: position error of cc_u and u1 (segment band EPOS/L)
    T SQDM;         // sqrt(2 DM): worst-case error of s = sqrt(m) near m = 0
    T E;            // z error bound of a well-conditioned contact
    T HPAD;         // rounding allowance of the profile lower bound (culling)
    T CUT;          // cull margin: 2 x the largest contact error bound of the mesh + HPAD
    // composite cutters (drop_comp.hpp): per sub-cutter kind, radius,
    // BullCutter radii and radius window, for the culling bounds
    int ck[2], cwin[2];  // cwin: window test of a CylCutter sub-cutter's edge contacts (at distance R)
    int cflat[2];        // the sub-cutter's flat bottom has composite height exactly 0 (z offset 0)
    T cR[2], cr1[2], cr2[2], clo[2], chi[2];
    int cull;       // culling by profile bound enabled.  Cones cull against the
                    // running best only (their facet-first rule makes the
                    // result order dependent, so no floor of the final height
                    // and no reordering): a triangle whose contacts all stay
                    // below the running height at its turn cannot lift, and a
                    // triangle that does not lift changes nothing
};

// Error model shared by the host launcher and the record builder.  Lane
// coordinates are relative to a box centre, so |x|,|y| <= 2R + box size, and
// heights are relative to a wave frame (|z| <= zmag).
struct SelErr {
    double u, R, zmag;
    OCL_HD double dm() const { return 12.0 * u * R * R; }
    // tie threshold: bound on |z_T - z_FP64| of a well-conditioned contact
    OCL_HD double E() const { return 6.0 * u * (zmag + 2.0 * R); }

    // s^2 band where d(height)/ds * ds exceeds the contact's error bound e,
    // for a height slope k = dz/ds
    OCL_HD double s2band(double k, double e) const {
        const double q = k * dm() / e;
        return fmax(dm(), q * q);
    }
    OCL_HD double slack(double k, double e) const { return k * sqrt(2.0 * dm()) + 2.0 * e; }
};

// ------------------------------------------------------------- math
template <class T> struct FM;
template <> struct FM<float> {
    OCL_HD static float sqrt_(float a) {
#if defined(__HIP_DEVICE_COMPILE__)
        return __builtin_amdgcn_sqrtf(a);
#else
        return std::sqrt(a);
#endif
    }
    OCL_HD static float rsqrt_(float a) {
#if defined(__HIP_DEVICE_COMPILE__)
        return __builtin_amdgcn_rsqf(a);
#else
        return 1.0f / std::sqrt(a);
#endif
    }
    OCL_HD static float rcp_(float a) {
#if defined(__HIP_DEVICE_COMPILE__)
        return __builtin_amdgcn_rcpf(a);
#else
        return 1.0f / a;
#endif
    }
    OCL_HD static float fma_(float a, float b, float c) { return fmaf(a, b, c); }
    OCL_HD static float max_(float a, float b) { return fmaxf(a, b); }
    OCL_HD static float min_(float a, float b) { return fminf(a, b); }
    static constexpr float inf() { return INFINITY; }
    static constexpr double eps() { return 5.9604644775390625e-08; }  // 2^-24
};
template <> struct FM<double> {
    OCL_HD static double sqrt_(double a) { return sqrt(a); }
    OCL_HD static double rsqrt_(double a) { return 1.0 / sqrt(a); }
    OCL_HD static double rcp_(double a) { return 1.0 / a; }
    OCL_HD static double fma_(double a, double b, double c) { return fma(a, b, c); }
    OCL_HD static double max_(double a, double b) { return fmax(a, b); }
    OCL_HD static double min_(double a, double b) { return fmin(a, b); }
    static constexpr double inf() { return INFINITY; }
    static constexpr double eps() { return 1.1102230246251565e-16; }  // 2^-53
};

// ------------------------------------------------------ record builder
// Builds the T record of one triangle for one cutter, in FP64, from the
// exact record (vertices + reference normal), relative to the box origin
// (ox, oy, oz), with the error bounds of the facet and edge contacts.  fid:
// the triangle's 6 feature IDs (FP32 records; null: distinct IDs 0..5).
template <class T>
OCL_HD void build_record(const TriX& t, const CutterDev& cd, double ox, double oy, double zmid, const SelErr& er,
                         T* out, const uint32_t* fid = nullptr) {
    const double u = FM<T>::eps();
    const double Eabs = er.E();
    const CutterP& c = cd.c;
    const double R = c.radius;
    int flags = 0;
    V3 p[3];
    for (int k = 0; k < 3; ++k) p[k] = v3(t.p[k].x - ox, t.p[k].y - oy, t.p[k].z - zmid);
    double bx0 = p[0].x, bx1 = p[0].x, by0 = p[0].y, by1 = p[0].y, bz1 = p[0].z;
    for (int k = 1; k < 3; ++k) {
        bx0 = fmin(bx0, p[k].x); bx1 = fmax(bx1, p[k].x);
        by0 = fmin(by0, p[k].y); by1 = fmax(by1, p[k].y);
        bz1 = fmax(bz1, p[k].z);
    }
    // lanes overlapping this triangle lie within R of its box
    const double xyb = fmax(fmax(fabs(bx0), fabs(bx1)), fmax(fabs(by0), fabs(by1))) + R;
    const double pad = 4 * u * (xyb + R) + 1e-300;
    out[R_BX0] = (T)(bx0 - pad);
    out[R_BX1] = (T)(bx1 + pad);
    out[R_BY0] = (T)(by0 - pad);
    out[R_BY1] = (T)(by1 + pad);
    out[R_BZ1] = (T)(bz1 + Eabs);
    // ---- facet (millingcutter.cpp:51-75, conecutter.cpp:70-105)
    V3 n = up_normal(t);
    const bool vertical = near0(n.z);
    const bool horiz = !vertical && near0(n.x) && near0(n.y);
    // barycentric affine forms in local coordinates (point.cpp:224-239)
    const V3 a = p[0], b = p[1], cc = p[2];
    const double du = a.y * cc.x - a.x * cc.y + (cc.y - a.y) * b.x + (a.x - cc.x) * b.y;
    const double dv = a.x * b.y - a.y * b.x + (a.y - b.y) * cc.x + (b.x - a.x) * cc.y;
    double U0 = (a.y * cc.x - a.x * cc.y) / du, U1 = (cc.y - a.y) / du, U2 = (a.x - cc.x) / du;
    double V0 = (a.x * b.y - a.y * b.x) / dv, V1 = (a.y - b.y) / dv, V2 = (b.x - a.x) / dv;
    double F0 = 0, GX = 0, GY = 0, offx = 0, offy = 0, F0B = 0, offbx = 0, offby = 0;
    if (vertical || !(fabs(du) > 0) || !(fabs(dv) > 0)) {
        flags |= FL_VERT;
        U0 = -4; U1 = 0; U2 = 0; V0 = -4; V1 = 0; V2 = 0;
    } else if (horiz) {
        flags |= FL_HORIZ;
        F0 = p[0].z;
    } else {
        const double nz = n.z;
        const V3 nn = vunit(n);
        const V3 xyn = vxyunit(v3(nn.x, nn.y, 0.0));
        GX = -nn.x / nn.z;
        GY = -nn.y / nn.z;
        // plane through p0: z(x,y) = p0.z + GX (x - p0.x) + GY (y - p0.y)
        const double zc = p[0].z - GX * p[0].x - GY * p[0].y;
        if (c.kind == K_CONE) {
            // tip: cc = cl, z = plane(cl); rim: cc = cl - R*xyn, z = plane(cc) - length
            F0 = zc;
            offbx = R * xyn.x;
            offby = R * xyn.y;
            F0B = zc - GX * offbx - GY * offby - c.length;
        } else if (c.kind != K_COMPOSITE) {
            const V3 rv = vadd(vscale(c.xyl, xyn), vscale(c.nl, nn));
            offx = rv.x;
            offy = rv.y;
            F0 = zc - GX * offx - GY * offy + rv.z - c.ch;
        }
        (void)nz;
    }
    out[R_F0] = (T)F0;
    out[R_GX] = (T)GX;
    out[R_GY] = (T)GY;
    out[R_U0] = (T)(U0 - U1 * offx - U2 * offy);
    out[R_U1] = (T)U1;
    out[R_U2] = (T)U2;
    out[R_V0] = (T)(V0 - V1 * offx - V2 * offy);
    out[R_V1] = (T)V1;
    out[R_V2] = (T)V2;
    out[R_F0B] = (T)F0B;
    out[R_U0B] = (T)(U0 - U1 * offbx - U2 * offby);
    out[R_V0B] = (T)(V0 - V1 * offbx - V2 * offby);
    const double ebf = 4 * u * (fabs(U0) + fabs(V0) + 2.0 + (fabs(U1) + fabs(U2) + fabs(V1) + fabs(V2)) * xyb);
    out[R_EBF] = (T)ebf;
    // facet height error bound (steep facets amplify the position rounding)
    const double ezf = 4 * u * (fabs(F0) + fabs(F0B) + 2.0 * (fabs(GX) + fabs(GY)) * xyb) + Eabs;
    out[R_EZF] = (T)ezf;
    // largest error bound and ambiguity slack over the contacts (culling);
    // vertex contacts use the cutter-wide constants of sel_params
    double emax = fmax(ezf, fmax(Eabs, 2e-9));
    double slk = fmax(emax, fmax(er.slack(1.0, Eabs), 2.0 * Eabs));
    // ---- vertices (heights relative to the cutter-specific base)
    for (int k = 0; k < 3; ++k) {
        out[R_P0X + 3 * k] = (T)p[k].x;
        out[R_P0Y + 3 * k] = (T)p[k].y;
        double zb = p[k].z;
        if (c.kind == K_BALL) zb -= R;
        if (c.kind == K_BULL) zb -= c.r2;
        out[R_P0Z + 3 * k] = (T)zb;
    }
    // ---- edges (millingcutter.cpp:79-136)
    for (int k = 0; k < 3; ++k) {
        T* e = out + R_E0 + kEdge * k;
        const V3 p1 = (k == 0) ? p[0] : (k == 1) ? p[1] : p[2];
        const V3 p2 = (k == 0) ? p[1] : (k == 1) ? p[2] : p[0];
        const double dx = p2.x - p1.x, dy = p2.y - p1.y, dz = p2.z - p1.z;
        const double L = sqrt(dx * dx + dy * dy);
        const bool degen = near0(t.p[k].x - t.p[(k + 1) % 3].x) && near0(t.p[k].y - t.p[(k + 1) % 3].y);
        if (degen || !(L > 0)) {
            flags |= (FL_EDEG0 << k);
            for (int q = 0; q < kEdge; ++q) e[q] = (T)0;
            e[E_C] = (T)1e30;  // d huge -> never valid
            continue;
        }
        const double ex = dx / L, ey = dy / L;
        const double M = dz / L;
        const double Q = sqrt(1.0 + M * M);
        double kslope = 0.0;  // d(height)/ds of the contact
        e[E_EX] = (T)ex;
        e[E_EY] = (T)ey;
        e[E_C] = (T)(ey * p1.x - ex * p1.y);
        e[E_A] = (T)(ex * p1.x + ey * p1.y);
        e[E_INVL] = (T)(1.0 / L);
        e[E_M] = (T)M;
        double Z1 = p1.z, K1 = 0, K2 = 0, K3 = 0;
        const bool hz = near0(p1.z - p2.z);
        switch (c.kind) {
        case K_CYL:  // cc_u = sign(M) s; z = Z1 - u1 M + |M| s
            K1 = fabs(M);
            K2 = (M > 0) ? 1.0 : -1.0;
            kslope = fabs(M);
            break;
        case K_BALL:  // cc_u = s M/Q; z = Z1 - u1 M + s Q - R
            K1 = Q;
            K2 = M / Q;
            Z1 -= R;
            kslope = Q;
            break;
        case K_BULL:  // horizontal: z = Z1 - h(d); else Newton in tau
            if (hz) flags |= (FL_EHOR0 << k);
            K1 = Q;
            K2 = (M > 0) ? 1.0 : -1.0;
            K3 = hz ? 0.0 : 1.0 / fabs(M);
            kslope = hz ? sqrt(c.r2 / R) : Q * sqrt(R / c.r2);
            break;
        case K_CONE: {  // hyperbola branch linear in d (conecutter.cpp:109-135)
            const double Lc = c.length;
            const double denom = Lc * Lc - R * R * M * M;
            const double KH = (denom > 0) ? fabs(M) * R / sqrt(denom) : 0.0;
            K1 = KH;
            K2 = (M > 0) ? 1.0 : -1.0;
            K3 = sqrt(1.0 + KH * KH);
            kslope = fabs(M);
            break;
        }
        default: break;
        }
        e[E_Z1] = (T)Z1;
        e[E_K1] = (T)K1;
        e[E_K2] = (T)K2;
        e[E_K3] = (T)K3;
        const double ez = 4 * u * (fabs(Z1) + 2.0 * fabs(M) * xyb + Q * R) + Eabs;
        e[E_EZE] = (T)ez;
        e[E_S2E] = (T)er.s2band(kslope, ez);
        e[E_EAE] = (T)er.slack(kslope, ez);
        emax = fmax(emax, ez);
        slk = fmax(slk, fmax(ez, er.slack(kslope, ez)));
    }
    out[R_FLAGS] = (T)flags;
    // rounded up: cull tests compare against this cap
    out[R_ZCAP] = (T)((bz1 + Eabs + slk) * (1.0 + 4.0 * u) + 4.0 * u);
    out[R_EMAX] = (T)(emax * (1.0 + 4.0 * u));
    // Plane cap.  The facet height F0 + GX x + GY y (without the inside
    // test) is the height at which the cutter first touches the triangle's
    // whole plane, and every point of the triangle lies on that plane, so no
    // vertex, edge or facet contact of the triangle sits above it.  The
    // margin adds, like R_ZCAP, the contact error bounds and slack, plus the
    // facet's own rounding once more for the evaluation at the lane.  The
    // BullCutter's exact edge heights come from Brent solves that may sit up
    // to about 1e-9 above the converged contact.  Vertical triangles and
    // cones (tip and rim facets) get no plane cap.
    const bool pcap = !(flags & FL_VERT) && (c.kind == K_CYL || c.kind == K_BALL || c.kind == K_BULL);
    const double pextra = (c.kind == K_BULL) ? 4e-9 : 0.0;
    out[R_PMAR] = pcap ? (T)((2.0 * ezf + Eabs + slk + pextra) * (1.0 + 4.0 * u) + 4.0 * u) : FM<T>::inf();
    for (int q = R_PMAR + 1; q < rec_len<T>(); ++q) out[q] = (T)0;
    if (sizeof(T) == 4)
        for (int q = 0; q < 6; ++q) {
            // The cylinder's contact on a horizontal edge is one end of the
            // chord under the disk, the end towards the edge's first vertex
            // (cylcutter.cpp:62-68 with equal heights), so it depends on the
            // edge's direction; every other contact does not.
            const int j = q - 3;
            const bool directed = j >= 0 && c.kind == K_CYL && t.p[j].z == t.p[(j + 1) % 3].z;
            uint32_t id = fid ? fid[q] : (uint32_t)q;
            if (!directed) id &= ~kEdgeReversed;
            out[R_FID0 + q] = (T)id_bits(id);
        }
}

// Selection constants for one cutter (host side, FP64 -> T).
template <class T>
OCL_HD SelParams<T> sel_params(const CutterDev& cd, const SelErr& er) {
    const CutterP& c = cd.c;
    const double R = c.radius;
    SelParams<T> P{};
    P.kind = c.kind;
    P.R = (T)R;
    P.R2 = (T)(R * R);
    P.r1 = (T)c.r1;
    P.r2 = (T)c.r2;
    P.r22 = (T)(c.r2 * c.r2);
    P.ir2 = (T)(c.r2 > 0 ? 1.0 / c.r2 : 0.0);
    P.cot = (T)(c.kind == K_CONE ? 1.0 / c.tan_a : 0.0);
    P.ch = (T)c.ch;
    P.LEN = (T)c.length;
    const bool sq = (c.kind == K_BALL || c.kind == K_BULL);  // vertex height has sqrt(m)
    P.DM = (T)er.dm();
    P.S2V = (T)(sq ? er.s2band(1.0, er.E()) : er.dm());
    P.EAV = (T)(sq ? er.slack(1.0, er.E()) : 2.0 * er.E());
    P.EPOS = (T)(16.0 * er.u * 4.0 * R + er.E());
    P.SQDM = (T)sqrt(2.0 * er.dm());
    P.E = (T)er.E();
    // BullCutter: the exact path solves the offset ellipse with Brent's
    // method to 1e-10 in diamond angle, so its heights differ from the
    // converged Newton solution by up to ~1e-9; ties closer than that are
    // decided by the exact path.
    if (c.kind == K_BULL) P.E = (T)fmax((double)P.E, 2e-9);
    P.HPAD = (T)(16.0 * er.u * R);
    P.CUT = (T)INFINITY;  // set by the launcher once the records' largest error is known
    P.cull = (c.kind == K_CYL || c.kind == K_BALL || c.kind == K_BULL || (c.kind == K_CONE && c.length >= 0.0)) ? 1 : 0;
    return P;
}

// ------------------------------------------------------------ culling
// Every contact of a triangle whose XY box lies at distance >= d from the
// CL point has height <= maxz - h(d) (h = the cutter profile height, which
// grows with the distance from the axis).  With the record's R_ZCAP (max z
// plus the largest error bound and ambiguity slack of its contacts), a lane
// whose final height is known to reach thr gains nothing from a triangle
// with R_ZCAP - h(d) < thr - CUT: none of its contacts can lift, tie or
// enter the ambiguity window of the winner.

// squared XY distance from (x, y) to the box [x0, x1] x [y0, y1]
template <class T>
OCL_HD T box_d2(T x, T y, T x0, T x1, T y0, T y1) {
    typedef FM<T> F;
    const T dx = F::max_(F::max_(x0 - x, x - x1), (T)0);
    const T dy = F::max_(F::max_(y0 - y, y - y1), (T)0);
    return F::fma_(dx, dx, dy * dy);
}

// profile height at XY distance sqrt(d2), rounded down (0 for the cylinder).
// Cone: every contact at XY distance r from the axis sits r/tan(angle) below
// its contact point (the vertex, edge and tip contacts, conecutter.cpp:
// 109-135 and 70-105, including the hyperbola case), except the rim facet
// contact, which sits the cutter length below a point at distance R: so
// min(r / tan(angle), length) bounds them all.
template <class T, int K>
OCL_HD T prof_lower(T d2, const SelParams<T>& P) {
    typedef FM<T> F;
    T h = (T)0;
    if (K == K_BALL) {
        h = P.R - F::sqrt_(F::max_(P.R2 - d2, (T)0));
    } else if (K == K_BULL) {
        const T w = F::max_(F::sqrt_(d2) - P.r1, (T)0);
        h = P.r2 - F::sqrt_(F::max_(P.r22 - w * w, (T)0));
    } else if (K == K_CONE) {
        h = F::min_(F::sqrt_(d2) * P.cot, P.LEN);
        h = h * ((T)1 - (T)(8.0 * F::eps()));
    }
    return F::max_(h - P.HPAD, (T)0);
}

// Upper bound of every contact's upper value (z + error bound, or z + slack)
// of the triangle of record r for a lane at (x, y) in the record's frame,
// whose XY distance to the triangle's box is sqrt(d2): the lower of the
// profile cap (R_ZCAP - h(d)) and the plane cap (facet height at the lane
// + R_PMAR).
template <class T, int K>
OCL_HD T tri_cap(const T* r, T x, T y, T d2, const SelParams<T>& P) {
    typedef FM<T> F;
    const T zp = r[R_ZCAP] - prof_lower<T, K>(d2, P);
    if (K == K_CONE) return zp;
    return F::min_(zp, F::fma_(r[R_GY], y, F::fma_(r[R_GX], x, r[R_F0])) + r[R_PMAR]);
}

// Contact height of the vertex v = (x, y, z) (cutter base at the vertex, see
// sel_triangle) and m = R^2 - q^2 (>= 0: the vertex is under the cutter).
template <class T, int K>
OCL_HD T vertex_height_at(const T* v, T x, T y, const SelParams<T>& P, T& m) {
    typedef FM<T> F;
    const T dx = v[0] - x, dy = v[1] - y;
    const T q2 = F::fma_(dx, dx, dy * dy);
    m = P.R2 - q2;
    switch (K) {
    case K_CYL: return v[2];
    case K_BALL: return v[2] + F::sqrt_(F::max_(m, (T)0));
    case K_BULL: {
        const T q = F::sqrt_(q2);
        const T w = F::max_(q - P.r1, (T)0);
        // flat part: z = p.z; torus: p.z - r2 + sqrt(r2^2 - w^2)  (base is p.z - r2)
        return v[2] + F::sqrt_(F::max_(P.r22 - w * w, (T)0));
    }
    default: return v[2] - F::sqrt_(q2) * P.cot;  // cone
    }
}

// Vertex j of a selection record.
template <class T, int K>
OCL_HD T vertex_height(const T* rec, int j, T x, T y, const SelParams<T>& P, T& m) {
    static_assert(R_P0Y == R_P0X + 1 && R_P0Z == R_P0X + 2 && R_P1X == R_P0X + 3, "vertex layout");
    return vertex_height_at<T, K>(rec + R_P0X + 3 * j, x, y, P, m);
}

// Lower bound of a lane's final height from one triangle's vertices (v: the
// 9 coordinates p0, p1, p2): a vertex clearly under the cutter is a contact
// the reference always reaches or exceeds (for Cyl/Ball/Bull the facet
// contact it may take instead is the maximum over the triangle).
template <class T, int K>
OCL_HD T vertex_floor_at(const T* v, T x, T y, const SelParams<T>& P, bool on) {
    typedef FM<T> F;
    T zc = -F::inf();
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        T m;
        const T z = vertex_height_at<T, K>(v + 3 * j, x, y, P, m);
        zc = F::max_(zc, (on & (m > P.S2V)) ? z - P.E : -F::inf());
    }
    return zc;
}

// The same from a selection record.
template <class T, int K>
OCL_HD T vertex_floor(const T* rec, T x, T y, const SelParams<T>& P, bool on) {
    return vertex_floor_at<T, K>(rec + R_P0X, x, y, P, on);
}

// --------------------------------------------------------- lane state
// Heights kept per type are upper bounds (z + error bound of the contact).
template <class T>
struct Sel {
    T zb;          // running best (box frame inside a box, wave frame outside)
    T eb;          // error bound of the running best
    uint32_t bt;   // rank index of the best triangle
    int bf;        // feature of the best (FT_*), type derived from it
    bool bamb;     // the best contact's validity test was within its margin
    T zt[4];       // per-type max of z + e: vertex, edge, facet(+tip), facet-rim
    T za;          // max of z + slack over ambiguous contacts
    T zr;          // FP32: max of z + e over valid contacts other than the winner
                   // and other copies of the winner's feature
    uint32_t wid;  // FP32: feature ID of the winner
    bool oamb;     // cone: an order-dependent skip decision was within error
    uint32_t pairs;
};

template <class T>
OCL_HD void sel_init(Sel<T>& S, T z0) {
    S.zb = z0;
    S.eb = (T)0;
    S.bt = 0xffffffffu;
    S.bf = FT_NONE;
    S.bamb = false;
    for (int i = 0; i < 4; ++i) S.zt[i] = -FM<T>::inf();
    S.za = -FM<T>::inf();
    S.zr = -FM<T>::inf();
    S.wid = 0xffffffffu;
    S.oamb = false;
    S.pairs = 0;
}

// true if b holds for any lane of the wavefront (for one lane on the host)
OCL_HD bool sel_any(bool b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __any(b) != 0;
#else
    return b;
#endif
}

// slot of the CC type of a feature: 0 vertex, 1 edge, 2 facet/tip, 3 rim
OCL_HD int feat_slot(int f) {
    const int k = f & 15;
    if (k == FT_FACET) return 2;
    if (k == FT_FACET_RIM) return 3;
    if (k >= FT_EDGE0) return 1;
    return 0;
}

// One edge contact of a record (edge j, cutter kind K) at the lane (x, y):
// height z, validity (under the cutter, CC on the segment) and ambiguity
// (validity within its margins, or an ill-conditioned grazing contact that
// may still lie on the segment), the height error bound eze and the slack
// eae of an ambiguous contact, and the contact's offset ccu along the edge
// from the lane's projection and the lane's signed distance sd from the
// edge's line.
template <class T>
struct EdgeC {
    T z, ccu, sd, eze, eae;
    bool val, amb;
};

template <class T, int K>
OCL_HD EdgeC<T> edge_contact(const T* rec, int j, int flags, T x, T y, const SelParams<T>& P, bool ve) {
    typedef FM<T> F;
    const T* e = rec + R_E0 + kEdge * j;
    const T s2e = e[E_S2E], invl = e[E_INVL];
    const T sd = F::fma_(e[E_EX], y, F::fma_(-e[E_EY], x, e[E_C]));
    const T m = F::fma_(-sd, sd, P.R2);
    const T u1 = F::fma_(-e[E_EY], y, F::fma_(-e[E_EX], x, e[E_A]));
    const T zl = F::fma_(-u1, e[E_M], e[E_Z1]);  // edge height at the CL's projection
    T z, ccu;
    switch (K) {
    case K_CYL:
    case K_BALL: {
        const T s = F::sqrt_(F::max_(m, (T)0));
        z = F::fma_(s, e[E_K1], zl);
        ccu = s * e[E_K2];
        break;
    }
    case K_BULL: {
        const T d = fabs(sd);
        // horizontal edge: cc at the projection, z = Z1 - h(d)
        const T wh = F::max_(d - P.r1, (T)0);
        const T zh = e[E_Z1] - P.r2 + F::sqrt_(F::max_(P.r22 - wh * wh, (T)0));
        // general edge: Newton on d = tau (r2 + r1 Q / sqrt(M^2 + tau^2))
        const T M = e[E_M], Q = e[E_K1], IAM = e[E_K3], AM = fabs(M), M2 = M * M;
        const T RQ = P.r1 * Q;
        T tau = F::max_(F::max_(d * F::rcp_(F::fma_(RQ, IAM, P.r2)), (d - P.r1) * P.ir2), (T)0);
#pragma unroll
        for (int it = 0; it < (sizeof(T) == 4 ? 6 : 12); ++it) {
            const T is = F::rsqrt_(F::fma_(tau, tau, M2));
            const T phi = F::fma_(tau, F::fma_(RQ, is, P.r2), -d);
            const T dphi = F::fma_(RQ * M2, is * is * is, P.r2);
            tau = F::min_(F::fma_(-phi, F::rcp_(dphi), tau), (T)1);
        }
        const T c = F::sqrt_(F::max_(F::fma_(-tau, tau, (T)1), (T)0));
        const T is = F::rsqrt_(F::fma_(tau, tau, M2));
        // With Q^2 = 1 + M^2 the contact's offset along the edge and its
        // height are c |M| (r1 is + r2 / Q) and c (r1 M^2 is + r2 Q):
        // no 1/|M| term, so a nearly horizontal edge (|M| just above the
        // reference's 1e-7 threshold) loses nothing to cancellation.
        const T zg = F::fma_(c, F::fma_(P.r1 * M2, is, P.r2 * Q), zl - P.r2);
        const T cg = e[E_K2] * c * AM * F::fma_(P.r1, is, P.r2 * F::rcp_(Q));
        const bool hz = (flags & (FL_EHOR0 << j)) != 0;
        z = hz ? zh : zg;
        ccu = hz ? (T)0 : cg;
        break;
    }
    default: {  // cone
        const T d = fabs(sd);
        const T s = F::sqrt_(F::max_(m, (T)0));
        const T AM = fabs(e[E_M]);
        const bool hyper = AM * P.R2 <= P.ch * s;
        const T KH = e[E_K1];
        const T zhyp = zl - P.ch + F::fma_(d, F::fma_(-e[E_K3], P.cot, AM * KH), P.R * P.cot);
        const T zrim = F::fma_(AM, s, zl - P.ch);
        z = hyper ? zhyp : zrim;
        ccu = e[E_K2] * (hyper ? KH * d : s);
        break;
    }
    }
    const T t = (ccu - u1) * invl;
    const T mt = F::min_(t, (T)1 - t);
    const T dt = P.EPOS * invl;
    EdgeC<T> r;
    r.z = z;
    r.ccu = ccu;
    r.sd = sd;
    r.eze = e[E_EZE];
    r.eae = e[E_EAE];
    r.val = ve & (m >= (T)0) & (mt >= (T)0);
    r.amb = ve & (m >= -P.DM) & ((fabs(mt) <= dt) | ((m <= s2e) & (mt >= -dt - P.SQDM * invl)));
    return r;
}

// One candidate triangle.  `on` = this lane overlaps the triangle in XY (the
// reference's overlaps() test).  Branch-free: every update is a select, so a
// wavefront runs the triangle as one straight-line block of VALU code.
// x, y: lane position relative to the record's box origin.
// hi (optional): the largest upper value (z + error bound, or z + slack of
// an ambiguous contact) of the triangle's live contacts, -inf if none.
template <class T, int K>
OCL_HD void sel_triangle(const T* rec, uint32_t k, T x, T y, const SelParams<T>& P, Sel<T>& S, bool on,
                         T* hi = nullptr) {
    typedef FM<T> F;
    const T NEG = -F::inf();
    const int flags = (int)rec[R_FLAGS];
    // below(): skip when the running z is clearly above the triangle
    const T bz1 = rec[R_BZ1];
    const bool live = on & (S.zb < bz1);
    S.pairs += live ? 1u : 0u;
    const T ebf = rec[R_EBF], ezf = rec[R_EZF];
    // ---------------- facet(s)
    const T u = F::fma_(rec[R_U2], y, F::fma_(rec[R_U1], x, rec[R_U0]));
    const T v = F::fma_(rec[R_V2], y, F::fma_(rec[R_V1], x, rec[R_V0]));
    const T mf = F::min_(F::min_(u, v), (T)1 - u - v);
    const T zf = F::fma_(rec[R_GY], y, F::fma_(rec[R_GX], x, rec[R_F0]));
    const bool fval = live & (mf > (T)0);
    const bool famb = live & (fabs(mf) <= ebf);
    S.zt[2] = F::max_(S.zt[2], fval ? zf + ezf : NEG);
    S.za = F::max_(S.za, famb ? zf + ezf : NEG);
    T th = (fval | famb) ? zf + ezf : NEG;
    const bool flift = fval & (zf > S.zb);
    bool rlift = false, ramb = false;
    T zr = NEG, urp = NEG;  // rim contact height and upper value (cone)
    if (K == K_CONE) {
        const T ur = F::fma_(rec[R_U2], y, F::fma_(rec[R_U1], x, rec[R_U0B]));
        const T vr = F::fma_(rec[R_V2], y, F::fma_(rec[R_V1], x, rec[R_V0B]));
        const T mr = F::min_(F::min_(ur, vr), (T)1 - ur - vr);
        zr = F::fma_(rec[R_GY], y, F::fma_(rec[R_GX], x, rec[R_F0B]));
        const bool rval = live & (mr > (T)0);
        ramb = live & (fabs(mr) <= ebf);
        urp = rval ? zr + ezf : NEG;
        S.zt[3] = F::max_(S.zt[3], urp);
        S.za = F::max_(S.za, ramb ? zr + ezf : NEG);
        th = F::max_(th, (rval | ramb) ? zr + ezf : NEG);
        rlift = !flift & rval & (zr > S.zb);
        // the cone's vertices/edges may rise above its facet contacts, so
        // whether the facet lifted (and skipped them) must be decided exactly
        const bool close = (fval & fabs(zf - S.zb) <= ezf + S.eb) | (rval & fabs(zr - S.zb) <= ezf + S.eb);
        S.oamb = S.oamb | famb | ramb | close;
    }
    // vertices and edges only count when no facet contact lifted
    const bool ve = live & !flift & !rlift;
    // ---------------- vertices
    constexpr bool kRun = sizeof(T) == 4;  // track the runner-up upper value (FP32)
    T upv[3], upe[3];                      // upper value of each valid contact (runner-up, FP32)
    T zv = NEG;
    int fv = FT_NONE;
    bool vamb = false;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        T m;
        const T z = vertex_height<T, K>(rec, j, x, y, P, m);
        const bool val = ve & (m >= (T)0);
        const bool amb = ve & (m <= P.S2V) & (m >= -P.DM);
        const T up = val ? z + P.E : NEG;
        S.zt[0] = F::max_(S.zt[0], up);
        S.za = F::max_(S.za, amb ? z + P.EAV : NEG);
        th = F::max_(th, F::max_(up, amb ? z + P.EAV : NEG));
        const bool take = val & (z > zv);
        if constexpr (kRun) upv[j] = up;
        zv = take ? z : zv;
        fv = take ? (FT_VTX0 + j) : fv;
        vamb = take ? amb : vamb;
    }
    // ---------------- edges
    T ze = NEG, ee = (T)0;
    int fe = FT_NONE;
    bool eamb = false;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const EdgeC<T> ec = edge_contact<T, K>(rec, j, flags, x, y, P, ve);
        const T z = ec.z, eze = ec.eze, eae = ec.eae;
        const bool val = ec.val, amb = ec.amb;
        const T up = val ? z + eze : NEG;
        S.zt[1] = F::max_(S.zt[1], up);
        S.za = F::max_(S.za, amb ? z + eae : NEG);
        th = F::max_(th, F::max_(up, amb ? z + eae : NEG));
        const bool take = val & (z > ze);
        if constexpr (kRun) upe[j] = up;
        ze = take ? z : ze;
        ee = take ? eze : ee;
        fe = take ? (FT_EDGE0 + j) : fe;
        eamb = take ? amb : eamb;
    }
    // vertices come before edges; strict '>' keeps the first of equal heights
    const bool eb_ = ze > zv;
    const T zw = eb_ ? ze : zv;
    const T ew = eb_ ? ee : P.E;
    const int fw = eb_ ? fe : fv;
    const bool aw = eb_ ? eamb : vamb;
    const bool wlift = (fw != FT_NONE) & (zw > S.zb);
    // runner-up: when the triangle lifts, every other valid contact of it and
    // the previous winner (unless it is a copy of the same feature); when it
    // does not, its valid contacts that are not copies of the winner's feature.
    // Skipped when every lane's contacts lie more than CUT below its running
    // best: they cannot reach the final winner's band (its lower end is at
    // least the running best less the winner's error bound, below CUT), and
    // no lane lifts.
    const bool any = flift | rlift | wlift;
    if constexpr (kRun) {
      if (sel_any(live & (th >= S.zb - P.CUT))) {
        T uvb = NEG, uvo = NEG, uvn = NEG, ueb = NEG, ueo = NEG, uen = NEG;
        uint32_t idv = 0, ide = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const uint32_t iv = id_of((float)rec[R_FID0 + j]), ie = id_of((float)rec[R_FID0 + 3 + j]);
            const bool bv = fv == FT_VTX0 + j, be = fe == FT_EDGE0 + j;
            uvb = bv ? upv[j] : uvb;
            uvo = F::max_(uvo, bv ? NEG : upv[j]);
            uvn = F::max_(uvn, iv != S.wid ? upv[j] : NEG);
            idv = bv ? iv : idv;
            ueb = be ? upe[j] : ueb;
            ueo = F::max_(ueo, be ? NEG : upe[j]);
            uen = F::max_(uen, ie != S.wid ? upe[j] : NEG);
            ide = be ? ie : ide;
        }
        const T uf = fval ? zf + ezf : NEG;
        // a facet or rim lift skips the vertices and edges (their values are
        // -inf then)
        const T uwin_v = F::max_(uvo, F::max_(ueb, ueo));   // a vertex lifts
        const T uwin_e = F::max_(F::max_(uvb, uvo), ueo);   // an edge lifts
        const T uexc = flift ? urp : (rlift ? uf : F::max_(F::max_(uf, urp), eb_ ? uwin_e : uwin_v));
        const T unew = F::max_(F::max_(uf, urp), F::max_(uvn, uen));
        const uint32_t wnew = flift ? facet_id(k, false) : (rlift ? facet_id(k, true) : (eb_ ? ide : idv));
        const T uprev = (S.bt != 0xffffffffu && S.wid != wnew) ? S.zb + S.eb : NEG;
        S.zr = F::max_(S.zr, any ? F::max_(uprev, uexc) : unew);
        S.wid = any ? wnew : S.wid;
      }
    }
    const T zn = flift ? zf : (rlift ? zr : zw);
    const T en = (flift | rlift) ? ezf : ew;
    const int fn = flift ? (int)FT_FACET : (rlift ? (int)FT_FACET_RIM : fw);
    const bool an = flift ? famb : (rlift ? ramb : aw);
    S.zb = any ? zn : S.zb;
    S.eb = any ? en : S.eb;
    S.bt = any ? k : S.bt;
    S.bf = any ? fn : S.bf;
    S.bamb = any ? an : S.bamb;
    if (hi) *hi = th;
}

// Escalation decision at the end of the walk (wave frame).  z0: initial CL
// height, whose "no lift" outcome competes like another contact type.
template <class T>
OCL_HD bool sel_escalate(const Sel<T>& S, T z0, const SelParams<T>& P) {
    if (S.bamb || S.oamb) return true;
    T other = -FM<T>::inf();
    if (S.bt == 0xffffffffu) {
        for (int i = 0; i < 4; ++i) other = FM<T>::max_(other, S.zt[i]);
        return other >= S.zb || S.za >= S.zb;
    }
    const int s = feat_slot(S.bf);
    for (int i = 0; i < 4; ++i)
        if (i != s) other = FM<T>::max_(other, S.zt[i]);
    if (sizeof(T) == 4) other = FM<T>::max_(other, S.zr);  // same-type contacts too
    other = FM<T>::max_(other, z0 + P.E);
    const T lo = S.zb - S.eb;
    return other >= lo || S.za >= lo;
}

}  // namespace ocl_amd
